// 3-D R2C / C2R FFT of the lf2mi library (PlanFFT<Imag> for the R2C case,
// LATfield2_PlanFFT_decl.hpp:731-1056), and the C2C transform of two complex fields
// (:608-719, :1059-1198; passes described at fft_x_c2c_kernel), written for gfx950.
//
// Layout contract (same as the reference k-field): k[((kx+kh) + R((kz+kh) + KG(ky+kh)))C + c]
// with R = N/2+1+2kh, KG = N+2kh.  Rows of the k-field are indexed by two "slots":
// slot1 = lattice dim 1 (kz), slot2 = lattice dim 2 (ky).
//
// Forward (in place in the k-field, no workspace):
//   x pass : two real rows (y, y+1 of plane z) as one complex FFT of length N, split into
//            the two half spectra; row (y, z) lands at slot1 = z, slot2 = y
//   z pass : complex FFT along slot1 (z -> kz), in place
//   y pass : complex FFT along slot2 (y -> ky), in place          => (kx, kz, ky)
// Backward (k-field preserved, plan workspace W with the same row structure):
//   y pass : slot2 ky -> y, k-field -> W;  z pass : slot1 kz -> z, in place on W;
//   x pass : Hermitian rows (real parts of kx = 0, N/2) -> two real rows per FFT.
// Each pass moves every element once through HBM; all butterflies run in LDS
// (Stockham auto-sort, radix 8/4), twiddles come from an FP64-generated table.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>

#include "common.hpp"

namespace lf2mi {

// ---------------------------------------------------------------- small DFTs
template <int S, typename V> __device__ __forceinline__ V mul_si(V a) {  // a * (S i)
    V r;
    if (S < 0) { r.x = a.y; r.y = -a.x; }
    else { r.x = -a.y; r.y = a.x; }
    return r;
}

template <int S, typename V> __device__ __forceinline__ void dft2(V &a0, V &a1) {
    V t = a1;
    a1 = csub(a0, t);
    a0 = cadd(a0, t);
}

template <int S, typename V> __device__ __forceinline__ void dft4(V &a0, V &a1, V &a2, V &a3) {
    V t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <int S, typename V> __device__ __forceinline__ void dft8(V *a) {
    using T = decltype(a[0].x);
    const T r2 = (T)0.70710678118654752440;
    dft4<S>(a[0], a[2], a[4], a[6]);
    dft4<S>(a[1], a[3], a[5], a[7]);
    // odd terms times w8^k, w8 = exp(S 2 pi i / 8)
    V o1, o2, o3;
    o1.x = r2 * (a[3].x - S * a[3].y);
    o1.y = r2 * (a[3].y + S * a[3].x);
    o2 = mul_si<S>(a[5]);
    o3.x = r2 * (-a[7].x - S * a[7].y);
    o3.y = r2 * (-a[7].y + S * a[7].x);
    V e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1];
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1);
    a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2);
    a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3);
    a[7] = csub(e3, o3);
}

// a1' = a0 + w a1 + w^2 a2, a2' = a0 + w^2 a1 + w a2 with w = exp(S 2 pi i / 3)
template <int S, typename V> __device__ __forceinline__ void dft3(V &a0, V &a1, V &a2) {
    using T = decltype(a0.x);
    const T h3 = (T)0.86602540378443864676;  // sqrt(3) / 2
    const V t1 = cadd(a1, a2);
    V t2;
    t2.x = a0.x - (T)0.5 * t1.x;
    t2.y = a0.y - (T)0.5 * t1.y;
    V t3 = mul_si<S>(csub(a1, a2));
    t3.x *= h3;
    t3.y *= h3;
    a0 = cadd(a0, t1);
    a1 = cadd(t2, t3);
    a2 = csub(t2, t3);
}

// y_k = sum_j a_j w^(jk), w = exp(S 2 pi i / 5): the pairs (a1, a4) and (a2, a3) combine
// through cos and sin of 2 pi / 5 and 4 pi / 5
template <int S, typename V> __device__ __forceinline__ void dft5(V *a) {
    using T = decltype(a[0].x);
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    const V t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]), t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
    V b1, b2, e1, e2;
    b1.x = a[0].x + c1 * t1.x + c2 * t2.x;
    b1.y = a[0].y + c1 * t1.y + c2 * t2.y;
    b2.x = a[0].x + c2 * t1.x + c1 * t2.x;
    b2.y = a[0].y + c2 * t1.y + c1 * t2.y;
    e1.x = s1 * t3.x + s2 * t4.x;
    e1.y = s1 * t3.y + s2 * t4.y;
    e2.x = s2 * t3.x - s1 * t4.x;
    e2.y = s2 * t3.y - s1 * t4.y;
    const V d1 = mul_si<S>(e1), d2 = mul_si<S>(e2);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(b1, d1);
    a[4] = csub(b1, d1);
    a[2] = cadd(b2, d2);
    a[3] = csub(b2, d2);
}

template <int R, int S, typename V> __device__ __forceinline__ void dft(V *a) {
    if constexpr (R == 2) dft2<S>(a[0], a[1]);
    else if constexpr (R == 3) dft3<S>(a[0], a[1], a[2]);
    else if constexpr (R == 5) dft5<S>(a);
    else if constexpr (R == 4) dft4<S>(a[0], a[1], a[2], a[3]);
    else dft8<S>(a);
}

// ---------------------------------------------------------------- Stockham passes
// Tile of L lines of length N.  Radix plans (passes in order):
//   powers of two: 8 = [8], 16 = [4,4], 32 = [8,4], 64 = [8,8], 128 = [8,4,4], 256 = [8,8,4],
//                  512 = [8,8,8], 1024 = [8,8,4,4];
//   3 * 2^k:       12 = [4,3], 24 = [8,3], 48 = [4,4,3], 96 = [8,4,3], 192 = [8,8,3],
//                  384 = [8,4,4,3], 768 = [8,8,4,3];
//   5 * 2^k:       10 = [2,5], 20 = [4,5], 40 = [8,5], 80 = [4,4,5], 160 = [8,4,5],
//                  320 = [8,8,5], 640 = [8,4,4,5].
constexpr int fft_np(int N) {
    return N == 8 ? 1
         : (N == 16 || N == 32 || N == 64 || N == 12 || N == 24 || N == 10 || N == 20 || N == 40) ? 2
         : (N == 128 || N == 256 || N == 512 || N == 48 || N == 96 || N == 192 || N == 80 || N == 160 ||
            N == 320) ? 3
         : 4;
}
constexpr int fft_radix(int N, int p) {
    return N == 16 ? 4
         : N == 32 ? (p == 0 ? 8 : 4)
         : N == 128 ? (p == 0 ? 8 : 4)
         : N == 256 ? (p < 2 ? 8 : 4)
         : N == 1024 ? (p < 2 ? 8 : 4)
         : N == 12 ? (p == 0 ? 4 : 3)
         : N == 24 ? (p == 0 ? 8 : 3)
         : N == 48 ? (p < 2 ? 4 : 3)
         : N == 96 ? (p == 0 ? 8 : (p == 1 ? 4 : 3))
         : N == 192 ? (p < 2 ? 8 : 3)
         : N == 384 ? (p == 0 ? 8 : (p < 3 ? 4 : 3))
         : N == 768 ? (p < 2 ? 8 : (p == 2 ? 4 : 3))
         : N == 10 ? (p == 0 ? 2 : 5)
         : N == 20 ? (p == 0 ? 4 : 5)
         : N == 40 ? (p == 0 ? 8 : 5)
         : N == 80 ? (p < 2 ? 4 : 5)
         : N == 160 ? (p == 0 ? 8 : (p == 1 ? 4 : 5))
         : N == 320 ? (p < 2 ? 8 : 5)
         : N == 640 ? (p == 0 ? 8 : (p < 3 ? 4 : 5))
         : 8;
}
// the product of the plan's radices is N
constexpr int fft_prod(int N, int p) { return p == fft_np(N) ? 1 : fft_radix(N, p) * fft_prod(N, p + 1); }
static_assert(fft_prod(8, 0) == 8 && fft_prod(16, 0) == 16 && fft_prod(32, 0) == 32 && fft_prod(64, 0) == 64 &&
                  fft_prod(128, 0) == 128 && fft_prod(256, 0) == 256 && fft_prod(512, 0) == 512 &&
                  fft_prod(1024, 0) == 1024 && fft_prod(12, 0) == 12 && fft_prod(24, 0) == 24 &&
                  fft_prod(48, 0) == 48 && fft_prod(96, 0) == 96 && fft_prod(192, 0) == 192 &&
                  fft_prod(384, 0) == 384 && fft_prod(768, 0) == 768 && fft_prod(10, 0) == 10 &&
                  fft_prod(20, 0) == 20 && fft_prod(40, 0) == 40 && fft_prod(80, 0) == 80 &&
                  fft_prod(160, 0) == 160 && fft_prod(320, 0) == 320 && fft_prod(640, 0) == 640,
              "radix plans");
constexpr int fft_ns(int N, int p) { return p == 0 ? 1 : fft_ns(N, p - 1) * fft_radix(N, p - 1); }

// One radix-R pass after sub-transforms of size NS (Stockham auto-sort): butterfly j of
// line l reads element e = j + k N/R, multiplies by W_N^{(j mod NS) k N/(NS R)}, does a
// radix-R DFT and writes element (j/NS) NS R + (j mod NS) + k NS.  Inputs come from
// load(l, e), outputs go to store(l, e, v): registers <- HBM on the first pass,
// HBM <- registers on the last (both use e = j + k N/R), LDS in between.  BAR puts a
// barrier between the reads and the writes (needed when both touch the same LDS tile).
// MAP 0 maps threads line-fastest (coalesced across lines: the y/z passes); MAP 1 maps
// them j-fastest (coalesced along a line: the x pass).
template <typename T, int N, int L, int WG, int S, int R, int NS, int MAP, bool BAR, class Load, class Store>
__device__ __forceinline__ void pass_g(const cplx<T> *__restrict__ tw, Load &&load, Store &&store) {
    using V = cplx<T>;
    constexpr int NBF = L * (N / R);
    constexpr int PER = (NBF + WG - 1) / WG;
    V a[PER][R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int bf = threadIdx.x + b * WG;
        if (NBF % WG == 0 || bf < NBF) {
            const int l = MAP == 0 ? bf % L : bf / (N / R);
            const int j = MAP == 0 ? bf / L : bf % (N / R);
#pragma unroll
            for (int k = 0; k < R; ++k) a[b][k] = load(l, j + k * (N / R));
        }
    }
    if constexpr (BAR) __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int bf = threadIdx.x + b * WG;
        if (NBF % WG == 0 || bf < NBF) {
            const int l = MAP == 0 ? bf % L : bf / (N / R);
            const int j = MAP == 0 ? bf / L : bf % (N / R);
            const int m = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int k = 1; k < R; ++k) {
                    V w = tw[m * k * (N / (NS * R))];
                    if (S < 0) w.y = -w.y;
                    a[b][k] = cmul(a[b][k], w);
                }
            }
            dft<R, S>(a[b]);
            const int base = (j / NS) * NS * R + m;
#pragma unroll
            for (int k = 0; k < R; ++k) store(l, base + k * NS, a[b][k]);
        }
    }
}

// Whole transform of a tile.  FIRST_LDS / LAST_LDS say whether load0 reads and storeLast
// writes the LDS tile (lds[e*ES + l]); otherwise they are HBM accessors.  MAPF / MAPL map the
// threads of the first / last pass (default MAP): a transposing pass reads line-fastest and
// writes element-fastest, or the other way round, straight from and to HBM.
template <typename T, int N, int L, int ES, int WG, int S, int MAP, bool FIRST_LDS, bool LAST_LDS, int MAPF = MAP,
          int MAPL = MAP, class Load0, class StoreLast>
__device__ __forceinline__ void tile_fft_io(cplx<T> *lds, const cplx<T> *__restrict__ tw, Load0 &&load0,
                                            StoreLast &&storeLast) {
    using V = cplx<T>;
    auto ld = [lds](int l, int e) { return lds[e * ES + l]; };
    auto st = [lds](int l, int e, V v) { lds[e * ES + l] = v; };
    constexpr int NP = fft_np(N);
    if constexpr (NP == 1) {
        pass_g<T, N, L, WG, S, fft_radix(N, 0), 1, MAPF, FIRST_LDS && LAST_LDS>(tw, load0, storeLast);
    } else {
        pass_g<T, N, L, WG, S, fft_radix(N, 0), 1, MAPF, FIRST_LDS>(tw, load0, st);
        __syncthreads();
        if constexpr (NP >= 3) {
            pass_g<T, N, L, WG, S, fft_radix(N, 1), fft_ns(N, 1), MAP, true>(tw, ld, st);
            __syncthreads();
        }
        if constexpr (NP >= 4) {
            pass_g<T, N, L, WG, S, fft_radix(N, 2), fft_ns(N, 2), MAP, true>(tw, ld, st);
            __syncthreads();
        }
        pass_g<T, N, L, WG, S, fft_radix(N, NP - 1), fft_ns(N, NP - 1), MAPL, LAST_LDS>(tw, ld, storeLast);
    }
    if constexpr (LAST_LDS) __syncthreads();
}

// ---------------------------------------------------------------- kernel arguments
template <typename T> struct StridedArgs {
    const cplx<T> *in;
    cplx<T> *out;
    int64_t in_col0, out_col0;  // offset of column 0
    int64_t in_es, out_es;      // stride between elements of a line (transform axis)
    int64_t in_bs, out_bs;      // stride between batches (the other row axis)
    int ncols;                  // valid columns per row
    const cplx<T> *tw;
    // Green fusion (MODE 1): element (col, e, batch) has kx = col / C, ky = e, kz = batch
    const double *green_s;      // s[k] = N^2 (2 - 2 cos(2 pi k / N)) = 4 N^2 sin^2(pi k / N), FP64 table
    const float *green_sf;      // the same table rounded to FP32 (FP32 path)
    double green_scale;
    int C;
};

template <typename T> struct XArgs {
    const void *real_in;  // forward: real field (T); backward: unused
    T *real_out;          // backward: real field
    const cplx<T> *kin;   // backward: row source
    cplx<T> *kout;        // forward: row destination
    int C;
    int rh;               // real ghost width
    int64_t k_col0;       // offset of column (kx = 0, c = 0) of row (slot1 = 0, slot2 = 0)
    int64_t k_s1, k_s2;   // strides of slot1 and slot2 (complex units)
    T shift;              // forward: subtracted from every real value
    T post;               // backward: multiplies every real output
    const cplx<T> *tw;
};

// ---------------------------------------------------------------- strided passes (y, z)
// grid: x = column tiles of L, y = batches.  MODE 0: one transform of sign S, HBM to HBM.
// MODE 1: forward transform, Green multiply, inverse transform (fused Poisson y pass).
// I32: every element offset of a tile, (N - 1) * es + L, fits in 32 bits (one integer
// multiply-add per element instead of a 64-bit one: 4-8 % of the strided passes at 512^3)
template <typename T, int N, int L, int WG, int S, int MODE, bool I32>
__global__ __launch_bounds__(WG) void fft_strided_kernel(StridedArgs<T> A) {
    using V = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    int bx = blockIdx.x, by = blockIdx.y;
    if constexpr (L * (int)sizeof(V) < 128) {
        // a tile row is part of a 128-byte line whose rest belongs to the next tiles: blocks are
        // dispatched round-robin over the 8 XCDs, so give each XCD a contiguous range of tiles
        // and the whole line one L2
        const int n = gridDim.x * gridDim.y, lin = blockIdx.x + gridDim.x * blockIdx.y;
        const int per = n / 8, rem = n % 8, xcd = lin % 8, slot = lin / 8;
        const int t = xcd * per + min(xcd, rem) + slot;
        bx = t % gridDim.x;
        by = t / gridDim.x;
    }
    const int col0 = bx * L;
    const int64_t batch = by;
    const V *in = A.in + A.in_col0 + batch * A.in_bs + col0;
    V *out = A.out + A.out_col0 + batch * A.out_bs + col0;
    const int ncols = A.ncols - col0;
    const int ies = (int)A.in_es, oes = (int)A.out_es;
    auto gload = [in, ncols, ies, &A](int l, int e) {
        if constexpr (I32) return l < ncols ? in[e * ies + l] : mk<T>(0, 0);
        else return l < ncols ? in[(int64_t)e * A.in_es + l] : mk<T>(0, 0);
    };
    auto gstore = [out, ncols, oes, &A](int l, int e, V v) {
        if constexpr (I32) {
            if (l < ncols) out[e * oes + l] = v;
        } else {
            if (l < ncols) out[(int64_t)e * A.out_es + l] = v;
        }
    };
    if constexpr (MODE == 0) {
        tile_fft_io<T, N, L, L, WG, S, 0, false, false>(lds, A.tw, gload, gstore);
    } else {
        // Green's factor -scale / K^2, K^2 = s[kx] + s[ky] + s[kz], zero mode 0, in the
        // transform's precision (FP32 tables for the FP32 path); applied as the forward
        // transform's last pass writes the tile
        const T skz = sizeof(T) == 4 ? (T)A.green_sf[batch] : (T)A.green_s[batch];
        const T gsc = (T)A.green_scale;
        const int C = A.C;
        auto green_store = [lds, skz, gsc, col0, C, &A](int l, int e, V v) {
            // padding lines past the last column (kx > N/2) hold zeros; keep their table
            // index inside the N-entry table
            const int kx = min((col0 + l) / C, N / 2);
            const T sx = sizeof(T) == 4 ? (T)A.green_sf[kx] : (T)A.green_s[kx];
            const T sy = sizeof(T) == 4 ? (T)A.green_sf[e] : (T)A.green_s[e];
            const T k2 = sx + sy + skz;
            // FP32: the hardware reciprocal (within 1 ulp) instead of a correctly rounded
            // division, 7 % of the pass; FP64 divides
            T g = T(0);
            if (k2 > T(0)) {
                if constexpr (sizeof(T) == 4) g = -gsc * __builtin_amdgcn_rcpf(k2);
                else g = -gsc / k2;
            }
            v.x *= g;
            v.y *= g;
            lds[e * L + l] = v;
        };
        auto lload = [lds](int l, int e) { return lds[e * L + l]; };
        // the twiddles in LDS after the tile (the launch adds N values): the two transforms'
        // passes read them there instead of through the vector cache (0.333 -> 0.273 ms at
        // 512^3; for the single transforms of the z and x passes the larger LDS footprint
        // costs more than it saves).  The first pass uses none, and the barrier after it
        // orders these stores before any read.
        V *twl = lds + N * L;
        for (int i = threadIdx.x; i < N; i += WG) twl[i] = A.tw[i];
        tile_fft_io<T, N, L, L, WG, -1, 0, false, true>(lds, twl, gload, green_store);
        tile_fft_io<T, N, L, L, WG, +1, 0, true, false>(lds, twl, lload, gstore);
    }
}

// ---------------------------------------------------------------- x pass, R2C
// Block = P row pairs x C components = P*C complex lines; line (p, c) carries rows
// (2p, 2p+1) of component c as re/im.  Row r = y + N z of the real interior.
// CC: the component count when it is known at compile time (1: the Poisson solve's field),
// else 0 and A.C
// LDS rows of a tile: L lines plus one of padding (bank conflicts), except for the six
// component tile at N >= 1024, where a third workgroup per CU counts for more
template <int N, int L, int CC> constexpr int x_es() { return (N >= 1024 && CC == 6) ? L : L + 1; }

template <typename T, int N, int L, int WG, int CC>
__global__ __launch_bounds__(WG) void fft_x_r2c_kernel(XArgs<T> A) {
    using V = cplx<T>;
    constexpr int ES = x_es<N, L, CC>();
    extern __shared__ unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    const int C = CC ? CC : A.C;
    const int P = L / C;
    const int64_t pair0 = (int64_t)blockIdx.x * P;
    const int64_t npairs = (int64_t)N * N / 2;
    const int rh = A.rh;
    const int64_t G = N + 2 * rh;
    const T *real = (const T *)A.real_in;
    const T shift = A.shift;
    // first pass straight from the two real rows of line (p, c)
    auto rload = [&](int l, int n) {
        const int p = l / C, c = l - p * C;
        const int64_t pr = pair0 + p;
        if (p >= P || pr >= npairs) return mk<T>(0, 0);
        const int64_t row = 2 * pr;
        const int y = (int)(row % N), z = (int)(row / N);
        const int64_t src = ((int64_t)rh + G * ((int64_t)(y + rh) + G * (z + rh))) * C + (int64_t)n * C + c;
        // rows y and y+1 of the same plane (N even) are G*C apart
        return mk<T>(real[src] - shift, real[src + G * C] - shift);
    };
    auto lstore = [lds](int l, int e, V v) { lds[e * ES + l] = v; };
    tile_fft_io<T, N, L, ES, WG, -1, 1, false, true>(lds, A.tw, rload, lstore);
    // split the two half spectra: A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = -i (Z[k] - conj Z[N-k]) / 2
    constexpr int H = N / 2 + 1;
    if constexpr (CC == 1) {
        // one output row at a time (its base address uniform), threads along kx
        for (int rr = 0; rr < 2 * L; ++rr) {
            const int p = rr >> 1, ab = rr & 1;
            const int64_t pr = pair0 + p;
            if (pr >= npairs) break;
            const int64_t row = 2 * pr + ab;
            const int y = (int)(row % N), z = (int)(row / N);
            V *dst = A.kout + A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)y * A.k_s2;
            for (int kx = threadIdx.x; kx < H; kx += WG) {
                const V zk = lds[kx * ES + p];
                const V zm = lds[(kx == 0 ? 0 : N - kx) * ES + p];
                V r;
                if (ab == 0) {
                    r.x = (T)0.5 * (zk.x + zm.x);
                    r.y = (T)0.5 * (zk.y - zm.y);
                } else {
                    r.x = (T)0.5 * (zk.y + zm.y);
                    r.y = (T)0.5 * (zm.x - zk.x);
                }
                dst[kx] = r;
            }
        }
        return;
    }
    const int HC = H * C;
    for (int i = threadIdx.x; i < P * 2 * HC; i += WG) {
        const int q = i % HC;
        const int rest = i / HC;
        const int ab = rest & 1, p = rest >> 1;
        const int64_t pr = pair0 + p;
        if (pr >= npairs) continue;
        const int kx = q / C, c = q % C;
        const int l = p * C + c;
        const V zk = lds[kx * ES + l];
        const V zm = lds[(kx == 0 ? 0 : N - kx) * ES + l];
        V r;
        if (ab == 0) {
            r.x = (T)0.5 * (zk.x + zm.x);
            r.y = (T)0.5 * (zk.y - zm.y);
        } else {
            r.x = (T)0.5 * (zk.y + zm.y);
            r.y = (T)0.5 * (zm.x - zk.x);
        }
        const int64_t row = 2 * pr + ab;
        const int y = (int)(row % N), z = (int)(row / N);
        A.kout[A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)y * A.k_s2 + q] = r;
    }
}

// ---------------------------------------------------------------- x pass, C2R
template <typename T, int N, int L, int WG, int CC>
__global__ __launch_bounds__(WG) void fft_x_c2r_kernel(XArgs<T> A) {
    using V = cplx<T>;
    constexpr int ES = x_es<N, L, CC>();
    extern __shared__ unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    const int C = CC ? CC : A.C;
    const int P = L / C;
    const int64_t pair0 = (int64_t)blockIdx.x * P;
    const int64_t npairs = (int64_t)N * N / 2;
    constexpr int H = N / 2 + 1;
    const int HC = H * C;
    // gather rows a, b into Z: Z[k] = A + iB (k in [0, N/2]), Z[N-k] = conj A + i conj B
    if constexpr (CC == 1) {
        // the (row pair, kx) items spread over the threads, kx fastest, every item's two
        // loads issued before any is used (one memory round trip for the tile instead of one
        // per row pair: 512^3 0.305 -> 0.299 ms, 1024^3 2.35 -> 2.07 ms)
        constexpr int NI = (L * H + WG - 1) / WG;
        V va[NI], vb[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = threadIdx.x + k * WG;
            const int p = i / H, kx = i - p * H;
            const int64_t pr = pair0 + p;
            va[k] = mk<T>(0, 0);
            vb[k] = mk<T>(0, 0);
            if (i < L * H && pr < npairs) {
                const int64_t ra = 2 * pr;
                const int ya = (int)(ra % N), z = (int)(ra / N);
                const V *srca = A.kin + A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)ya * A.k_s2;
                va[k] = srca[kx];
                vb[k] = srca[A.k_s2 + kx];  // row ya + 1 of the same plane (N even)
            }
        }
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = threadIdx.x + k * WG;
            const int p = i / H, kx = i - p * H;
            if (i >= L * H || pair0 + p >= npairs) continue;
            V a = va[k], b = vb[k];
            if (kx == 0 || kx == N / 2) {
                a.y = 0;
                b.y = 0;
            }
            V z1;
            z1.x = a.x - b.y;
            z1.y = a.y + b.x;
            lds[kx * ES + p] = z1;
            if (kx != 0 && kx != N / 2) {
                V z2;
                z2.x = a.x + b.y;
                z2.y = b.x - a.y;
                lds[(N - kx) * ES + p] = z2;
            }
        }
    } else if constexpr (CC > 1) {
        // the same with the component count known: all of the tile's loads in flight at once
        constexpr int PC = L / CC, HCC = H * CC, NI = (PC * HCC + WG - 1) / WG;
        V va[NI], vb[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = threadIdx.x + k * WG;
            const int p = i / HCC, q = i - p * HCC;
            const int64_t pr = pair0 + p;
            va[k] = mk<T>(0, 0);
            vb[k] = mk<T>(0, 0);
            if (i < PC * HCC && pr < npairs) {
                const int64_t ra = 2 * pr;
                const int ya = (int)(ra % N), z = (int)(ra / N);
                const V *srca = A.kin + A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)ya * A.k_s2;
                va[k] = srca[q];
                vb[k] = srca[A.k_s2 + q];
            }
        }
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int i = threadIdx.x + k * WG;
            const int p = i / HCC, q = i - p * HCC;
            if (i >= PC * HCC || pair0 + p >= npairs) continue;
            const int kx = q / CC, c = q - kx * CC;
            const int l = p * CC + c;
            V a = va[k], b = vb[k];
            if (kx == 0 || kx == N / 2) {
                a.y = 0;
                b.y = 0;
            }
            V z1;
            z1.x = a.x - b.y;
            z1.y = a.y + b.x;
            lds[kx * ES + l] = z1;
            if (kx != 0 && kx != N / 2) {
                V z2;
                z2.x = a.x + b.y;
                z2.y = b.x - a.y;
                lds[(N - kx) * ES + l] = z2;
            }
        }
    } else
    for (int i = threadIdx.x; i < P * HC; i += WG) {
        const int q = i % HC;
        const int p = i / HC;
        const int64_t pr = pair0 + p;
        if (pr >= npairs) continue;
        const int kx = q / C, c = q % C;
        const int l = p * C + c;
        const int64_t ra = 2 * pr, rb = ra + 1;
        const int ya = (int)(ra % N), yb = (int)(rb % N), z = (int)(ra / N);
        V a = A.kin[A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)ya * A.k_s2 + q];
        V b = A.kin[A.k_col0 + (int64_t)z * A.k_s1 + (int64_t)yb * A.k_s2 + q];
        if (kx == 0 || kx == N / 2) {
            a.y = 0;
            b.y = 0;
        }
        V z1;
        z1.x = a.x - b.y;
        z1.y = a.y + b.x;
        lds[kx * ES + l] = z1;
        if (kx != 0 && kx != N / 2) {
            V z2;
            z2.x = a.x + b.y;
            z2.y = b.x - a.y;
            lds[(N - kx) * ES + l] = z2;
        }
    }
    __syncthreads();
    const int rh = A.rh;
    const int64_t G = N + 2 * rh;
    T *real = A.real_out;
    const T post = A.post;
    auto lload = [lds](int l, int e) { return lds[e * ES + l]; };
    // last pass straight into the two real rows of line (p, c)
    auto rstore = [&](int l, int n, V v) {
        const int p = l / C, c = l - p * C;
        const int64_t pr = pair0 + p;
        if (p >= P || pr >= npairs) return;
        const int64_t row = 2 * pr;
        const int y = (int)(row % N), z = (int)(row / N);
        const int64_t dst = ((int64_t)rh + G * ((int64_t)(y + rh) + G * (z + rh))) * C + (int64_t)n * C + c;
        real[dst] = v.x * post;
        real[dst + G * C] = v.y * post;
    };
    tile_fft_io<T, N, L, ES, WG, +1, 1, true, false>(lds, A.tw, lload, rstore);
}

// ---------------------------------------------------------------- C2C passes
// PlanFFT<Imag> for two complex fields (LATfield2_PlanFFT_decl.hpp:608-719, :1059-1198): the
// k-field holds mode (kx, ky, kz) at storage (d0, d1, d2) = (kz, kx, ky) (cKSite,
// LATfield2_Site.hpp:219-220).  Three passes, each reading and writing every element once:
//   x : rows along x, complex field <-> W[z][y][kx][c] (dense workspace);
//   y : fft_strided_kernel along y in place on W;
//   z : columns (kx, c) of W at one ky, transformed along z and written with kz fastest
//       into the k-field (forward), or read with kz fastest and written into W (backward).
template <typename T> struct C2CArgs {
    const cplx<T> *in;
    cplx<T> *out;
    int C;
    int rh;                 // ghost width of the complex position-space field
    int64_t k_col0;         // k-field: offset of (kz, kx, ky, c) = 0
    int64_t k_s1, k_s2;     // k-field strides of kx (storage dim 1) and ky (storage dim 2)
    const cplx<T> *tw;
};

// x pass: block = P rows x C components = P*C lines along x; row r = y + N z.  DIR 0
// (forward): field -> W; DIR 1 (backward): W -> field.
template <typename T, int N, int L, int WG, int S, int DIR>
__global__ __launch_bounds__(WG) void fft_x_c2c_kernel(C2CArgs<T> A) {
    using V = cplx<T>;
    constexpr int ES = L + 1;
    extern __shared__ unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    const int C = A.C;
    const int P = L / C;
    const int64_t row0 = (int64_t)blockIdx.x * P;
    const int64_t nrows = (int64_t)N * N;
    const int rh = A.rh;
    const int64_t G = N + 2 * rh;
    // element 0 of line l in the field and in W, or -1 past the last row
    auto field_base = [&](int l) -> int64_t {
        const int p = l / C, c = l - p * C;
        const int64_t r = row0 + p;
        if (p >= P || r >= nrows) return -1;
        const int y = (int)(r % N), z = (int)(r / N);
        return ((int64_t)rh + G * ((int64_t)(y + rh) + G * (z + rh))) * C + c;
    };
    auto work_base = [&](int l) -> int64_t {
        const int p = l / C, c = l - p * C;
        const int64_t r = row0 + p;
        if (p >= P || r >= nrows) return -1;
        return r * N * C + c;
    };
    auto gload = [&](int l, int e) {
        const int64_t b = DIR == 0 ? field_base(l) : work_base(l);
        return b < 0 ? mk<T>(0, 0) : A.in[b + (int64_t)e * C];
    };
    auto gstore = [&](int l, int e, V v) {
        const int64_t b = DIR == 0 ? work_base(l) : field_base(l);
        if (b >= 0) A.out[b + (int64_t)e * C] = v;
    };
    tile_fft_io<T, N, L, ES, WG, S, 1, false, false>(lds, A.tw, gload, gstore);
}

// z pass: tile = L consecutive columns (kx, c) of W at ky = blockIdx.y, N elements along z.
// The W side is read or written line-fastest (coalesced across the columns), the k-field side
// element-fastest (each column's N values along kz are consecutive, stride C); the first and
// last butterfly passes do the HBM accesses with those two thread maps.
template <typename T, int N, int L, int WG, int S, int DIR>
__global__ __launch_bounds__(WG) void fft_z_c2c_kernel(C2CArgs<T> A) {
    using V = cplx<T>;
    constexpr int ES = L + 1;
    extern __shared__ unsigned char smem_raw[];
    V *lds = reinterpret_cast<V *>(smem_raw);
    const int C = A.C;
    const int NC = N * C;
    const int col0 = blockIdx.x * L;
    const int ky = blockIdx.y;
    const int ncols = min(L, NC - col0);
    const int64_t wbase = (int64_t)ky * NC + col0;  // W[z][ky][col]: element z at + z N^2 C
    const int64_t wes = (int64_t)N * NC;
    auto kaddr = [&](int l, int e) {
        const int q = col0 + l, kx = q / C, c = q - kx * C;
        return A.k_col0 + (int64_t)e * C + (int64_t)kx * A.k_s1 + (int64_t)ky * A.k_s2 + c;
    };
    if constexpr (DIR == 0) {
        auto wload = [&](int l, int e) { return l < ncols ? A.in[wbase + (int64_t)e * wes + l] : mk<T>(0, 0); };
        auto kstore = [&](int l, int e, V v) {
            if (l < ncols) A.out[kaddr(l, e)] = v;
        };
        tile_fft_io<T, N, L, ES, WG, S, 1, false, false, 0, 1>(lds, A.tw, wload, kstore);
    } else {
        auto kload = [&](int l, int e) { return l < ncols ? A.in[kaddr(l, e)] : mk<T>(0, 0); };
        auto wstore = [&](int l, int e, V v) {
            if (l < ncols) A.out[wbase + (int64_t)e * wes + l] = v;
        };
        tile_fft_io<T, N, L, ES, WG, S, 1, false, false, 1, 0>(lds, A.tw, kload, wstore);
    }
}

// ---------------------------------------------------------------- launch configuration
template <typename T> struct Tune;
template <> struct Tune<float> {
    static constexpr int WG = 256;
    // 16 columns (128-byte rows); at N = 1024 8 columns (64 KB tiles, two workgroups per CU
    // instead of one; the 64-byte rows are remapped over the XCDs): 1024^3 z / y passes 3.1 ->
    // 2.3 ms for one component, 16.6 -> 14.2 ms for six (4 columns: 2.8 / 17.1 ms)
    template <int N> static constexpr int strided_L() { return N >= 1024 ? 8 : 16; }
    template <int N> static constexpr int green_L() { return N >= 512 ? 8 : 16; }
    // x pass: lines per tile for one component (row pairs), and the widest tile, used when
    // the field has more components than the narrow tile holds
    template <int N> static constexpr int x_L() { return N >= 512 ? 4 : (N >= 128 ? 12 : 24); }
    template <int N> static constexpr int x_L_max() { return N >= 512 ? 6 : (N >= 128 ? 12 : 24); }
    // x pass workgroup: 512 threads at N = 1024 (1024^3 forward / backward x passes 2.87 / 2.63
    // -> 2.41 / 2.35 ms for one component, 21.6 / 23.1 -> 20.6 / 20.9 ms for six)
    template <int N> static constexpr int x_WG() { return N >= 1024 ? 512 : 256; }
    // C2C z pass: 128-byte column reads (padded LDS tile N x (L+1), up to 136 KB at 1024)
    template <int N> static constexpr int c2c_z_L() { return 16; }
    // C2C x pass: lines per tile (8 at N >= 512: 1.65 -> 1.55 ms forward at 512^3)
    template <int N> static constexpr int c2c_x_L() { return N >= 512 ? 8 : (N >= 128 ? 12 : 24); }
};
template <> struct Tune<double> {
    static constexpr int WG = 256;
    template <int N> static constexpr int strided_L() { return 8; }
    template <int N> static constexpr int green_L() { return N >= 1024 ? 4 : 8; }
    template <int N> static constexpr int x_L() { return N >= 512 ? 3 : (N >= 128 ? 6 : 12); }
    template <int N> static constexpr int x_L_max() { return N >= 512 ? 3 : (N >= 128 ? 6 : 12); }
    template <int N> static constexpr int x_WG() { return 256; }
    template <int N> static constexpr int c2c_z_L() { return 8; }
    template <int N> static constexpr int c2c_x_L() { return x_L_max<N>(); }
};

}  // namespace lf2mi

// ---------------------------------------------------------------- plan object
struct lf2mi_fft_plan_s {
    int N, C, rh, kh, dtype;
    int c2c;         // 1: complex position-space field (PlanFFT<Imag> of two Field<Imag>)
    int device;
    void *tw;        // N complex twiddles (dtype)
    double *green_s; // N doubles
    float *green_sf; // N floats
    void *work;      // R2C backward / Poisson workspace: N * N rows of wcols complex; C2C:
                     // W[z][y][kx][c], N^3 C complex
    int64_t wcols;   // padded columns per workspace row
    int64_t bytes;
};

namespace lf2mi {

template <typename T, int N>
static int launch_strided(const StridedArgs<T> &A, int S, int mode, int nbatch, hipStream_t st) {
    constexpr int WG = Tune<T>::WG;
    constexpr int L = Tune<T>::template strided_L<N>();
    constexpr int LG = Tune<T>::template green_L<N>();
    LF2MI_PROF(mode == 1 ? "fft_y_green" : (S < 0 ? "fft_strided_fwd" : "fft_strided_inv"), st);
    const int64_t es = std::max(A.in_es, A.out_es);
    const bool i32 = (int64_t)(N - 1) * es + std::max(L, LG) < ((int64_t)1 << 31);
    if (mode == 1) {
        dim3 grid((A.ncols + LG - 1) / LG, nbatch);
        const size_t shm = sizeof(cplx<T>) * N * (LG + 1);  // the tile and the twiddles
        if (i32) fft_strided_kernel<T, N, LG, WG, -1, 1, true><<<grid, WG, shm, st>>>(A);
        else fft_strided_kernel<T, N, LG, WG, -1, 1, false><<<grid, WG, shm, st>>>(A);
    } else {
        dim3 grid((A.ncols + L - 1) / L, nbatch);
        const size_t shm = sizeof(cplx<T>) * N * L;
        if (S < 0) {
            if (i32) fft_strided_kernel<T, N, L, WG, -1, 0, true><<<grid, WG, shm, st>>>(A);
            else fft_strided_kernel<T, N, L, WG, -1, 0, false><<<grid, WG, shm, st>>>(A);
        } else {
            if (i32) fft_strided_kernel<T, N, L, WG, +1, 0, true><<<grid, WG, shm, st>>>(A);
            else fft_strided_kernel<T, N, L, WG, +1, 0, false><<<grid, WG, shm, st>>>(A);
        }
    }
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

template <typename T, int N, int L>
static int launch_x_tile(const XArgs<T> &A, int S, hipStream_t st) {
    constexpr int WG = Tune<T>::template x_WG<N>();
    const int P = L / A.C;
    const int64_t npairs = (int64_t)N * N / 2;
    const unsigned grid = (unsigned)((npairs + P - 1) / P);
    const bool c6 = A.C == 6 && L >= 6;  // the CC = 6 kernels below
    const size_t shm = sizeof(cplx<T>) * N * (c6 ? x_es<N, L, 6>() : x_es<N, L, 1>());
    LF2MI_PROF(S < 0 ? "fft_x_r2c" : "fft_x_c2r", st);
    // the component counts of the PM step (1) and of vector (3) and symmetric tensor (6) fields
    // are compile-time constants, so that the tile's line -> (pair, component) splits are
    // multiplications
    if (S < 0) {
        if (A.C == 1) fft_x_r2c_kernel<T, N, L, WG, 1><<<grid, WG, shm, st>>>(A);
        else if (A.C == 3 && L >= 3) fft_x_r2c_kernel<T, N, L, WG, (L >= 3 ? 3 : 0)><<<grid, WG, shm, st>>>(A);
        else if (A.C == 6 && L >= 6) fft_x_r2c_kernel<T, N, L, WG, (L >= 6 ? 6 : 0)><<<grid, WG, shm, st>>>(A);
        else fft_x_r2c_kernel<T, N, L, WG, 0><<<grid, WG, shm, st>>>(A);
    } else {
        if (A.C == 1) fft_x_c2r_kernel<T, N, L, WG, 1><<<grid, WG, shm, st>>>(A);
        else if (A.C == 3 && L >= 3) fft_x_c2r_kernel<T, N, L, WG, (L >= 3 ? 3 : 0)><<<grid, WG, shm, st>>>(A);
        else if (A.C == 6 && L >= 6) fft_x_c2r_kernel<T, N, L, WG, (L >= 6 ? 6 : 0)><<<grid, WG, shm, st>>>(A);
        else fft_x_c2r_kernel<T, N, L, WG, 0><<<grid, WG, shm, st>>>(A);
    }
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

template <typename T, int N>
static int launch_x(const XArgs<T> &A, int S, hipStream_t st) {
    constexpr int L = Tune<T>::template x_L<N>();
    constexpr int LMAX = Tune<T>::template x_L_max<N>();
    if (A.C <= L) return launch_x_tile<T, N, L>(A, S, st);
    if constexpr (LMAX > L) {
        if (A.C <= LMAX) return launch_x_tile<T, N, LMAX>(A, S, st);
    }
    return fail(LF2MI_EUNSUPPORTED, "FFT with %d components exceeds the x-pass tile (%d)", A.C, LMAX);
}

template <typename T, int N>
static int launch_c2c(const C2CArgs<T> &A, int pass, int S, hipStream_t st) {
    constexpr int WG = Tune<T>::WG;
    if (pass == 0) {
        constexpr int L = Tune<T>::template c2c_x_L<N>();
        if (A.C > L) return fail(LF2MI_EUNSUPPORTED, "FFT with %d components exceeds the x-pass tile (%d)", A.C, L);
        const int P = L / A.C;
        const unsigned grid = (unsigned)(((int64_t)N * N + P - 1) / P);
        const size_t shm = sizeof(cplx<T>) * N * (L + 1);
        static_assert(sizeof(cplx<T>) * N * (L + 1) <= 163840, "x tile within the 160 KB of LDS of a CU");
        LF2MI_PROF(S < 0 ? "fft_x_c2c_fwd" : "fft_x_c2c_inv", st);
        if (S < 0)
            fft_x_c2c_kernel<T, N, L, WG, -1, 0><<<grid, WG, shm, st>>>(A);
        else
            fft_x_c2c_kernel<T, N, L, WG, +1, 1><<<grid, WG, shm, st>>>(A);
    } else {
        constexpr int L = Tune<T>::template c2c_z_L<N>();
        static_assert(sizeof(cplx<T>) * N * (L + 1) <= 163840, "z tile within the 160 KB of LDS of a CU");
        dim3 grid((N * A.C + L - 1) / L, N);
        const size_t shm = sizeof(cplx<T>) * N * (L + 1);
        LF2MI_PROF(S < 0 ? "fft_z_c2c_fwd" : "fft_z_c2c_inv", st);
        if (S < 0)
            fft_z_c2c_kernel<T, N, L, WG, -1, 0><<<grid, WG, shm, st>>>(A);
        else
            fft_z_c2c_kernel<T, N, L, WG, +1, 1><<<grid, WG, shm, st>>>(A);
    }
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

#define LF2MI_DISPATCH_N(N_, CALL)                              \
    switch (N_) {                                                \
        case 8: { constexpr int NN = 8; return CALL; }           \
        case 16: { constexpr int NN = 16; return CALL; }         \
        case 32: { constexpr int NN = 32; return CALL; }         \
        case 64: { constexpr int NN = 64; return CALL; }         \
        case 128: { constexpr int NN = 128; return CALL; }       \
        case 256: { constexpr int NN = 256; return CALL; }       \
        case 512: { constexpr int NN = 512; return CALL; }       \
        case 1024: { constexpr int NN = 1024; return CALL; }     \
        case 12: { constexpr int NN = 12; return CALL; }         \
        case 24: { constexpr int NN = 24; return CALL; }         \
        case 48: { constexpr int NN = 48; return CALL; }         \
        case 96: { constexpr int NN = 96; return CALL; }         \
        case 192: { constexpr int NN = 192; return CALL; }       \
        case 384: { constexpr int NN = 384; return CALL; }       \
        case 768: { constexpr int NN = 768; return CALL; }       \
        case 10: { constexpr int NN = 10; return CALL; }         \
        case 20: { constexpr int NN = 20; return CALL; }         \
        case 40: { constexpr int NN = 40; return CALL; }         \
        case 80: { constexpr int NN = 80; return CALL; }         \
        case 160: { constexpr int NN = 160; return CALL; }       \
        case 320: { constexpr int NN = 320; return CALL; }       \
        case 640: { constexpr int NN = 640; return CALL; }       \
        default: return fail(LF2MI_ERANGE, "FFT size %d not supported", N_); \
    }

template <typename T> static int strided(int N, const StridedArgs<T> &A, int S, int mode, int nb, hipStream_t st) {
    LF2MI_DISPATCH_N(N, (launch_strided<T, NN>(A, S, mode, nb, st)));
}
template <typename T> static int xpass(int N, const XArgs<T> &A, int S, hipStream_t st) {
    LF2MI_DISPATCH_N(N, (launch_x<T, NN>(A, S, st)));
}
template <typename T> static int c2cpass(int N, const C2CArgs<T> &A, int pass, int S, hipStream_t st) {
    LF2MI_DISPATCH_N(N, (launch_c2c<T, NN>(A, pass, S, st)));
}

// k-field geometry
struct KGeo {
    int64_t col0, s1, s2;  // complex units
};
static KGeo kfield_geo(const lf2mi_fft_plan_s *p) {
    const int64_t R = p->N / 2 + 1 + 2 * p->kh, KG = p->N + 2 * p->kh;
    KGeo g;
    g.s1 = R * p->C;
    g.s2 = R * KG * p->C;
    g.col0 = (int64_t)p->kh * p->C + (int64_t)p->kh * g.s1 + (int64_t)p->kh * g.s2;
    return g;
}
static KGeo work_geo(const lf2mi_fft_plan_s *p) {
    KGeo g;
    g.col0 = 0;
    g.s1 = p->wcols;
    g.s2 = p->wcols * p->N;
    return g;
}

template <typename T>
static int forward_x_only(lf2mi_fft_plan_s *p, const void *rfield, cplx<T> *kout, KGeo kg, T shift, hipStream_t st) {
    XArgs<T> X{};
    X.real_in = rfield;
    X.kout = kout;
    X.C = p->C;
    X.rh = p->rh;
    X.k_col0 = kg.col0;
    X.k_s1 = kg.s1;
    X.k_s2 = kg.s2;
    X.shift = shift;
    X.post = T(1);
    X.tw = (const cplx<T> *)p->tw;
    return xpass<T>(p->N, X, -1, st);
}

// The forward transform goes through the workspace: the x pass reads the real field into W,
// the z pass works in place on W and the y pass writes the k-field.  Every read of the real
// field thus precedes every write of the k-field, so the two may share storage
// (FFT_IN_PLACE, LATfield2_PlanFFT_decl.hpp:804-824); the bytes moved are the same.
template <typename T>
static int forward(lf2mi_fft_plan_s *p, const void *rfield, cplx<T> *kout, KGeo kg, T shift, hipStream_t st) {
    const int N = p->N;
    cplx<T> *W = (cplx<T> *)p->work;
    const KGeo wg = work_geo(p);
    if (int rc = forward_x_only<T>(p, rfield, W, wg, shift, st)) return rc;
    StridedArgs<T> A{};
    A.in = W;
    A.out = W;
    A.in_col0 = A.out_col0 = wg.col0;
    A.ncols = (N / 2 + 1) * p->C;
    A.tw = (const cplx<T> *)p->tw;
    A.C = p->C;
    // z pass in place on W: along slot1, batches over slot2
    A.in_es = A.out_es = wg.s1;
    A.in_bs = A.out_bs = wg.s2;
    if (int rc = strided<T>(N, A, -1, 0, N, st)) return rc;
    // y pass W -> k-field: along slot2, batches over slot1
    A.out = kout;
    A.out_col0 = kg.col0;
    A.in_es = wg.s2;
    A.out_es = kg.s2;
    A.in_bs = wg.s1;
    A.out_bs = kg.s1;
    return strided<T>(N, A, -1, 0, N, st);
}

template <typename T>
static int backward(lf2mi_fft_plan_s *p, const cplx<T> *kin, KGeo kg, void *rfield, T post, hipStream_t st) {
    const int N = p->N;
    cplx<T> *W = (cplx<T> *)p->work;
    const KGeo wg = work_geo(p);
    StridedArgs<T> A{};
    A.ncols = (N / 2 + 1) * p->C;
    A.tw = (const cplx<T> *)p->tw;
    A.C = p->C;
    // y pass: k-field -> W along slot2
    A.in = kin;
    A.out = W;
    A.in_col0 = kg.col0;
    A.out_col0 = wg.col0;
    A.in_es = kg.s2;
    A.out_es = wg.s2;
    A.in_bs = kg.s1;
    A.out_bs = wg.s1;
    if (int rc = strided<T>(N, A, +1, 0, N, st)) return rc;
    // z pass in place on W along slot1
    A.in = W;
    A.in_col0 = wg.col0;
    A.in_es = A.out_es = wg.s1;
    A.in_bs = A.out_bs = wg.s2;
    if (int rc = strided<T>(N, A, +1, 0, N, st)) return rc;
    XArgs<T> X{};
    X.kin = W;
    X.real_out = (T *)rfield;
    X.C = p->C;
    X.rh = p->rh;
    X.k_col0 = wg.col0;
    X.k_s1 = wg.s1;
    X.k_s2 = wg.s2;
    X.post = post;
    X.tw = (const cplx<T> *)p->tw;
    return xpass<T>(N, X, +1, st);
}

// C2C: k-field strides of the cKSite layout (kz fastest, then kx, then ky)
template <typename T>
static C2CArgs<T> c2c_args(const lf2mi_fft_plan_s *p) {
    C2CArgs<T> A{};
    const int64_t KG = p->N + 2 * p->kh;
    A.C = p->C;
    A.rh = p->rh;
    A.k_s1 = KG * p->C;
    A.k_s2 = KG * KG * p->C;
    A.k_col0 = (int64_t)p->kh * (p->C + A.k_s1 + A.k_s2);
    A.tw = (const cplx<T> *)p->tw;
    return A;
}

template <typename T>
static int c2c_forward(lf2mi_fft_plan_s *p, const cplx<T> *rfield, cplx<T> *kfield, hipStream_t st) {
    const int N = p->N;
    cplx<T> *W = (cplx<T> *)p->work;
    C2CArgs<T> A = c2c_args<T>(p);
    A.in = rfield;
    A.out = W;
    if (int rc = c2cpass<T>(N, A, 0, -1, st)) return rc;
    StridedArgs<T> Y{};
    Y.in = W;
    Y.out = W;
    Y.ncols = N * p->C;
    Y.in_es = Y.out_es = (int64_t)N * p->C;
    Y.in_bs = Y.out_bs = (int64_t)N * N * p->C;
    Y.tw = A.tw;
    Y.C = p->C;
    if (int rc = strided<T>(N, Y, -1, 0, N, st)) return rc;
    A.in = W;
    A.out = kfield;
    return c2cpass<T>(N, A, 1, -1, st);
}

template <typename T>
static int c2c_backward(lf2mi_fft_plan_s *p, const cplx<T> *kfield, cplx<T> *rfield, hipStream_t st) {
    const int N = p->N;
    cplx<T> *W = (cplx<T> *)p->work;
    C2CArgs<T> A = c2c_args<T>(p);
    A.in = kfield;
    A.out = W;
    if (int rc = c2cpass<T>(N, A, 1, +1, st)) return rc;
    StridedArgs<T> Y{};
    Y.in = W;
    Y.out = W;
    Y.ncols = N * p->C;
    Y.in_es = Y.out_es = (int64_t)N * p->C;
    Y.in_bs = Y.out_bs = (int64_t)N * N * p->C;
    Y.tw = A.tw;
    Y.C = p->C;
    if (int rc = strided<T>(N, Y, +1, 0, N, st)) return rc;
    A.in = W;
    A.out = rfield;
    return c2cpass<T>(N, A, 0, +1, st);
}

template <typename T>
static int poisson(lf2mi_fft_plan_s *p, const void *rho, void *phi, double shift, double scale, hipStream_t st) {
    const int N = p->N;
    cplx<T> *W = (cplx<T> *)p->work;
    // from N = 1024 the workspace rows are transposed against the k-field's: kz (slot1) at
    // the long stride, ky (slot2) at the short one, so that the fused y + Green + y^-1 pass,
    // whose tiles are the narrow ones (two transforms in LDS), reads rows a few KB apart and
    // the plain z passes take the long stride with their wider tiles (1024^3: 15.4 -> 14.9 ms;
    // at 512^3 the z passes lose more than the y pass gains)
    KGeo wg = work_geo(p);
    if (N >= 1024) std::swap(wg.s1, wg.s2);
    if (int rc = forward_x_only<T>(p, rho, W, wg, (T)shift, st)) return rc;
    StridedArgs<T> A{};
    A.in = W;
    A.out = W;
    A.in_col0 = A.out_col0 = wg.col0;
    A.ncols = (N / 2 + 1) * p->C;
    A.tw = (const cplx<T> *)p->tw;
    A.C = p->C;
    A.green_s = p->green_s;
    A.green_sf = p->green_sf;
    A.green_scale = scale;
    // z forward along slot1
    A.in_es = A.out_es = wg.s1;
    A.in_bs = A.out_bs = wg.s2;
    if (int rc = strided<T>(N, A, -1, 0, N, st)) return rc;
    // y forward + Green + y inverse along slot2 (batch = slot1 = kz)
    A.in_es = A.out_es = wg.s2;
    A.in_bs = A.out_bs = wg.s1;
    if (int rc = strided<T>(N, A, -1, 1, N, st)) return rc;
    // z inverse along slot1
    A.in_es = A.out_es = wg.s1;
    A.in_bs = A.out_bs = wg.s2;
    if (int rc = strided<T>(N, A, +1, 0, N, st)) return rc;
    XArgs<T> X{};
    X.kin = W;
    X.real_out = (T *)phi;
    X.C = p->C;
    X.rh = p->rh;
    X.k_col0 = wg.col0;
    X.k_s1 = wg.s1;
    X.k_s2 = wg.s2;
    X.post = T(1);
    X.tw = (const cplx<T> *)p->tw;
    return xpass<T>(N, X, +1, st);
}

// ---------------------------------------------------------------- Green kernel (drop-in)
// lambda(k) = 2 (sum_i cos(2 pi k_i / N) - 3) / res^2 (examples/poissonSolver.cpp:100-104),
// evaluated as -sum_i 4 sin^2(pi k_i / N) / res^2, the same operator without the cancellation
// of the cosine sum against 3 (at N = 1024 and k = (1,0,0) that cancellation alone costs
// 2e-11 relative in FP64); s4[q] = 4 sin^2(pi q / N)
template <typename T>
__global__ void green_kernel(cplx<T> *__restrict__ k, int N, int C, int kh, double inv_res2, double scale,
                             const double *__restrict__ s4) {
    const int H = N / 2 + 1;
    const int64_t R = H + 2 * kh, KG = N + 2 * kh;
    const int64_t total = (int64_t)H * N * N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int kx = (int)(t % H);
        const int kz = (int)((t / H) % N);
        const int ky = (int)(t / ((int64_t)H * N));
        const int64_t o = ((kx + kh) + R * ((kz + kh) + KG * (ky + kh))) * C;
        double f = 0.0;
        if (kx | ky | kz) {
            const double lam = -(s4[kx] + s4[ky] + s4[kz]) * inv_res2;
            f = scale / lam;
        }
        for (int c = 0; c < C; ++c) {
            cplx<T> v = k[o + c];
            v.x = (T)((double)v.x * f);
            v.y = (T)((double)v.y * f);
            k[o + c] = v;
        }
    }
}

// ---------------------------------------------------------------- Green's multiply, no allocation
// `scratch` holds the eigenvalue table s4[q] = 4 sin^2(pi q / N) of lf2mi_poisson_green_noalloc
// (at most GREEN_SCRATCH_N entries, 8 KiB).  Each GPU the library runs on holds its own copy
// until the process ends; every call on a GPU overwrites it, so calls that overlap on one GPU
// (on different streams) race.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
static constexpr int GREEN_SCRATCH_N = 1024;
static constexpr int GREEN_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[GREEN_SCRATCH_N];
static double *g_scratch_addr[GREEN_MAX_DEVICES];  // device address of `scratch`, per device

// the scratch table's address on the current device, looked up once per device
int green_scratch_init() {
    int dev = 0;
    LF2MI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= GREEN_MAX_DEVICES)
        return fail(LF2MI_ERANGE, "device %d: the library keeps scratch addresses for %d devices", dev, GREEN_MAX_DEVICES);
    if (g_scratch_addr[dev]) return LF2MI_OK;
    void *ptr = nullptr;
    LF2MI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_addr[dev] = static_cast<double *>(ptr);
    return LF2MI_OK;
}

__global__ void green_table_kernel(double *__restrict__ s4, int N) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const double sh = sin(3.141592653589793 * (double)q / (double)N);
    s4[q] = 4.0 * sh * sh;
}

}  // namespace lf2mi

using namespace lf2mi;

extern "C" {

static int plan_create(int N, int C, int rhalo, int khalo, int dtype, int c2c, lf2mi_fft_plan *plan) {
    if (!plan) return fail(LF2MI_EINVAL, "lf2mi_fft_plan_create: null plan pointer");
    *plan = nullptr;
    const bool pow2 = is_pow2(N), three = N % 3 == 0 && is_pow2(N / 3), five = N % 5 == 0 && is_pow2(N / 5);
    if (!(pow2 || three || five) || N < 8 || N > 1024)
        return fail(LF2MI_ERANGE, "FFT size N=%d: 2^k, 3 * 2^k or 5 * 2^k in [8, 1024] is required", N);
    if (C < 1 || rhalo < 0 || khalo < 0) return fail(LF2MI_EINVAL, "bad plan geometry C=%d rhalo=%d khalo=%d", C, rhalo, khalo);
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    // components one x-pass tile holds (Tune<T>::x_L_max, and c2c_x_L for C2C plans)
    const int xl = dtype == LF2MI_F32 ? (N >= 512 ? (c2c ? 8 : 6) : (N >= 128 ? 12 : 24))
                                      : (N >= 512 ? 3 : (N >= 128 ? 6 : 12));
    if (C > xl) return fail(LF2MI_EUNSUPPORTED, "FFT with %d components exceeds the x-pass tile (%d)", C, xl);
    lf2mi_fft_plan_s *p = new (std::nothrow) lf2mi_fft_plan_s();
    if (!p) return fail(LF2MI_ENOMEM, "host allocation failed");
    p->N = N;
    p->C = C;
    p->rh = rhalo;
    p->kh = khalo;
    p->dtype = dtype;
    p->c2c = c2c;
    hipGetDevice(&p->device);
    const size_t esz = dtype == LF2MI_F32 ? sizeof(float2) : sizeof(double2);
    // twiddles and Green eigenvalues in FP64
    double *twd = (double *)malloc(sizeof(double) * 2 * N);
    double *gs = (double *)malloc(sizeof(double) * N);
    float *twf = (float *)malloc(sizeof(float) * 2 * N);
    if (!twd || !gs || !twf) {
        free(twd); free(gs); free(twf); delete p;
        return fail(LF2MI_ENOMEM, "host allocation failed");
    }
    for (int q = 0; q < N; ++q) {
        const double a = 2.0 * M_PI * (double)q / (double)N;
        twd[2 * q] = cos(a);
        twd[2 * q + 1] = sin(a);
        twf[2 * q] = (float)twd[2 * q];
        twf[2 * q + 1] = (float)twd[2 * q + 1];
        // 2 - 2 cos(a) evaluated as 4 sin^2(a / 2): no cancellation for small k
        const double sh = sin(M_PI * (double)q / (double)N);
        gs[q] = (double)N * N * 4.0 * sh * sh;
    }
    // workspace rows padded to a multiple of 16 complex values (128 B for float2)
    p->wcols = (((int64_t)(N / 2 + 1) * C + 15) / 16) * 16;
    const int64_t wbytes = c2c ? (int64_t)esz * N * N * N * C : (int64_t)esz * p->wcols * N * N;
    int rc = LF2MI_OK;
    hipError_t e;
    if ((e = hipMalloc(&p->tw, esz * N)) != hipSuccess) rc = hip_fail(e, "hipMalloc(twiddles)", __FILE__, __LINE__);
    if (!rc && (e = hipMalloc((void **)&p->green_s, sizeof(double) * N)) != hipSuccess) rc = hip_fail(e, "hipMalloc(green)", __FILE__, __LINE__);
    if (!rc && (e = hipMalloc(&p->work, wbytes)) != hipSuccess) rc = hip_fail(e, "hipMalloc(workspace)", __FILE__, __LINE__);
    if (!rc && (e = hipMalloc((void **)&p->green_sf, sizeof(float) * N)) != hipSuccess) rc = hip_fail(e, "hipMalloc(green32)", __FILE__, __LINE__);
    if (!rc) {
        float *gsf = (float *)malloc(sizeof(float) * N);
        if (!gsf) rc = fail(LF2MI_ENOMEM, "host allocation failed");
        else {
            for (int q = 0; q < N; ++q) gsf[q] = (float)gs[q];
            if ((e = hipMemcpy(p->green_sf, gsf, sizeof(float) * N, hipMemcpyHostToDevice)) != hipSuccess)
                rc = hip_fail(e, "hipMemcpy(green32)", __FILE__, __LINE__);
            free(gsf);
        }
    }
    if (!rc && (e = hipMemcpy(p->tw, dtype == LF2MI_F32 ? (void *)twf : (void *)twd, esz * N, hipMemcpyHostToDevice)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy(twiddles)", __FILE__, __LINE__);
    if (!rc && (e = hipMemcpy(p->green_s, gs, sizeof(double) * N, hipMemcpyHostToDevice)) != hipSuccess)
        rc = hip_fail(e, "hipMemcpy(green)", __FILE__, __LINE__);
    free(twd);
    free(gs);
    free(twf);
    if (rc) {
        hipFree(p->tw);
        hipFree(p->green_s);
        hipFree(p->green_sf);
        hipFree(p->work);
        delete p;
        return rc;
    }
    p->bytes = (int64_t)esz * N + (int64_t)(sizeof(double) + sizeof(float)) * N + wbytes;
    *plan = p;
    return LF2MI_OK;
}

int lf2mi_fft_plan_create(int N, int C, int rhalo, int khalo, int dtype, lf2mi_fft_plan *plan) {
    return plan_create(N, C, rhalo, khalo, dtype, 0, plan);
}

int lf2mi_fft_plan_create_c2c(int N, int C, int rhalo, int khalo, int dtype, lf2mi_fft_plan *plan) {
    return plan_create(N, C, rhalo, khalo, dtype, 1, plan);
}

int lf2mi_fft_plan_destroy(lf2mi_fft_plan p) {
    if (!p) return LF2MI_OK;
    hipFree(p->tw);
    hipFree(p->green_s);
    hipFree(p->green_sf);
    hipFree(p->work);
    delete p;
    return LF2MI_OK;
}

int lf2mi_fft_plan_bytes(lf2mi_fft_plan p, int64_t *bytes) {
    if (!p || !bytes) return fail(LF2MI_EINVAL, "lf2mi_fft_plan_bytes: null argument");
    *bytes = p->bytes;
    return LF2MI_OK;
}

int lf2mi_fft_execute(lf2mi_fft_plan p, int direction, void *rfield, void *kfield, void *stream) {
    if (!p) return fail(LF2MI_EINVAL, "lf2mi_fft_execute: null plan");
    if (direction != LF2MI_FFT_FORWARD && direction != LF2MI_FFT_BACKWARD) return LF2MI_OK;
    if (!rfield || !kfield) return fail(LF2MI_EINVAL, "lf2mi_fft_execute: null field");
    hipStream_t st = as_stream(stream);
    if (p->c2c) {
        if (p->dtype == LF2MI_F32) {
            if (direction == LF2MI_FFT_FORWARD) return c2c_forward<float>(p, (const float2 *)rfield, (float2 *)kfield, st);
            return c2c_backward<float>(p, (const float2 *)kfield, (float2 *)rfield, st);
        }
        if (direction == LF2MI_FFT_FORWARD) return c2c_forward<double>(p, (const double2 *)rfield, (double2 *)kfield, st);
        return c2c_backward<double>(p, (const double2 *)kfield, (double2 *)rfield, st);
    }
    const KGeo kg = kfield_geo(p);
    if (p->dtype == LF2MI_F32) {
        if (direction == LF2MI_FFT_FORWARD) return forward<float>(p, rfield, (float2 *)kfield, kg, 0.0f, st);
        return backward<float>(p, (const float2 *)kfield, kg, rfield, 1.0f, st);
    }
    if (direction == LF2MI_FFT_FORWARD) return forward<double>(p, rfield, (double2 *)kfield, kg, 0.0, st);
    return backward<double>(p, (const double2 *)kfield, kg, rfield, 1.0, st);
}

int lf2mi_poisson_green(void *kfield, int N, int C, int khalo, double res, double scale, int dtype, void *stream) {
    if (!kfield || N < 1 || C < 1 || khalo < 0 || !(res > 0)) return fail(LF2MI_EINVAL, "lf2mi_poisson_green: bad argument");
    double *s4v = nullptr;
    double *host = (double *)malloc(sizeof(double) * N);
    if (!host) return fail(LF2MI_ENOMEM, "host allocation failed");
    for (int q = 0; q < N; ++q) {
        const double sh = sin(M_PI * (double)q / (double)N);
        host[q] = 4.0 * sh * sh;
    }
    hipStream_t st = as_stream(stream);
    hipError_t e = hipMallocAsync((void **)&s4v, sizeof(double) * N, st);
    if (e != hipSuccess) { free(host); return hip_fail(e, "hipMallocAsync", __FILE__, __LINE__); }
    e = hipMemcpyAsync(s4v, host, sizeof(double) * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    free(host);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
    const int64_t total = (int64_t)(N / 2 + 1) * N * N;
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    LF2MI_PROF("poisson_green", st);
    if (dtype == LF2MI_F32)
        green_kernel<float><<<g, 256, 0, st>>>((float2 *)kfield, N, C, khalo, 1.0 / (res * res), scale, s4v);
    else if (dtype == LF2MI_F64)
        green_kernel<double><<<g, 256, 0, st>>>((double2 *)kfield, N, C, khalo, 1.0 / (res * res), scale, s4v);
    else {
        hipFreeAsync(s4v, st);
        return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    }
    LF2MI_LAUNCHED();
    LF2MI_HIP(hipFreeAsync(s4v, st));
    return LF2MI_OK;
}

int lf2mi_poisson_green_noalloc(void *kfield, int N, int C, int khalo, double res, double scale, int dtype,
                                void *stream) {
    if (!kfield || N < 1 || C < 1 || khalo < 0 || !(res > 0))
        return fail(LF2MI_EINVAL, "lf2mi_poisson_green_noalloc: bad argument");
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    if (N > GREEN_SCRATCH_N)
        return fail(LF2MI_ERANGE, "lf2mi_poisson_green_noalloc: N = %d exceeds its %d-entry scratch table", N,
                    GREEN_SCRATCH_N);
    if (int rc = green_scratch_init()) return rc;
    int dev = 0;
    LF2MI_HIP(hipGetDevice(&dev));
    double *s4v = g_scratch_addr[dev];
    hipStream_t st = as_stream(stream);
    green_table_kernel<<<(N + 255) / 256, 256, 0, st>>>(s4v, N);
    LF2MI_LAUNCHED();
    const int64_t total = (int64_t)(N / 2 + 1) * N * N;
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    LF2MI_PROF("poisson_green", st);
    if (dtype == LF2MI_F32)
        green_kernel<float><<<g, 256, 0, st>>>((float2 *)kfield, N, C, khalo, 1.0 / (res * res), scale, s4v);
    else
        green_kernel<double><<<g, 256, 0, st>>>((double2 *)kfield, N, C, khalo, 1.0 / (res * res), scale, s4v);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

int lf2mi_poisson_solve(lf2mi_fft_plan p, const void *rho, void *phi, double shift, double scale, void *stream) {
    if (!p || !rho || !phi) return fail(LF2MI_EINVAL, "lf2mi_poisson_solve: null argument");
    if (p->C != 1 || p->c2c) return fail(LF2MI_EINVAL, "lf2mi_poisson_solve needs a one-component R2C plan");
    if (p->dtype == LF2MI_F32) return poisson<float>(p, rho, phi, shift, scale, as_stream(stream));
    return poisson<double>(p, rho, phi, shift, scale, as_stream(stream));
}

}  // extern "C"
