// Particle container, CIC deposits and the fused gather + kick + drift of the lf2mi
// library (Particles<part_simple,...> and particles/projections.hpp on one gfx950 device).
//
// F32 representation: per particle an int32 cell index x + N(y + N z) and an FP32 offset
// in cell units in [0,1) per axis (position = (cell + offset)/N, formed in FP64 when the
// particle moves), FP32 velocity, optional FP32 mass and int64 ID, in SoA arrays.
// Particles are bucketed by bricks of B^3 cells (B = 8 when N allows); a bucket is the
// unit of locality that the reference gets from its per-cell std::forward_list
// (particles/LATfield2_Particles.hpp:97-103,333).  Deposits and gathers work on "items":
// chunks of at most ITEM_P particles of one bucket, each staged through an LDS window
// of the brick's nodes.  After a drift the container is re-bucketed with a counting sort
// (histogram -> scan -> scatter into the second buffer).
//
// F64 representation (the plain path): absolute FP64 positions and velocities in upload
// order; deposits are per-particle FP64 atomics into the field and the gather reads phi
// directly.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"

namespace lf2mi {

static constexpr int ITEM_P = 1024;     // particles per item
static constexpr int DEP_WG = 256;      // deposit / gather workgroup size
static constexpr int MAXB = 8;          // largest brick edge

// ---------------------------------------------------------------- RNG (counter based)
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform double in [0,1) from (seed, index, stream)
__host__ __device__ inline double urand(uint64_t seed, uint64_t idx, uint32_t stream) {
    const uint64_t h = mix64(seed ^ mix64(idx * 16ull + stream));
    return (double)(h >> 11) * 0x1.0p-53;
}
__device__ inline double grand(uint64_t seed, uint64_t idx, uint32_t stream) {
    // Box-Muller on two uniforms of the same counter
    double u1 = urand(seed, idx, stream);
    const double u2 = urand(seed, idx, stream + 8);
    if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// ---------------------------------------------------------------- exact FP64 helpers
// Rounded add / multiply that the compiler may not fuse into an FMA: the drift must
// round exactly like the reference's `pos += dtau * vel` (one product, one sum).
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double sub_rn(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// wave-aggregated atomic increment: lanes with the same key share one atomic; returns
// each lane's slot (base + rank among the lanes with its key)
__device__ __forceinline__ int agg_inc(int *counters, int key, bool active) {
    unsigned long long todo = __ballot(active);
    int slot = 0;
    const int lane = lane_id();
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int lkey = __shfl(key, leader);
        const unsigned long long grp = __ballot(active && key == lkey);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[lkey], __popcll(grp));
        base = __shfl(base, leader);
        if (active && key == lkey) slot = base + __popcll(grp & ((1ull << lane) - 1ull));
        todo &= ~grp;
    }
    return slot;
}

// ---------------------------------------------------------------- geometry helpers
struct Geo {
    int N, B, NB;
    int64_t nbk;
    int lgN, lgB;  // log2 of N and B when they are powers of two, else -1
};
// cell index x + N(y + N z) -> coordinates; shifts when N is a power of two
__device__ __forceinline__ void decode_cell(int cell, const Geo &g, int &x, int &y, int &z) {
    if (g.lgN >= 0) {
        const int m = g.N - 1;
        x = cell & m;
        y = (cell >> g.lgN) & m;
        z = cell >> (2 * g.lgN);
    } else {
        x = cell % g.N;
        y = (cell / g.N) % g.N;
        z = cell / (g.N * g.N);
    }
}
__device__ __forceinline__ int bucket_of(int cell, const Geo &g) {
    int x, y, z;
    decode_cell(cell, g, x, y, z);
    if (g.lgB >= 0) return (x >> g.lgB) + g.NB * ((y >> g.lgB) + g.NB * (z >> g.lgB));
    return (x / g.B) + g.NB * ((y / g.B) + g.NB * (z / g.B));
}
__device__ __forceinline__ int wrapi(int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); }

// ---------------------------------------------------------------- F32 SoA buffer
struct Soa32 {
    int32_t *cell;
    float *off[3];
    float *vel[3];
    float *mass;   // may be null
    int64_t *id;   // may be null
    unsigned long long *pl;  // payload: plw 8-byte words per particle (user structs), may be null
};

// ---------------------------------------------------------------- binning (FP64 -> cell, offset)
// Particles::getPartCoord: (int)floor(pos*N) % N (particles/LATfield2_Particles.hpp:464-467).
// The offset is pos*N - floor(pos*N) rounded to FP32 and kept below 1.
__device__ __forceinline__ bool bin_one(const double p[3], int N, int &cell, float off[3]) {
    int c[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
        const double t = p[d] * (double)N;
        const double fl = floor(t);
        c[d] = (int)fl % N;
        if (c[d] < 0) ok = false;
        float o = (float)(t - fl);
        if (!(o < 1.0f)) o = 0x1.fffffep-1f;
        off[d] = o;
    }
    cell = ok ? c[0] + N * (c[1] + N * c[2]) : -1;
    return ok;
}

// staged FP64 input: AoS pos/vel (host layout), optional mass/id
__global__ void bin_upload_kernel(int64_t np, const double *__restrict__ pos, const double *__restrict__ vel,
                                  const double *__restrict__ mass, const int64_t *__restrict__ id,
                                  const unsigned long long *__restrict__ pl, int plw, Soa32 dst, Geo g,
                                  int *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < np;
    int cell = -1;
    if (in) {
        const double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        float off[3];
        bin_one(p, g.N, cell, off);
        dst.cell[i] = cell;
        for (int d = 0; d < 3; ++d) {
            dst.off[d][i] = off[d];
            dst.vel[d][i] = (float)vel[3 * i + d];
        }
        if (dst.mass) dst.mass[i] = mass ? (float)mass[i] : 1.0f;
        if (dst.id) dst.id[i] = id ? id[i] : i;
        for (int w = 0; w < plw; ++w) dst.pl[i * plw + w] = pl[i * plw + w];
    }
    const bool act = in && cell >= 0;
    agg_inc(counts, act ? bucket_of(cell, g) : 0, act);
}

// ---------------------------------------------------------------- synthetic ICs (F32 and F64)
struct IcArgs {
    int kind;  // 0 lattice, 1 uniform, 2 blobs
    int64_t np;
    int ng;
    double sigma_pos;  // box units
    double sigma_v;
    uint64_t seed;
    int nblobs;
    int64_t nclustered;
    const double *centres;   // 3 * nblobs
    const int64_t *bstart;   // nblobs + 1 prefix of blob occupancies
};

__device__ __forceinline__ void ic_particle(const IcArgs &A, int64_t i, double p[3], double v[3]) {
    if (A.kind == 0) {
        const int64_t ng = A.ng;
        const int64_t ix = i % ng, iy = (i / ng) % ng, iz = i / (ng * ng);
        const double base[3] = {(double)ix / ng, (double)iy / ng, (double)iz / ng};
        for (int d = 0; d < 3; ++d) p[d] = base[d] + A.sigma_pos * grand(A.seed, i, d);
    } else if (A.kind == 1 || i >= A.nclustered) {
        for (int d = 0; d < 3; ++d) p[d] = urand(A.seed, i, 4 + d);
    } else {
        // blob of particle i: largest b with bstart[b] <= i
        int lo = 0, hi = A.nblobs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (A.bstart[mid] <= i) lo = mid;
            else hi = mid;
        }
        for (int d = 0; d < 3; ++d) p[d] = A.centres[3 * lo + d] + A.sigma_pos * grand(A.seed, i, d);
    }
    for (int d = 0; d < 3; ++d) {
        p[d] -= floor(p[d]);  // periodic wrap into [0, 1)
        if (p[d] >= 1.0) p[d] = 0.0;
        v[d] = A.sigma_v * grand(A.seed, i, 12 + d);
    }
}

__global__ void ic_f32_kernel(IcArgs A, Soa32 dst, Geo g, int *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < A.np;
    int cell = -1;
    if (in) {
        double p[3], v[3];
        ic_particle(A, i, p, v);
        float off[3];
        bin_one(p, g.N, cell, off);
        dst.cell[i] = cell;
        for (int d = 0; d < 3; ++d) {
            dst.off[d][i] = off[d];
            dst.vel[d][i] = (float)v[d];
        }
        if (dst.mass) dst.mass[i] = 1.0f;
        if (dst.id) dst.id[i] = i;
    }
    const bool act = in && cell >= 0;
    agg_inc(counts, act ? bucket_of(cell, g) : 0, act);
}

struct Soa64 {
    int32_t *cell;
    double *pos[3];
    double *vel[3];
    double *mass;
    int64_t *id;
    unsigned long long *pl;  // payload, as in Soa32
};

__global__ void ic_f64_kernel(IcArgs A, Soa64 dst, int N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    double p[3], v[3];
    ic_particle(A, i, p, v);
    int c[3];
    for (int d = 0; d < 3; ++d) {
        dst.pos[d][i] = p[d];
        dst.vel[d][i] = v[d];
        c[d] = (int)floor(p[d] * (double)N) % N;
    }
    dst.cell[i] = c[0] + N * (c[1] + N * c[2]);
    if (dst.mass) dst.mass[i] = 1.0;
    if (dst.id) dst.id[i] = i;
}

__global__ void upload_f64_kernel(int64_t np, const double *__restrict__ pos, const double *__restrict__ vel,
                                  const double *__restrict__ mass, const int64_t *__restrict__ id, Soa64 dst, int N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int c[3];
    for (int d = 0; d < 3; ++d) {
        const double p = pos[3 * i + d];
        dst.pos[d][i] = p;
        dst.vel[d][i] = vel[3 * i + d];
        c[d] = (int)floor(p * (double)N) % N;
    }
    dst.cell[i] = c[0] + N * (c[1] + N * c[2]);
    if (dst.mass) dst.mass[i] = mass ? mass[i] : 1.0;
    if (dst.id) dst.id[i] = id ? id[i] : i;
}

// ---------------------------------------------------------------- scans
// exclusive scans of count[] (-> start[]) and of ceil(count / ITEM_P) (-> istart[]),
// n <= 4096 * 1024, each with the total at index n.
static constexpr int SCAN_T = 256, SCAN_PER = 4, SCAN_CHUNK = SCAN_T * SCAN_PER;

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total) {
    // Hillis-Steele over SCAN_T threads
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_T; o <<= 1) {
        const int add = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[SCAN_T - 1];
    const int incl = sh[t];
    __syncthreads();
    return incl - v;
}

// exclusive block scan with wave shuffles: two barriers (blockDim.x = SCAN_T = 256)
__device__ __forceinline__ int block_excl_scan_fast(int v, int *wsum, int &total) {
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < SCAN_T / 64; ++w) {
        const int t = wsum[w];
        if (w < wave) before += t;
        total += t;
    }
    __syncthreads();
    return before + x - v;
}

__global__ void scan_local_kernel(const int *__restrict__ count, int64_t n, int *__restrict__ start,
                                  int *__restrict__ istart, int *__restrict__ bsum) {
    __shared__ int sh[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
    int c[SCAN_PER], it[SCAN_PER];
    int sc = 0, si = 0;
    for (int k = 0; k < SCAN_PER; ++k) {
        const int64_t i = base + k;
        c[k] = i < n ? count[i] : 0;
        it[k] = (c[k] + ITEM_P - 1) / ITEM_P;
        sc += c[k];
        si += it[k];
    }
    int tc, ti;
    int oc = block_excl_scan(sc, sh, tc);
    int oi = block_excl_scan(si, sh, ti);
    for (int k = 0; k < SCAN_PER; ++k) {
        const int64_t i = base + k;
        if (i < n) {
            start[i] = oc;
            istart[i] = oi;
        }
        oc += c[k];
        oi += it[k];
    }
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = tc;
        bsum[2 * blockIdx.x + 1] = ti;
    }
}

__global__ void scan_sums_kernel(int *__restrict__ bsum, int nblocks, int *__restrict__ start,
                                 int *__restrict__ istart, int64_t n) {
    // one block of 1024 threads, up to 4 block sums per thread
    __shared__ int sc[1024], si[1024];
    const int t = threadIdx.x;
    int vc[4], vi[4], tc = 0, ti = 0;
    for (int k = 0; k < 4; ++k) {
        const int b = t * 4 + k;
        vc[k] = b < nblocks ? bsum[2 * b] : 0;
        vi[k] = b < nblocks ? bsum[2 * b + 1] : 0;
        tc += vc[k];
        ti += vi[k];
    }
    sc[t] = tc;
    si[t] = ti;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int ac = t >= o ? sc[t - o] : 0;
        const int ai = t >= o ? si[t - o] : 0;
        __syncthreads();
        sc[t] += ac;
        si[t] += ai;
        __syncthreads();
    }
    int oc = sc[t] - tc, oi = si[t] - ti;
    for (int k = 0; k < 4; ++k) {
        const int b = t * 4 + k;
        if (b < nblocks) {
            bsum[2 * b] = oc;
            bsum[2 * b + 1] = oi;
        }
        oc += vc[k];
        oi += vi[k];
    }
    if (t == 1023) {
        start[n] = sc[1023];
        istart[n] = si[1023];
    }
}

__global__ void scan_add_kernel(int *__restrict__ start, int *__restrict__ istart, const int *__restrict__ bsum,
                                int64_t n) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
    const int ac = bsum[2 * blockIdx.x], ai = bsum[2 * blockIdx.x + 1];
    for (int k = threadIdx.x; k < SCAN_CHUNK; k += SCAN_T) {
        const int64_t i = base + k;
        if (i < n) {
            start[i] += ac;
            istart[i] += ai;
        }
    }
}

// ---------------------------------------------------------------- scatter (counting sort)
// Counting-sort scatter, SCAT_PER particles per thread (1024 per block): every load is in
// flight before the slots are claimed.  Slots: per wave, lanes with equal new buckets form
// groups (leader loop) whose leader adds the group size to the bucket's entry in an LDS hash
// table (one CAS to insert the key); after a barrier one global atomic per distinct bucket of
// the block reserves the range, and after a second barrier every particle stores at
// start[b] + base + rank.  A re-bucketing block sees a few buckets (the source is bucket
// ordered); a block of an unordered source (initial sort) at most 1024, for which the
// 2048-slot table has room.
static constexpr int SCAT_PER = 4;
static constexpr int SCAT_HT = 2048;

__global__ __launch_bounds__(256) void scatter_kernel(int64_t np, Soa32 src, Soa32 dst, Geo g,
                                                      const int *__restrict__ start, int *__restrict__ cursor,
                                                      int plw) {
    __shared__ int hkey[SCAT_HT];
    __shared__ int hval[SCAT_HT];  // group counts, then the bucket's reserved base
    for (int i = threadIdx.x; i < SCAT_HT; i += 256) {
        hkey[i] = -1;
        hval[i] = 0;
    }
    const int64_t i0 = (int64_t)blockIdx.x * (256 * SCAT_PER) + threadIdx.x;
    int cell[SCAT_PER];
    float o[SCAT_PER][3], v[SCAT_PER][3], m[SCAT_PER];
    int64_t id[SCAT_PER];
#pragma unroll
    for (int k = 0; k < SCAT_PER; ++k) {
        const int64_t i = i0 + k * 256;
        cell[k] = i < np ? src.cell[i] : -1;
        m[k] = 0.f;
        id[k] = 0;
        if (cell[k] >= 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                o[k][d] = src.off[d][i];
                v[k][d] = src.vel[d][i];
            }
            if (src.mass) m[k] = src.mass[i];
            if (src.id) id[k] = src.id[i];
        }
    }
    __syncthreads();
    const int lane = lane_id();
    int bk[SCAT_PER], ent[SCAT_PER], rank[SCAT_PER];
#pragma unroll
    for (int k = 0; k < SCAT_PER; ++k) {
        const bool act = cell[k] >= 0;
        bk[k] = act ? bucket_of(cell[k], g) : 0;
        ent[k] = 0;
        rank[k] = 0;
        unsigned long long todo = __ballot(act);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int lkey = __shfl(bk[k], leader);
            const unsigned long long grp = __ballot(act && bk[k] == lkey);
            int e = 0, r0 = 0;
            if (lane == leader) {
                e = (int)(((unsigned)lkey * 2654435761u) >> 21) & (SCAT_HT - 1);
                for (;;) {
                    const int old = atomicCAS(&hkey[e], -1, lkey);
                    if (old == -1 || old == lkey) break;
                    e = (e + 1) & (SCAT_HT - 1);
                }
                r0 = atomicAdd(&hval[e], __popcll(grp));
            }
            e = __shfl(e, leader);
            r0 = __shfl(r0, leader);
            if (act && bk[k] == lkey) {
                ent[k] = e;
                rank[k] = r0 + __popcll(grp & ((1ull << lane) - 1ull));
            }
            todo &= ~grp;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SCAT_HT; i += 256) {
        const int key = hkey[i];
        if (key >= 0) hval[i] = start[key] + atomicAdd(&cursor[key], hval[i]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAT_PER; ++k) {
        if (cell[k] < 0) continue;
        const int64_t j = (int64_t)hval[ent[k]] + rank[k];
        dst.cell[j] = cell[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            dst.off[d][j] = o[k][d];
            dst.vel[d][j] = v[k][d];
        }
        if (src.mass) dst.mass[j] = m[k];
        if (src.id) dst.id[j] = id[k];
        const int64_t i = i0 + k * 256;
        for (int w = 0; w < plw; ++w) dst.pl[j * plw + w] = src.pl[i * plw + w];
    }
}

// ---------------------------------------------------------------- items
// item -> bucket table, filled after every scan (a bucket of c particles owns
// ceil(c / ITEM_P) consecutive items)
__global__ void item_bucket_kernel(const int *__restrict__ istart, int64_t nbk, int *__restrict__ item_bucket) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbk) return;
    for (int k = istart[b]; k < istart[b + 1]; ++k) item_bucket[k] = (int)b;
}

struct ItemRange {
    int b;        // bucket
    int p0, p1;   // particle range
    int ox, oy, oz;  // brick origin (cells)
};
__device__ __forceinline__ bool item_range(int item, const int *__restrict__ start, const int *__restrict__ istart,
                                           const int *__restrict__ item_bucket, const Geo &g, ItemRange &r) {
    if (item >= istart[g.nbk]) return false;
    const int b = item_bucket[item];
    const int k = item - istart[b];
    r.b = b;
    r.p0 = start[b] + k * ITEM_P;
    r.p1 = min(r.p0 + ITEM_P, start[b + 1]);
    r.ox = (b % g.NB) * g.B;
    r.oy = ((b / g.NB) % g.NB) * g.B;
    r.oz = (int)(b / ((int64_t)g.NB * g.NB)) * g.B;
    return true;
}

// ---------------------------------------------------------------- deposit (F32)
// The reference accumulates, per cell, the contributions of the cell's particles to the
// nodes around it (localCube[8] / vi[36] / tii[24] + tij[54], particles/projections.hpp:
// 140-179, :1034-1122, :1421-1521) and then adds them to the field.  Here one workgroup
// takes an item (a chunk of at most ITEM_P particles of one brick) and
//   A. counting-sorts the chunk by cell in LDS (one int LDS atomic per particle);
//   B. each lane takes a run of DEP_PER consecutive sorted particles and sums their node
//      terms per cell in registers; a run's inner cells go to the LDS node window with LDS
//      atomics, its last cell is combined across the wave with a segmented DPP sum (equal
//      cells are adjacent lanes after the sort) and added once per run of lanes, so a
//      crowded cell costs the same as a sparse one.  The window holds 64-bit fixed point
//      (integer LDS atomics: the FP32 LDS atomic add is many times slower on gfx950): per
//      item and component a power-of-two scale puts the bound sum_p |q_p| (from the item's
//      maxima of |q| and |v|) just below 2^51, so the window sums are exact and the only
//      rounding past the FP32 run sums is the quantum 2^-51 of that bound;
//   C. the node window (all components, component-fastest) is added into the field with
//      one global atomic per non-zero value (periodic wrap or the unfolded ghost layer).
// Node offsets per axis: {0, 1} on the vertex grid, {-1, 0, +1} on the link grid (nodes at
// n + 1/2) and {0} on a nearest-grid-point axis, so 8 (scalar, T_ii), 12 (vector component),
// 18 (T_ij), 4 (CIC-NGP vector component) or 2 (CIC-NGP T_ij) terms per particle and
// component; the window spans nodes [LO, B] per axis, LO = -1 when a component of the kind
// sits on links.
template <int KIND> struct KindC;
// CW: components the node window holds at a time (VecVec runs two passes of three, which
// keeps its LDS at the vector kind's)
template <> struct KindC<0> { static constexpr int C = 1, CW = 1, LO = 0; };
template <> struct KindC<1> { static constexpr int C = 3, CW = 3, LO = -1; };
template <> struct KindC<2> { static constexpr int C = 6, CW = 3, LO = -1; };
// the CIC-NGP hybrids (vertex and nearest-grid-point axes only)
template <> struct KindC<3> { static constexpr int C = 3, CW = 3, LO = 0; };
template <> struct KindC<4> { static constexpr int C = 6, CW = 3, LO = 0; };

// link-grid weights of one axis as a 3-vector over node offsets {-1, 0, +1}
// (projections.hpp:1058-1071: the branch at the link midpoint)
__device__ __forceinline__ void link3(float u, float &wm, float &w0, float &wp) {
    const float t = u - 0.5f;
    if (t < 0.0f) {
        const float a = -t;
        wm = a;
        w0 = 1.0f - a;
        wp = 0.0f;
    } else {
        wm = 0.0f;
        w0 = 1.0f - t;
        wp = t;
    }
}

// component c of kind KIND: stagger mask (bit d: link grid along axis d), nearest-grid-point
// mask (bit d: the single node of the particle's cell along axis d, weight 1) and velocity
// factors; an axis in neither mask takes the vertex weights
template <int KIND, int COMP> struct Comp;
template <> struct Comp<0, 0> { static constexpr int MASK = 0, NGP = 0, VA = -1, VB = -1; };
template <> struct Comp<1, 0> { static constexpr int MASK = 1, NGP = 0, VA = 0, VB = -1; };
template <> struct Comp<1, 1> { static constexpr int MASK = 2, NGP = 0, VA = 1, VB = -1; };
template <> struct Comp<1, 2> { static constexpr int MASK = 4, NGP = 0, VA = 2, VB = -1; };
template <> struct Comp<2, 0> { static constexpr int MASK = 0, NGP = 0, VA = 0, VB = 0; };  // (0,0)
template <> struct Comp<2, 1> { static constexpr int MASK = 3, NGP = 0, VA = 0, VB = 1; };  // (0,1)
template <> struct Comp<2, 2> { static constexpr int MASK = 5, NGP = 0, VA = 0, VB = 2; };  // (0,2)
template <> struct Comp<2, 3> { static constexpr int MASK = 0, NGP = 0, VA = 1, VB = 1; };  // (1,1)
template <> struct Comp<2, 4> { static constexpr int MASK = 6, NGP = 0, VA = 1, VB = 2; };  // (1,2)
template <> struct Comp<2, 5> { static constexpr int MASK = 0, NGP = 0, VA = 2, VB = 2; };  // (2,2)
// vectorProjectionCICNGP_project (projections.hpp:422-549): component i is NGP along i
template <> struct Comp<3, 0> { static constexpr int MASK = 0, NGP = 1, VA = 0, VB = -1; };
template <> struct Comp<3, 1> { static constexpr int MASK = 0, NGP = 2, VA = 1, VB = -1; };
template <> struct Comp<3, 2> { static constexpr int MASK = 0, NGP = 4, VA = 2, VB = -1; };
// symtensorProjectionCICNGP_project (projections.hpp:709-847): T_ii on the vertices, T_ij
// NGP along i and j and CIC along the third axis
template <> struct Comp<4, 0> { static constexpr int MASK = 0, NGP = 0, VA = 0, VB = 0; };  // (0,0)
template <> struct Comp<4, 1> { static constexpr int MASK = 0, NGP = 3, VA = 0, VB = 1; };  // (0,1)
template <> struct Comp<4, 2> { static constexpr int MASK = 0, NGP = 5, VA = 0, VB = 2; };  // (0,2)
template <> struct Comp<4, 3> { static constexpr int MASK = 0, NGP = 0, VA = 1, VB = 1; };  // (1,1)
template <> struct Comp<4, 4> { static constexpr int MASK = 0, NGP = 6, VA = 1, VB = 2; };  // (1,2)
template <> struct Comp<4, 5> { static constexpr int MASK = 0, NGP = 0, VA = 2, VB = 2; };  // (2,2)

// nodes per axis d of a component: 3 on links, 1 for NGP, 2 on vertices
template <int MASK, int NGP, int D>
__host__ __device__ constexpr int axis_nodes() {
    return (MASK >> D) & 1 ? 3 : ((NGP >> D) & 1 ? 1 : 2);
}

struct DepArgs {
    Soa32 s;
    const int *start;
    const int *istart;
    const int *item_bucket;
    Geo g;
    float mass_nv;           // global mass * N^3 (mass / res^3)
    float *field;            // native accumulation target
    unsigned long long *fx;  // fixed-point target (KIND 0 only)
    double fxs[6];           // fixed-point scale per component (scalar: 2^32 / mass_nv, so the
                             // fixed point counts dimensionless weights)
    int h;
    int tgt;
    const int *dlist;        // deposit_f32_kernel: a list of items (with its count at dcount)
    const int *dcount;       // instead of one item per workgroup
};

template <typename ACC> struct AccOps;
template <> struct AccOps<float> {
    static __device__ __forceinline__ float term(float q, float w, double) { return q * w; }
    static __device__ __forceinline__ void flush(const DepArgs &A, int64_t idx, float v) {
        atomicAdd(&A.field[idx], v);
    }
};
template <> struct AccOps<unsigned long long> {
    // each particle's term is quantised on its own, so the integer sums do not depend on order
    static __device__ __forceinline__ unsigned long long term(float q, float w, double fxscale) {
        return (unsigned long long)llrint((double)(q * w) * fxscale);
    }
    static __device__ __forceinline__ void flush(const DepArgs &A, int64_t idx, unsigned long long v) {
        atomicAdd(&A.fx[idx], v);
    }
};

// ---- segmented sum over the lanes of a wave (DPP)
// update_dpp with bound_ctrl off: lanes whose source is outside the row, or whose row the
// row mask leaves out, get `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i32(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_val(float v) {
    return __int_as_float(dpp_i32<CTRL, ROWMASK>(0, __float_as_int(v)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned long long dpp_val(unsigned long long v) {
    const unsigned lo = (unsigned)dpp_i32<CTRL, ROWMASK>(0, (int)(unsigned)v);
    const unsigned hi = (unsigned)dpp_i32<CTRL, ROWMASK>(0, (int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
template <int CTRL, int ROWMASK, typename T, int NS>
__device__ __forceinline__ void seg_step(int key, T (&v)[NS]) {
    const bool same = dpp_i32<CTRL, ROWMASK>(-2, key) == key;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const T s = dpp_val<CTRL, ROWMASK>(v[k]);
        v[k] += same ? s : T(0);
    }
}
// Inclusive segmented scan over the 64 lanes (all active): lanes with equal keys must be
// adjacent (keys >= 0, the sentinel -2 never matches); afterwards the last lane of each run
// of equal keys holds the run's sum.  Within rows of 16: shifts by 1, 2, 4, 8; then lane 15
// of rows 0 and 2 into rows 1 and 3, and lane 31 into rows 2 and 3.
template <typename T, int NS>
__device__ __forceinline__ void seg_scan(int key, T (&v)[NS]) {
    seg_step<0x111, 0xf>(key, v);  // row_shr:1
    seg_step<0x112, 0xf>(key, v);  // row_shr:2
    seg_step<0x114, 0xf>(key, v);  // row_shr:4
    seg_step<0x118, 0xf>(key, v);  // row_shr:8
    seg_step<0x142, 0xa>(key, v);  // row_bcast:15
    seg_step<0x143, 0xc>(key, v);  // row_bcast:31
}

static constexpr int DEP_PER = ITEM_P / DEP_WG;  // sorted particles per lane run

// window value of a run sum: FP32 sums in the item's fixed point (nearest), fixed-point terms
// as they are.  |x scale| < 2^51, so adding 1.5 2^52 leaves the rounded integer in the low
// mantissa bits of the sum (two's complement once the constant's bits are taken off).
__device__ __forceinline__ unsigned long long to_window(float x, double scale) {
    const double t = fma((double)x, scale, 0x1.8p52);
    return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000ll);
}
__device__ __forceinline__ unsigned long long to_window(unsigned long long x, double) { return x; }

// add a cell's NS node terms into the LDS node window of one component (edge NWE,
// origin LO): node offsets per axis start at -1 on link axes, 0 on the others
template <int MASK, int NGP, int LO, typename ACC, int NS>
__device__ __forceinline__ void window_add(unsigned long long *nd, int cell, int B, int NWE, const ACC (&acc)[NS],
                                           double scale) {
    constexpr int SX = axis_nodes<MASK, NGP, 0>(), SY = axis_nodes<MASK, NGP, 1>(), SZ = axis_nodes<MASK, NGP, 2>();
    const int lx = cell % B, ly = (cell / B) % B, lz = cell / (B * B);
    const int base = (lx + ((MASK & 1) ? -1 : 0) - LO) + NWE * ((ly + ((MASK & 2) ? -1 : 0) - LO) +
                                                              NWE * (lz + ((MASK & 4) ? -1 : 0) - LO));
#pragma unroll
    for (int kz = 0; kz < SZ; ++kz)
#pragma unroll
        for (int ky = 0; ky < SY; ++ky)
#pragma unroll
            for (int kx = 0; kx < SX; ++kx)
                atomicAdd(&nd[base + kx + NWE * (ky + NWE * kz)], to_window(acc[kx + SX * (ky + SY * kz)], scale));
}

// phase B for one component: the lane's run of sorted particles (cells key[k], -1 past the
// item's end), its last cell carried across the wave
template <int KIND, int COMP, typename ACC, int BC, int PER = DEP_PER>
__device__ __forceinline__ void deposit_component(const DepArgs &A, int B, int NWE, const int (&key)[PER],
                                                  const float4 (&pd)[PER], const float4 (&pv)[PER],
                                                  int ckey, bool any_shared, bool carry_end, unsigned long long *nodes,
                                                  double scale) {
    constexpr int MASK = Comp<KIND, COMP>::MASK, NGP = Comp<KIND, COMP>::NGP;
    constexpr int VA = Comp<KIND, COMP>::VA, VB = Comp<KIND, COMP>::VB;
    constexpr int SX = axis_nodes<MASK, NGP, 0>(), SY = axis_nodes<MASK, NGP, 1>(), SZ = axis_nodes<MASK, NGP, 2>();
    constexpr int NS = SX * SY * SZ;
    constexpr int LO = KindC<KIND>::LO;
    constexpr int NWMAX = MAXB + 1 - LO;
    unsigned long long *nd = nodes + (COMP % KindC<KIND>::CW) * NWMAX * NWMAX * NWMAX;
    ACC acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = ACC(0);
    int cur = -1;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int c = key[p];
        if (c < 0) continue;  // past the end: every later particle of the run is too
        if (c != cur) {
            if (cur >= 0) {
                window_add<MASK, NGP, LO>(nd, cur, B, NWE, acc, scale);
#pragma unroll
                for (int k = 0; k < NS; ++k) acc[k] = ACC(0);
            }
            cur = c;
        }
        const float4 d = pd[p];
        float q = d.w;
        if constexpr (VA >= 0) {
            const float4 v = pv[p];
            q *= VA == 0 ? v.x : (VA == 1 ? v.y : v.z);
            if constexpr (VB >= 0) q *= (VB == 0 ? v.x : (VB == 1 ? v.y : v.z));
        }
        float wx[3], wy[3], wz[3];
        if constexpr (MASK & 1) link3(d.x, wx[0], wx[1], wx[2]);
        else if constexpr (NGP & 1) wx[0] = 1.0f;
        else { wx[0] = 1.0f - d.x; wx[1] = d.x; }
        if constexpr (MASK & 2) link3(d.y, wy[0], wy[1], wy[2]);
        else if constexpr (NGP & 2) wy[0] = 1.0f;
        else { wy[0] = 1.0f - d.y; wy[1] = d.y; }
        if constexpr (MASK & 4) link3(d.z, wz[0], wz[1], wz[2]);
        else if constexpr (NGP & 4) wz[0] = 1.0f;
        else { wz[0] = 1.0f - d.z; wz[1] = d.z; }
#pragma unroll
        for (int kz = 0; kz < SZ; ++kz)
#pragma unroll
            for (int ky = 0; ky < SY; ++ky)
#pragma unroll
                for (int kx = 0; kx < SX; ++kx)
                    acc[kx + SX * (ky + SY * kz)] += AccOps<ACC>::term(q, wx[kx] * wy[ky] * wz[kz], A.fxs[COMP]);
    }
    // the run's last cell: summed over the lanes that share it (uniform branch)
    if (any_shared) seg_scan(ckey, acc);
    if (carry_end) window_add<MASK, NGP, LO>(nd, cur, B, NWE, acc, scale);
}

// item maxima: |q| and |v_d| as the bits of non-negative floats (ordered like the values)
struct DepMax {
    unsigned int q, v[3];
};

// per-item fixed-point scale of one component: 2^(51 - e) with sum_p |q_p v_a v_b| < 2^e
template <int KIND, int COMP>
__device__ __forceinline__ double window_scale(int n, const DepMax &m) {
    constexpr int VA = Comp<KIND, COMP>::VA, VB = Comp<KIND, COMP>::VB;
    double bound = (double)n * (double)__uint_as_float(m.q);
    if constexpr (VA >= 0) bound *= (double)__uint_as_float(m.v[VA]);
    if constexpr (VB >= 0) bound *= (double)__uint_as_float(m.v[VB]);
    if (!(bound > 0.0) || !(bound < 1e300)) return 1.0;  // all zero (or not finite)
    int e;
    frexp(bound, &e);
    return ldexp(1.0, 51 - e);
}

// grid: one workgroup per item.  LDS is what limits the workgroups per CU (the kernel
// slows down steeply with fewer), so the sorted particles are kept as planes of floats
// without a mass plane when the mass is global (PM false), with 16-bit cell keys, and the
// cell counts turn into starts in place.
template <int KIND, typename ACC, int BC, bool PM>
__device__ __forceinline__ void deposit_item(const DepArgs &A, int item) {
    constexpr int C = KindC<KIND>::C, CW = KindC<KIND>::CW;
    constexpr int LO = KindC<KIND>::LO;
    constexpr int NWMAX = MAXB + 1 - LO;
    constexpr int NW3 = NWMAX * NWMAX * NWMAX;
    constexpr int NCMAX = MAXB * MAXB * MAXB;
    constexpr bool FX = sizeof(ACC) == 8;  // global fixed point: terms quantised per particle
    static_assert(2 * DEP_WG >= NCMAX, "two cells per thread in the scan");
    static_assert(C % CW == 0, "whole window passes");
    // the sorted particles (offset planes, mass plane if PM, velocity planes if KIND > 0,
    // keys), then (aliased) the node window
    constexpr size_t PL_B = sizeof(float) * ITEM_P;
    constexpr int NPL = 3 + (PM ? 1 : 0) + (KIND > 0 ? 3 : 0);
    constexpr size_t SK_B = sizeof(unsigned short) * ITEM_P;
    constexpr size_t ND_B = sizeof(unsigned long long) * CW * NW3;
    constexpr size_t POOL_B = NPL * PL_B + SK_B > ND_B ? NPL * PL_B + SK_B : ND_B;
    __shared__ int cst[NCMAX];  // counts, then (in place) exclusive starts
    __shared__ int scan_sh[DEP_WG / 64];
    __shared__ DepMax smax;
    __shared__ __align__(16) unsigned char pool[POOL_B];
    float *pl = reinterpret_cast<float *>(pool);  // plane k at pl + k * ITEM_P: u.x u.y u.z [q] [v.x v.y v.z]
    constexpr int QPL = 3, VPL = PM ? 4 : 3;
    unsigned short *skey = reinterpret_cast<unsigned short *>(pool + NPL * PL_B);
    unsigned long long *nodes = reinterpret_cast<unsigned long long *>(pool);
    ItemRange r;
    if (!item_range(item, A.start, A.istart, A.item_bucket, A.g, r)) return;
    const int p0 = r.p0, n = r.p1 - r.p0;
    const int B = BC ? BC : A.g.B, N = A.g.N;
    const int NC = B * B * B, NWE = B + 1 - LO;
    const int lane = lane_id();
    for (int i = threadIdx.x; i < NC; i += DEP_WG) cst[i] = 0;
    if (threadIdx.x == 0) smax = DepMax{0u, {0u, 0u, 0u}};
    __syncthreads();
    // A. load and rank by cell; item maxima of |q| and |v|
    {
        int lc[DEP_PER], rk[DEP_PER];
        float4 qd[DEP_PER], qv[DEP_PER];
        float mq = 0.0f, mv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < DEP_PER; ++k) {
            const int i = threadIdx.x + k * DEP_WG;
            lc[k] = -1;
            if (i < n) {
                const int p = p0 + i;
                const int cell = A.s.cell[p];
                int lx, ly, lz;
                decode_cell(cell, A.g, lx, ly, lz);
                lc[k] = (lx - r.ox) + B * ((ly - r.oy) + B * (lz - r.oz));
                qd[k] = make_float4(A.s.off[0][p], A.s.off[1][p], A.s.off[2][p],
                                    PM ? A.s.mass[p] * A.mass_nv : A.mass_nv);
                mq = fmaxf(mq, fabsf(qd[k].w));
                if constexpr (KIND > 0) {
                    qv[k] = make_float4(A.s.vel[0][p], A.s.vel[1][p], A.s.vel[2][p], 0.0f);
                    mv[0] = fmaxf(mv[0], fabsf(qv[k].x));
                    mv[1] = fmaxf(mv[1], fabsf(qv[k].y));
                    mv[2] = fmaxf(mv[2], fabsf(qv[k].z));
                }
                rk[k] = atomicAdd(&cst[lc[k]], 1);
            }
        }
        if constexpr (!FX) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                mq = fmaxf(mq, __shfl_xor(mq, o));
                if constexpr (KIND > 0) {
                    mv[0] = fmaxf(mv[0], __shfl_xor(mv[0], o));
                    mv[1] = fmaxf(mv[1], __shfl_xor(mv[1], o));
                    mv[2] = fmaxf(mv[2], __shfl_xor(mv[2], o));
                }
            }
            if (lane == 0) {
                atomicMax(&smax.q, __float_as_uint(mq));
                if constexpr (KIND > 0) {
                    atomicMax(&smax.v[0], __float_as_uint(mv[0]));
                    atomicMax(&smax.v[1], __float_as_uint(mv[1]));
                    atomicMax(&smax.v[2], __float_as_uint(mv[2]));
                }
            }
        }
        __syncthreads();
        // exclusive scan of the NC <= 512 counts in place, two cells per thread (the scan's
        // first barrier orders every read before any write)
        {
            const int c0 = 2 * threadIdx.x, c1 = c0 + 1;
            const int a0 = c0 < NC ? cst[c0] : 0, a1 = c1 < NC ? cst[c1] : 0;
            int total;
            const int ex = block_excl_scan_fast(a0 + a1, scan_sh, total);
            if (c0 < NC) cst[c0] = ex;
            if (c1 < NC) cst[c1] = ex + a0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < DEP_PER; ++k) {
            if (lc[k] >= 0) {
                const int j = cst[lc[k]] + rk[k];
                skey[j] = (unsigned short)lc[k];
                pl[0 * ITEM_P + j] = qd[k].x;
                pl[1 * ITEM_P + j] = qd[k].y;
                pl[2 * ITEM_P + j] = qd[k].z;
                if constexpr (PM) pl[QPL * ITEM_P + j] = qd[k].w;
                if constexpr (KIND > 0) {
                    pl[(VPL + 0) * ITEM_P + j] = qv[k].x;
                    pl[(VPL + 1) * ITEM_P + j] = qv[k].y;
                    pl[(VPL + 2) * ITEM_P + j] = qv[k].z;
                }
            }
        }
    }
    __syncthreads();
    // B. lane runs of DEP_PER consecutive sorted particles, held in registers (the node
    // window then takes over their LDS)
    // (whole 16-byte rows of each plane: the run starts at a multiple of DEP_PER = 4; values
    // past the item's end are masked by their key)
    static_assert(DEP_PER == 4, "runs are read as 16-byte rows");
    int key[DEP_PER];
    float4 pd[DEP_PER], pv[DEP_PER];
    const int j0 = DEP_PER * threadIdx.x;
    {
        const float4 ux = *reinterpret_cast<const float4 *>(pl + 0 * ITEM_P + j0);
        const float4 uy = *reinterpret_cast<const float4 *>(pl + 1 * ITEM_P + j0);
        const float4 uz = *reinterpret_cast<const float4 *>(pl + 2 * ITEM_P + j0);
        float4 uq = make_float4(A.mass_nv, A.mass_nv, A.mass_nv, A.mass_nv);
        if constexpr (PM) uq = *reinterpret_cast<const float4 *>(pl + QPL * ITEM_P + j0);
        const ushort4 kk = *reinterpret_cast<const ushort4 *>(skey + j0);
        pd[0] = make_float4(ux.x, uy.x, uz.x, uq.x);
        pd[1] = make_float4(ux.y, uy.y, uz.y, uq.y);
        pd[2] = make_float4(ux.z, uy.z, uz.z, uq.z);
        pd[3] = make_float4(ux.w, uy.w, uz.w, uq.w);
        key[0] = j0 + 0 < n ? (int)kk.x : -1;
        key[1] = j0 + 1 < n ? (int)kk.y : -1;
        key[2] = j0 + 2 < n ? (int)kk.z : -1;
        key[3] = j0 + 3 < n ? (int)kk.w : -1;
        if constexpr (KIND > 0) {
            const float4 vx = *reinterpret_cast<const float4 *>(pl + (VPL + 0) * ITEM_P + j0);
            const float4 vy = *reinterpret_cast<const float4 *>(pl + (VPL + 1) * ITEM_P + j0);
            const float4 vz = *reinterpret_cast<const float4 *>(pl + (VPL + 2) * ITEM_P + j0);
            pv[0] = make_float4(vx.x, vy.x, vz.x, 0.f);
            pv[1] = make_float4(vx.y, vy.y, vz.y, 0.f);
            pv[2] = make_float4(vx.z, vy.z, vz.z, 0.f);
            pv[3] = make_float4(vx.w, vy.w, vz.w, 0.f);
        }
    }
    const DepMax m = smax;
    __syncthreads();
    for (int i = threadIdx.x; i < CW * NW3; i += DEP_WG) nodes[i] = 0ull;
    int ckey = -1;  // the run's last cell
#pragma unroll
    for (int k = 0; k < DEP_PER; ++k) ckey = key[k] >= 0 ? key[k] : ckey;
    const int next = __shfl_down(ckey, 1);
    const int prev = __shfl_up(ckey, 1);
    const bool any_shared = __ballot(lane > 0 && ckey >= 0 && prev == ckey) != 0ull;
    const bool carry_end = ckey >= 0 && (lane == 63 || next != ckey);
    __syncthreads();
    double sc[C];
#define LF2MI_DEP_COMP(K, I)                                                                               \
    sc[I] = FX ? 1.0 : window_scale<K, I>(n, m);                                                           \
    deposit_component<K, I, ACC, BC>(A, B, NWE, key, pd, pv, ckey, any_shared, carry_end, nodes, sc[I]);
    if constexpr (KIND == 0) {
        LF2MI_DEP_COMP(0, 0)
    } else {
        LF2MI_DEP_COMP(KIND, 0)
        LF2MI_DEP_COMP(KIND, 1)
        LF2MI_DEP_COMP(KIND, 2)
    }
    __syncthreads();
    // C. the node window into the field, component-fastest (consecutive addresses of the
    // interleaved field), zeros skipped; components c0 .. c0 + CW - 1
    const int nn = NWE * NWE * NWE;
    auto flush = [&](int c0) {
        for (int i = threadIdx.x; i < CW * nn; i += DEP_WG) {
            const int mnode = i / CW, cw = i - mnode * CW, c = c0 + cw;
            const unsigned long long w = nodes[cw * NW3 + mnode];
            if (w == 0ull) continue;
            const int ax = mnode % NWE + LO, ay = (mnode / NWE) % NWE + LO, az = mnode / (NWE * NWE) + LO;
            int gx = r.ox + ax, gy = r.oy + ay, gz = r.oz + az;
            if (A.tgt == LF2MI_DEP_PERIODIC) {
                gx = wrapi(gx, N);
                gy = wrapi(gy, N);
                gz = wrapi(gz, N);
            }
            const int64_t idx = site_index(N, A.h, gx, gy, gz) * C + c;
            if constexpr (FX) {
                AccOps<ACC>::flush(A, idx, w);
            } else {
                double inv = 1.0;
#pragma unroll
                for (int k = 0; k < C; ++k)
                    if (k == c) inv = 1.0 / sc[k];
                atomicAdd(&A.field[idx], (float)((double)(long long)w * inv));
            }
        }
    };
    flush(0);
    if constexpr (C == 6) {
        // the window's second pass: T_11, T_12, T_22
        __syncthreads();
        for (int i = threadIdx.x; i < CW * NW3; i += DEP_WG) nodes[i] = 0ull;
        __syncthreads();
        LF2MI_DEP_COMP(KIND, 3)
        LF2MI_DEP_COMP(KIND, 4)
        LF2MI_DEP_COMP(KIND, 5)
        __syncthreads();
        flush(3);
    }
#undef LF2MI_DEP_COMP
}

// one item per workgroup, or (A.dlist) the items of a work list, grid-stride
template <int KIND, typename ACC, int BC, bool PM>
__global__ __launch_bounds__(DEP_WG) void deposit_f32_kernel(DepArgs A) {
    if (A.dlist) {
        const int cnt = *A.dcount;
        for (int it = blockIdx.x; it < cnt; it += gridDim.x) {
            deposit_item<KIND, ACC, BC, PM>(A, A.dlist[it]);
            __syncthreads();  // the LDS is reused
        }
    } else {
        deposit_item<KIND, ACC, BC, PM>(A, blockIdx.x);
    }
}

// Scalar deposit with a global mass: a workgroup takes up to DEP_MERGE consecutive items of
// one brick (the first of each group; the others exit) and sorts them one after the other
// into the same node window, so a dense brick's nodes are added to the field once per group
// rather than once per item.  With one mass the fixed-point bound is n m, known before any
// particle is read.  Phases as in deposit_f32_kernel; the window has its own LDS here.
static constexpr int DEP_MERGE = 8;
// consecutive items of a group sorted together (a lane's run holds DEP_SORT * DEP_PER sorted
// particles): 1, since the larger LDS planes of 2 or 4 cost more occupancy than the longer
// runs save
static constexpr int DEP_SORT = 1;

template <typename ACC, int BC>
__global__ __launch_bounds__(DEP_WG) void deposit_merged_kernel(DepArgs A) {
    constexpr int NWMAX = MAXB + 1;
    constexpr int NW3 = NWMAX * NWMAX * NWMAX;
    constexpr int NCMAX = MAXB * MAXB * MAXB;
    constexpr bool FX = sizeof(ACC) == 8;
    constexpr int SP = DEP_SORT * ITEM_P;  // particles per sort
    constexpr int PER = SP / DEP_WG;       // sorted particles per lane run
    static_assert(PER % 4 == 0, "runs are read as 16-byte rows");
    static_assert(DEP_MERGE % DEP_SORT == 0, "whole sorts per group");
    static_assert(2 * DEP_WG >= NCMAX, "two cells per thread in the scan");
    __shared__ int cst[NCMAX];
    __shared__ int scan_sh[DEP_WG / 64];
    __shared__ __align__(16) float pl[3][SP];
    __shared__ __align__(16) unsigned short skey[SP];
    __shared__ unsigned long long nodes[NW3];
    ItemRange r;
    if (!item_range(blockIdx.x, A.start, A.istart, A.item_bucket, A.g, r)) return;
    const int bstart = A.start[r.b], bend = A.start[r.b + 1];
    if (((r.p0 - bstart) / ITEM_P) % DEP_MERGE != 0) return;  // taken by the group's first item
    const int end = min(bend, r.p0 + DEP_MERGE * ITEM_P);
    const int B = BC ? BC : A.g.B, N = A.g.N;
    const int NC = B * B * B, NWE = B + 1;
    const int lane = lane_id();
    const DepMax m{__float_as_uint(A.mass_nv), {0u, 0u, 0u}};
    const double sc = FX ? 1.0 : window_scale<0, 0>(end - r.p0, m);
    for (int i = threadIdx.x; i < NW3; i += DEP_WG) nodes[i] = 0ull;
    // the next sort's particles are loaded before the current sort's lane runs
    int nc[PER];
    float3 nq[PER];
    auto fetch = [&](int q0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int p = q0 + threadIdx.x + k * DEP_WG;
            nc[k] = -1;
            if (p < end) {
                nc[k] = A.s.cell[p];
                nq[k] = make_float3(A.s.off[0][p], A.s.off[1][p], A.s.off[2][p]);
            }
        }
    };
    fetch(r.p0);
    for (int p0 = r.p0; p0 < end; p0 += SP) {
        const int n = min(SP, end - p0);
        for (int i = threadIdx.x; i < NC; i += DEP_WG) cst[i] = 0;
        __syncthreads();
        // A. rank by cell
        {
            int lc[PER], rk[PER];
            float3 qd[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                lc[k] = -1;
                qd[k] = nq[k];
                if (nc[k] >= 0) {
                    int lx, ly, lz;
                    decode_cell(nc[k], A.g, lx, ly, lz);
                    lc[k] = (lx - r.ox) + B * ((ly - r.oy) + B * (lz - r.oz));
                    rk[k] = atomicAdd(&cst[lc[k]], 1);
                }
            }
            __syncthreads();
            {
                const int c0 = 2 * threadIdx.x, c1 = c0 + 1;
                const int a0 = c0 < NC ? cst[c0] : 0, a1 = c1 < NC ? cst[c1] : 0;
                int total;
                const int ex = block_excl_scan_fast(a0 + a1, scan_sh, total);
                if (c0 < NC) cst[c0] = ex;
                if (c1 < NC) cst[c1] = ex + a0;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (lc[k] >= 0) {
                    const int j = cst[lc[k]] + rk[k];
                    skey[j] = (unsigned short)lc[k];
                    pl[0][j] = qd[k].x;
                    pl[1][j] = qd[k].y;
                    pl[2][j] = qd[k].z;
                }
            }
        }
        __syncthreads();
        // B. lane runs into the window (16-byte rows of each plane; values past the end
        // are masked by their key)
        int key[PER];
        float4 pd[PER], pv[PER];
        const int j0 = PER * threadIdx.x;
        const float q = A.mass_nv;
#pragma unroll
        for (int h = 0; h < PER / 4; ++h) {
            const int jr = j0 + 4 * h;
            const float4 ux = *reinterpret_cast<const float4 *>(&pl[0][jr]);
            const float4 uy = *reinterpret_cast<const float4 *>(&pl[1][jr]);
            const float4 uz = *reinterpret_cast<const float4 *>(&pl[2][jr]);
            const ushort4 kk = *reinterpret_cast<const ushort4 *>(&skey[jr]);
            pd[4 * h + 0] = make_float4(ux.x, uy.x, uz.x, q);
            pd[4 * h + 1] = make_float4(ux.y, uy.y, uz.y, q);
            pd[4 * h + 2] = make_float4(ux.z, uy.z, uz.z, q);
            pd[4 * h + 3] = make_float4(ux.w, uy.w, uz.w, q);
            key[4 * h + 0] = jr + 0 < n ? (int)kk.x : -1;
            key[4 * h + 1] = jr + 1 < n ? (int)kk.y : -1;
            key[4 * h + 2] = jr + 2 < n ? (int)kk.z : -1;
            key[4 * h + 3] = jr + 3 < n ? (int)kk.w : -1;
        }
        if (p0 + SP < end) fetch(p0 + SP);
        int ckey = -1;
#pragma unroll
        for (int k = 0; k < PER; ++k) ckey = key[k] >= 0 ? key[k] : ckey;
        const int next = __shfl_down(ckey, 1);
        const int prev = __shfl_up(ckey, 1);
        const bool any_shared = __ballot(lane > 0 && ckey >= 0 && prev == ckey) != 0ull;
        const bool carry_end = ckey >= 0 && (lane == 63 || next != ckey);
        deposit_component<0, 0, ACC, BC, PER>(A, B, NWE, key, pd, pv, ckey, any_shared, carry_end, nodes, sc);
        __syncthreads();  // the planes and counts are reused by the next sort
    }
    // C. the window into the field
    const int nn = NWE * NWE * NWE;
    const double inv = 1.0 / sc;
    for (int i = threadIdx.x; i < nn; i += DEP_WG) {
        const unsigned long long w = nodes[i];
        if (w == 0ull) continue;
        const int ax = i % NWE, ay = (i / NWE) % NWE, az = i / (NWE * NWE);
        int gx = r.ox + ax, gy = r.oy + ay, gz = r.oz + az;
        if (A.tgt == LF2MI_DEP_PERIODIC) {
            gx = wrapi(gx, N);
            gy = wrapi(gy, N);
            gz = wrapi(gz, N);
        }
        const int64_t idx = site_index(N, A.h, gx, gy, gz);
        if constexpr (FX) AccOps<ACC>::flush(A, idx, w);
        else atomicAdd(&A.field[idx], (float)((double)(long long)w * inv));
    }
}

// ---------------------------------------------------------------- scalar deposit, periodic target
// Three kernels replace the item kernels when the target is periodic, the brick edge 8, the
// number of bricks per row a multiple of TILE_B and the mass global (the PM step's deposit).
// A brick's nodes are [0,8)^3 relative to its origin; its particles also reach the first
// node layer of the bricks above it along x, y and z.
//  * deposit_tiles_kernel: one workgroup per tile of TILE_B bricks along x, one wave per
//    brick.  A brick with fewer than `dense_min` particles sums their terms in the wave's LDS
//    window of 9^3 nodes (a denser brick leaves it empty).  Within the tile each brick's
//    x = 8 plane is added into the next brick's x = 0 plane on chip; the tile's own nodes are
//    then written once, TILE_X consecutive x per row (stored, or added to the field).  The layer
//    that belongs to the bricks above in y or z goes to two shell buffers and the tile's last
//    x = 8 plane to a seam buffer, one value per node.  No global atomics, no sort.
//  * dense_groups_kernel + dense_body: the bricks with dense_min or more particles,
//    in groups of up to DEP_MERGE items (a work list shared out statically among persistent
//    workgroups); each particle's 8 terms go straight into a 9^3 LDS window (no sort: a
//    wave's 64 particles of a crowded brick rarely share a cell), added to the field with one
//    global atomic per non-zero node after the tiles kernel's stores.
//  * fold_node: every node in the lowest y or z layer of a brick, or in the first x plane of a
//    tile, adds the shell and seam values meant for it, one thread per node, x fastest.  The
//    first FOLD_BLOCKS workgroups of the dense kernel's launch do it with global atomics
//    while the others deposit the dense bricks (deposit_dense_fold_kernel): one is memory
//    bound, the other LDS bound.
// Window sums are exact 64-bit fixed point: per term a power-of-two scale 2^(51 - e) of a
// bound n q < 2^e (ACC float: n = dense_min for the tiles, whose x = 0 planes then take at
// most twice that, the group's count for the dense kernel), or the term quantised on its own
// (ACC fixed point, order independent).
static constexpr int TILE_B = 8;           // bricks per tile (one wave each)
static constexpr int TILE_X = 8 * TILE_B;  // nodes per tile row

// a particle's term for node weight w: float accumulation takes q w times the window scale
// in FP64 (qsc = q * scale, one FMA against 1.5 2^52 as in to_window); the fixed point
// quantises fl(q w) to the global scale on its own
template <typename ACC>
__device__ __forceinline__ unsigned long long dep_term(float q, float w, double qsc, double fxs) {
    if constexpr (sizeof(ACC) == 8) {
        return AccOps<ACC>::term(q, w, fxs);
    } else {
        const double t = fma((double)w, qsc, 0x1.8p52);
        return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000ll);
    }
}

// scale of a window whose bound is n particles of value q (float accumulation)
__device__ __forceinline__ double bound_scale(int n, float q) {
    const double bound = (double)n * (double)fabsf(q);
    if (!(bound > 0.0) || !(bound < 1e300)) return 1.0;
    int e;
    frexp(bound, &e);
    return ldexp(1.0, 51 - e);
}

// Shell and seam buffers (row = by + NB bz, X global, tile = X / TILE_X):
//   ysh[row][z in 0..8][X]: nodes with y = 8 of the row's bricks (they belong to row by + 1);
//   zsh[row][y in 0..7][X]: nodes with z = 8, y < 8 (row bz + 1);
//   xsh[row][tile][y + 9 z]: the x = 8 plane of the tile's last brick, addressed by the tile
//   whose first plane it is (the next one, periodically).
template <typename ACC>
__global__ __launch_bounds__(64 * TILE_B) void deposit_tiles_kernel(DepArgs A, int dense_min, int set,
                                                                    ACC *__restrict__ ysh, ACC *__restrict__ zsh,
                                                                    ACC *__restrict__ xsh) {
    constexpr int B = 8, NWE = 9, NW3 = NWE * NWE * NWE, PL = NWE * NWE, SP = 4;
    constexpr bool FX = sizeof(ACC) == 8;
    static_assert(TILE_X == 64, "one lane per x of the tile");
    __shared__ unsigned long long win[TILE_B][NW3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int NB = A.g.NB, N = A.g.N, G = N + 2 * A.h;
    const int ntile = NB / TILE_B;
    const int tile = blockIdx.x % ntile, row = blockIdx.x / ntile;
    const int by = row % NB, bz = row / NB;
    const int64_t brick = (int64_t)row * NB + tile * TILE_B + wave;  // bricks are numbered x fastest
    unsigned long long *w = win[wave];
    for (int i = lane; i < NW3; i += 64) w[i] = 0ull;
    const float q = A.mass_nv;
    const double sc = FX ? 1.0 : bound_scale(dense_min, q), qsc = (double)q * sc, inv = 1.0 / sc;
    const int a = A.start[brick], n = A.start[brick + 1] - a;
    if (n < dense_min) {
        for (int j0 = 0; j0 < n; j0 += 64 * SP) {
            int c[SP];
            float ux[SP], uy[SP], uz[SP];
#pragma unroll
            for (int k = 0; k < SP; ++k) {
                const int j = j0 + lane + 64 * k;
                c[k] = -1;
                if (j < n) {
                    c[k] = A.s.cell[a + j];
                    ux[k] = A.s.off[0][a + j];
                    uy[k] = A.s.off[1][a + j];
                    uz[k] = A.s.off[2][a + j];
                }
            }
#pragma unroll
            for (int k = 0; k < SP; ++k) {
                if (c[k] < 0) continue;
                int x, y, z;
                decode_cell(c[k], A.g, x, y, z);
                const int base = (x & (B - 1)) + NWE * ((y & (B - 1)) + NWE * (z & (B - 1)));
                const float wx[2] = {1.0f - ux[k], ux[k]}, wy[2] = {1.0f - uy[k], uy[k]},
                            wz[2] = {1.0f - uz[k], uz[k]};
#pragma unroll
                for (int cz = 0; cz < 2; ++cz)
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                        for (int cx = 0; cx < 2; ++cx)
                            atomicAdd(&w[base + cx + NWE * (cy + NWE * cz)],
                                      dep_term<ACC>(q, wx[cx] * wy[cy] * wz[cz], qsc, A.fxs[0]));
            }
        }
    }
    __syncthreads();
    // the previous brick's x = 8 plane joins this brick's x = 0 plane
    if (wave > 0)
        for (int yz = lane; yz < PL; yz += 64) w[NWE * yz] += win[wave - 1][B + NWE * yz];
    __syncthreads();
    auto cvt = [&](unsigned long long v) -> ACC {
        if constexpr (FX) return v;
        else return (float)((double)(long long)v * inv);
    };
    const int X0 = tile * TILE_X;
    // own nodes: lane = x within the tile, wave = y, rows z = 0 .. 7
    {
        const int64_t i0 = site_index(N, A.h, X0 + lane, by * B + wave, bz * B);
        const unsigned long long *src = &win[lane / B][(lane % B) + NWE * wave];
#pragma unroll
        for (int z = 0; z < B; ++z) {
            const unsigned long long v = src[NWE * NWE * z];
            const int64_t idx = i0 + (int64_t)z * G * G;
            if constexpr (FX) {
                if (set) A.fx[idx] = v;
                else if (v) A.fx[idx] += v;
            } else {
                const float f = cvt(v);
                if (set) A.field[idx] = f;
                else if (v) A.field[idx] += f;
            }
        }
    }
    // y = 8 layer (rows z = 0 .. 8), then z = 8 layer (rows y = 0 .. 7): row r = wave + 8 j
    for (int r = wave; r < NWE + B; r += TILE_B) {
        const int y = r < NWE ? B : r - NWE, z = r < NWE ? r : B;
        const ACC v = cvt(win[lane / B][(lane % B) + NWE * (y + NWE * z)]);
        if (r < NWE) ysh[((int64_t)row * NWE + z) * N + X0 + lane] = v;
        else zsh[((int64_t)row * B + y) * N + X0 + lane] = v;
    }
    // the tile's last x = 8 plane, for the next tile's first plane
    const int nt = tile + 1 == ntile ? 0 : tile + 1;
    for (int yz = threadIdx.x; yz < PL; yz += 64 * TILE_B)
        xsh[((int64_t)row * ntile + nt) * PL + yz] = cvt(win[TILE_B - 1][B + NWE * yz]);
}

// one thread per node that receives shell or seam values, x fastest: the lowest y layer of
// each brick (N x NB x N nodes), then the lowest z layer without it (N x 7 NB x NB), then the
// first x plane of each tile without both (ntile x 7 NB x 7 NB, x slowest)
// (with global atomics, so that the dense deposit may run at the same time)
template <typename ACC>
__device__ __forceinline__ void fold_node(const DepArgs &A, const ACC *__restrict__ ysh, const ACC *__restrict__ zsh,
                                          const ACC *__restrict__ xsh, unsigned t) {
    constexpr int B = 8, NWE = 9, PL = NWE * NWE;
    constexpr bool FX = sizeof(ACC) == 8;
    const unsigned N = A.g.N, NB = A.g.NB, ntile = NB / TILE_B;
    const unsigned ny = N * NB * N, nz = N * (N - NB) * NB, nx = ntile * (N - NB) * (N - NB);
    if (t >= ny + nz + nx) return;
    // divisions by N and NB as shifts when N is a power of two (wave-uniform branch)
    const int lg = A.g.lgN;
    auto divN = [&](unsigned v) { return lg >= 0 ? v >> lg : v / N; };
    auto divNB = [&](unsigned v) { return lg >= 0 ? v >> (lg - 3) : v / NB; };
    unsigned X, Y, Z;
    if (t < ny) {
        const unsigned r = divN(t);
        X = t - r * N;
        Z = divNB(r);
        Y = (r - Z * NB) * B;
    } else if (t < ny + nz) {
        // r = yy + 7 NB zb: yy = 7 yb + k, k in 0..6
        const unsigned u = t - ny, r = divN(u);
        X = u - r * N;
        const unsigned zb = divNB(r) / (B - 1), yy = r - zb * (N - NB), yb = yy / (B - 1);
        Y = yb * B + 1 + (yy - yb * (B - 1));
        Z = zb * B;
    } else {
        const unsigned u = t - ny - nz, m = N - NB;
        const unsigned r = divNB(u) / (B - 1), yy = u - r * m;  // r = zz + m tile
        const unsigned tl = divNB(r) / (B - 1), zz = r - tl * m;
        const unsigned yb = yy / (B - 1), zb = zz / (B - 1);
        X = tl * TILE_X;
        Y = yb * B + 1 + (yy - yb * (B - 1));
        Z = zb * B + 1 + (zz - zb * (B - 1));
    }
    const unsigned by = Y / B, bz = Z / B, ly = Y % B, lz = Z % B;
    const unsigned bym = by == 0 ? NB - 1 : by - 1, bzm = bz == 0 ? NB - 1 : bz - 1;
    ACC v = ACC(0);
    if (ly == 0) {
        v += ysh[((int64_t)(bym + NB * bz) * NWE + lz) * N + X];
        if (lz == 0) v += ysh[((int64_t)(bym + NB * bzm) * NWE + B) * N + X];
    }
    if (lz == 0) v += zsh[((int64_t)(by + NB * bzm) * B + ly) * N + X];
    if (X % TILE_X == 0) {
        const unsigned tl = X / TILE_X;
        v += xsh[((int64_t)(by + NB * bz) * ntile + tl) * PL + ly + NWE * lz];
        if (ly == 0) v += xsh[((int64_t)(bym + NB * bz) * ntile + tl) * PL + B + NWE * lz];
        if (lz == 0) v += xsh[((int64_t)(by + NB * bzm) * ntile + tl) * PL + ly + NWE * B];
        if (ly == 0 && lz == 0) v += xsh[((int64_t)(bym + NB * bzm) * ntile + tl) * PL + B + NWE * B];
    }
    if (v == ACC(0)) return;
    const int64_t idx = site_index((int)N, A.h, (int)X, (int)Y, (int)Z);
    if constexpr (FX) atomicAdd(&A.fx[idx], v);
    else atomicAdd(&A.field[idx], v);
}

// number of nodes fold_node serves
__host__ __device__ inline unsigned fold_nodes(unsigned N, unsigned NB) {
    return N * NB * N + N * (N - NB) * NB + (NB / TILE_B) * (N - NB) * (N - NB);
}


// work list of the dense kernel: (brick, group) pairs of the bricks with dense_min or more
// particles, a group being up to DEP_MERGE * ITEM_P consecutive particles of the brick
// one atomic per workgroup of DG_WG threads (a scan over the workgroup)
static constexpr int DG_WG = 1024;
__global__ __launch_bounds__(DG_WG) void dense_groups_kernel(const int *__restrict__ start, int64_t nbk,
                                                             int dense_min, int2 *__restrict__ groups,
                                                             int *__restrict__ ngroups) {
    constexpr int GP = DEP_MERGE * ITEM_P;
    __shared__ int wsum[DG_WG / 64];
    __shared__ int wbase;
    const int64_t b = (int64_t)blockIdx.x * DG_WG + threadIdx.x;
    int g = 0;
    if (b < nbk) {
        const int n = start[b + 1] - start[b];
        if (n >= dense_min) g = (n + GP - 1) / GP;
    }
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    int incl = g;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < DG_WG / 64; ++w) {
            const int v = wsum[w];
            wsum[w] = tot;
            tot += v;
        }
        wbase = tot > 0 ? atomicAdd(ngroups, tot) : 0;
    }
    __syncthreads();
    const int base = wbase + wsum[wave] + incl - g;
    for (int k = 0; k < g; ++k) groups[base + k] = make_int2((int)b, k);
}

// work list of the item kernel: the items of the bricks with dense_min or more particles
// (one atomic per workgroup of DG_WG threads)
__global__ __launch_bounds__(DG_WG) void dense_items_kernel(const int *__restrict__ start,
                                                            const int *__restrict__ istart, int64_t nbk,
                                                            int dense_min, int *__restrict__ items,
                                                            int *__restrict__ nitems) {
    __shared__ int wsum[DG_WG / 64];
    __shared__ int wbase;
    const int64_t b = (int64_t)blockIdx.x * DG_WG + threadIdx.x;
    int g = 0, i0 = 0;
    if (b < nbk && start[b + 1] - start[b] >= dense_min) {
        i0 = istart[b];
        g = istart[b + 1] - i0;
    }
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    int incl = g;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < DG_WG / 64; ++w) {
            const int v = wsum[w];
            wsum[w] = tot;
            tot += v;
        }
        wbase = tot > 0 ? atomicAdd(nitems, tot) : 0;
    }
    __syncthreads();
    const int base = wbase + wsum[wave] + incl - g;
    for (int k = 0; k < g; ++k) items[base + k] = i0 + k;
}

static constexpr int DENSE_PER = 4;  // particles per thread and round

// persistent: workgroup w takes entries w, w + grid, ... of the work list (full groups are
// alike, so a static share balances); each round's loads are issued before the previous
// round's terms are added
template <typename ACC>
__device__ __forceinline__ void dense_body(const DepArgs &A, const int2 *__restrict__ groups,
                                           const int *__restrict__ ngroups, int bid, int nbid) {
    constexpr int B = 8, NWE = B + 1, NW3 = NWE * NWE * NWE;
    constexpr bool FX = sizeof(ACC) == 8;
    constexpr int GP = DEP_MERGE * ITEM_P;
    constexpr int RP = DENSE_PER * DEP_WG;  // particles per round
    __shared__ unsigned long long nodes[NW3];
    const int N = A.g.N, NB = A.g.NB;
    const float q = A.mass_nv;
    const int total = ngroups[0];
    for (int gi = bid; gi < total; gi += nbid) {
        for (int i = threadIdx.x; i < NW3; i += DEP_WG) nodes[i] = 0ull;
        const int2 gr = groups[gi];
        const int b = gr.x;
        const int bstart = A.start[b], bend = A.start[b + 1];
        const int p0g = bstart + gr.y * GP, end = min(bend, p0g + GP);
        const int ox = (b % NB) * B, oy = ((b / NB) % NB) * B, oz = (int)(b / ((int64_t)NB * NB)) * B;
        const double sc = FX ? 1.0 : bound_scale(end - p0g, q), qsc = (double)q * sc;
        int c[DENSE_PER];
        float ux[DENSE_PER], uy[DENSE_PER], uz[DENSE_PER];
        auto fetch = [&](int p0) {
#pragma unroll
            for (int k = 0; k < DENSE_PER; ++k) {
                const int p = p0 + threadIdx.x + k * DEP_WG;
                c[k] = -1;
                if (p < end) {
                    c[k] = A.s.cell[p];
                    ux[k] = A.s.off[0][p];
                    uy[k] = A.s.off[1][p];
                    uz[k] = A.s.off[2][p];
                }
            }
        };
        fetch(p0g);
        __syncthreads();  // the window is cleared
        for (int p0 = p0g; p0 < end; p0 += RP) {
            int cc[DENSE_PER];
            float vx[DENSE_PER], vy[DENSE_PER], vz[DENSE_PER];
#pragma unroll
            for (int k = 0; k < DENSE_PER; ++k) {
                cc[k] = c[k];
                vx[k] = ux[k];
                vy[k] = uy[k];
                vz[k] = uz[k];
            }
            if (p0 + RP < end) fetch(p0 + RP);
#pragma unroll
            for (int k = 0; k < DENSE_PER; ++k) {
                if (cc[k] < 0) continue;
                int x, y, z;
                decode_cell(cc[k], A.g, x, y, z);
                const int base = (x & (B - 1)) + NWE * ((y & (B - 1)) + NWE * (z & (B - 1)));
                const float wx[2] = {1.0f - vx[k], vx[k]}, wy[2] = {1.0f - vy[k], vy[k]},
                            wz[2] = {1.0f - vz[k], vz[k]};
#pragma unroll
                for (int cz = 0; cz < 2; ++cz)
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                        for (int cx = 0; cx < 2; ++cx)
                            atomicAdd(&nodes[base + cx + NWE * (cy + NWE * cz)],
                                      dep_term<ACC>(q, wx[cx] * wy[cy] * wz[cz], qsc, A.fxs[0]));
            }
        }
        __syncthreads();
        const double inv = 1.0 / sc;
        for (int i = threadIdx.x; i < NW3; i += DEP_WG) {
            const unsigned long long w = nodes[i];
            if (w == 0ull) continue;
            const int ax = i % NWE, ay = (i / NWE) % NWE, az = i / (NWE * NWE);
            const int64_t idx = site_index(N, A.h, wrapi(ox + ax, N), wrapi(oy + ay, N), wrapi(oz + az, N));
            if constexpr (FX) AccOps<ACC>::flush(A, idx, w);
            else atomicAdd(&A.field[idx], (float)((double)(long long)w * inv));
        }
        __syncthreads();  // the window is reused
    }
}

// the first `fold_blocks` workgroups add the shells with atomics, the others run the dense
// deposit: the two overlap (one memory bound, the other LDS bound)
template <typename ACC>
__global__ __launch_bounds__(DEP_WG) void deposit_dense_fold_kernel(DepArgs A, const int2 *__restrict__ groups,
                                                                    const int *__restrict__ ngroups,
                                                                    const ACC *__restrict__ ysh,
                                                                    const ACC *__restrict__ zsh,
                                                                    const ACC *__restrict__ xsh, int fold_blocks) {
    static_assert(DEP_WG == 256, "workgroup size of both parts");
    // without dense bricks (an even spread of particles) every workgroup folds
    const bool all_fold = ngroups[0] == 0;
    if ((int)blockIdx.x < fold_blocks || all_fold) {
        const unsigned n = fold_nodes(A.g.N, A.g.NB);
        const unsigned stride = (unsigned)(all_fold ? (int)gridDim.x : fold_blocks) * DEP_WG;
        for (unsigned t = blockIdx.x * DEP_WG + threadIdx.x; t < n; t += stride)
            fold_node<ACC>(A, ysh, zsh, xsh, t);
        return;
    }
    dense_body<ACC>(A, groups, ngroups, (int)blockIdx.x - fold_blocks, (int)gridDim.x - fold_blocks);
}

// ---------------------------------------------------------------- multi-component tiles
// The vector, VecVec and CIC-NGP kinds, and the scalar kind with per-particle masses, into
// a periodic target: the scheme of deposit_tiles_kernel with the components in passes.
// One wave per brick, TC_TB bricks along x per workgroup.  The brick's first 64 TC_P
// particles stay in registers across the passes (the rest are read again in each pass),
// and each pass
//   1. sums one component into the wave's LDS window (64-bit fixed point with integer LDS
//      atomics, the scale from the tile's maxima of |q| and |v|, as in deposit_tiles_kernel);
//   2. adds the x halo planes of the windows into the neighbouring waves' windows;
//   3. keeps the bricks' own nodes of that component in registers (each thread owns a
//      column of 8 nodes along z), which are stored (set) or added (without atomics) with
//      all components of a site together after the last pass, so that every line of the
//      interleaved field is written once;
//   4. writes the window nodes outside the bricks in y or z (shell rows of the tile's
//      consecutive x) and the tile's outer x halo planes (seams) to a buffer, which
//      fold_rows_kernel and fold_edges_kernel then add to the nodes they land on.
// A component's window spans nodes [LO, HI] per axis relative to the brick.  On a link axis
// the particle's two nonzero link weights sit on nodes x - 1, x or on x, x + 1
// (projections.hpp:1058-1071, whose third weight is zero), so LO = -1 and HI = 8; a vertex
// axis spans [0, 8], a nearest-grid-point axis [0, 7].
// Bricks with dense_min or more particles are left out (their windows stay zero) and are
// deposited afterwards by the item kernel over a work list of their items
// (dense_items_kernel, deposit_f32_kernel with A.dlist).
// (measured at 1024^3 with about 512 particles per brick: 4 bricks per tile and 2 particles
// per lane in registers, at most 128 VGPRs, against 8 bricks per tile or 8 particles per
// lane: within 5 %, the bricks' own nodes written once per line being what counted)
static constexpr int TC_P = 2;       // particles per lane kept in registers across the passes
static constexpr int TC_WMAX = 900;  // largest window: two link axes, 10 x 10 x 9 nodes
static constexpr int TC_TB = 4;      // bricks per tile (one wave each)

// halo layers below and above the brick along axis d, and the window edge
__host__ __device__ constexpr int tc_nl(int mask, int d) { return (mask >> d) & 1; }
__host__ __device__ constexpr int tc_nh(int ngp, int d) { return ((ngp >> d) & 1) ? 0 : 1; }
__host__ __device__ constexpr int tc_w(int mask, int ngp, int d) { return 8 + tc_nl(mask, d) + tc_nh(ngp, d); }
// shell rows of a component per brick row: the y halo layers over the window's z range,
// then the z halo layers over y = 0 .. 7
__host__ __device__ constexpr int tc_rows(int mask, int ngp) {
    return (tc_nl(mask, 1) + tc_nh(ngp, 1)) * tc_w(mask, ngp, 2) + (tc_nl(mask, 2) + tc_nh(ngp, 2)) * 8;
}
// seam values per tile: its x halo planes (y and z over the window)
__host__ __device__ constexpr int tc_seam(int mask, int ngp) {
    return (tc_nl(mask, 0) + tc_nh(ngp, 0)) * tc_w(mask, ngp, 1) * tc_w(mask, ngp, 2);
}
// Comp<KIND, c>::MASK and ::NGP as tables for the host
static constexpr int TC_MASK[5][6] = {{0}, {1, 2, 4}, {0, 3, 5, 0, 6, 0}, {0, 0, 0}, {0, 0, 0, 0, 0, 0}};
static constexpr int TC_NGP[5][6] = {{0}, {0, 0, 0}, {0, 0, 0, 0, 0, 0}, {1, 2, 4}, {0, 3, 5, 0, 6, 0}};
template <int KIND, int COMP> constexpr bool tc_table_matches() {
    return TC_MASK[KIND][COMP] == Comp<KIND, COMP>::MASK && TC_NGP[KIND][COMP] == Comp<KIND, COMP>::NGP;
}
static_assert(tc_table_matches<0, 0>() && tc_table_matches<1, 0>() && tc_table_matches<1, 1>() &&
                  tc_table_matches<1, 2>() && tc_table_matches<2, 0>() && tc_table_matches<2, 1>() &&
                  tc_table_matches<2, 2>() && tc_table_matches<2, 3>() && tc_table_matches<2, 4>() &&
                  tc_table_matches<2, 5>() && tc_table_matches<3, 0>() && tc_table_matches<3, 1>() &&
                  tc_table_matches<3, 2>() && tc_table_matches<4, 0>() && tc_table_matches<4, 1>() &&
                  tc_table_matches<4, 2>() && tc_table_matches<4, 3>() && tc_table_matches<4, 4>() &&
                  tc_table_matches<4, 5>(),
              "host tables of the component masks");

struct TcArgs {
    int64_t shell_off[6];  // first value of component c's shell rows in the buffer
    int64_t seam_off[6];   // first value of its seams
};

// the particles a lane keeps: cell in the brick (x + 8 (y + 8 z), -1 for none), offsets,
// velocity (vector kinds) and q = m N^3
struct TcRegs {
    static constexpr int P = TC_P > 0 ? TC_P : 1;
    int lc[P];
    float u[3][P], v[3][P], q[P];
};

template <int KIND, bool PM>
__device__ __forceinline__ void tc_load(const DepArgs &A, int p, int &lc, float &u0, float &u1, float &u2, float &v0,
                                        float &v1, float &v2, float &q) {
    int x, y, z;
    decode_cell(A.s.cell[p], A.g, x, y, z);
    lc = (x & 7) | ((y & 7) << 3) | ((z & 7) << 6);
    u0 = A.s.off[0][p];
    u1 = A.s.off[1][p];
    u2 = A.s.off[2][p];
    q = PM ? A.s.mass[p] * A.mass_nv : A.mass_nv;
    if constexpr (KIND > 0) {
        v0 = A.s.vel[0][p];
        v1 = A.s.vel[1][p];
        v2 = A.s.vel[2][p];
    } else {
        v0 = v1 = v2 = 0.0f;
    }
}

// weights of one axis on nodes b, b + 1 (b moves to x - 1 below a link's midpoint)
template <int MASK, int NGP, int D>
__device__ __forceinline__ void tc_weights(float u, int &b, float (&wt)[2]) {
    if constexpr ((MASK >> D) & 1) {
        const float t = u - 0.5f;
        if (t < 0.0f) {
            const float a = -t;
            b -= 1;
            wt[0] = a;
            wt[1] = 1.0f - a;
        } else {
            wt[0] = 1.0f - t;
            wt[1] = t;
        }
    } else if constexpr ((NGP >> D) & 1) {
        wt[0] = 1.0f;
        wt[1] = 0.0f;
    } else {
        wt[0] = 1.0f - u;
        wt[1] = u;
    }
}

// the window scale of a component: 2^(51 - e) with 2 n max|q| max|v_a| max|v_b| < 2^e, so
// that a node's sum, with the x halo of the neighbouring brick added, stays below 2^52 and
// each term below 2^51 (the FMA conversion of dep_term)
template <int KIND, int COMP>
__device__ __forceinline__ double tc_scale(int n, float mq, const float (&mv)[3]) {
    constexpr int VA = Comp<KIND, COMP>::VA, VB = Comp<KIND, COMP>::VB;
    double bound = 2.0 * (double)n * (double)mq;
    if constexpr (VA >= 0) bound *= (double)mv[VA];
    if constexpr (VB >= 0) bound *= (double)mv[VB];
    if (!(bound > 0.0) || !(bound < 1e300)) return 1.0;
    int e;
    frexp(bound, &e);
    return ldexp(1.0, 51 - e);
}

template <int KIND, int COMP, bool PM, int TB>
__device__ __forceinline__ void tc_pass(const DepArgs &A, const TcArgs &T, float *__restrict__ shell,
                                        unsigned long long (*win)[TC_WMAX], const TcRegs &R, int a, int n, double sc,
                                        float (&own)[KindC<KIND>::C][8], int tile, int row, int ntile) {
    constexpr int MASK = Comp<KIND, COMP>::MASK, NGP = Comp<KIND, COMP>::NGP;
    constexpr int VA = Comp<KIND, COMP>::VA, VB = Comp<KIND, COMP>::VB;
    constexpr int TX = 8 * TB;  // nodes per tile row
    constexpr int LX = -tc_nl(MASK, 0), LY = -tc_nl(MASK, 1), LZ = -tc_nl(MASK, 2);
    constexpr int WX = tc_w(MASK, NGP, 0), WY = tc_w(MASK, NGP, 1), WZ = tc_w(MASK, NGP, 2);
    constexpr int W3 = WX * WY * WZ, PL = WY * WZ;
    constexpr int SX = tc_nh(NGP, 0) ? 2 : 1, SY = tc_nh(NGP, 1) ? 2 : 1, SZ = tc_nh(NGP, 2) ? 2 : 1;
    static_assert(W3 <= TC_WMAX, "the window fits");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int N = A.g.N;
    unsigned long long *w = win[wave];
    __syncthreads();  // the previous pass is done reading every window
    for (int i = lane; i < W3; i += 64) w[i] = 0ull;
    auto add = [&](int c, float x, float y, float z, float q, float v0, float v1, float v2) {
        float qc = q;
        if constexpr (VA >= 0) qc *= VA == 0 ? v0 : (VA == 1 ? v1 : v2);
        if constexpr (VB >= 0) qc *= VB == 0 ? v0 : (VB == 1 ? v1 : v2);
        const double qsc = (double)qc * sc;
        int bx = c & 7, byy = (c >> 3) & 7, bzz = c >> 6;
        float wx[2], wy[2], wz[2];
        tc_weights<MASK, NGP, 0>(x, bx, wx);
        tc_weights<MASK, NGP, 1>(y, byy, wy);
        tc_weights<MASK, NGP, 2>(z, bzz, wz);
        const int base = (bx - LX) + WX * ((byy - LY) + WY * (bzz - LZ));
#pragma unroll
        for (int kz = 0; kz < SZ; ++kz)
#pragma unroll
            for (int ky = 0; ky < SY; ++ky)
#pragma unroll
                for (int kx = 0; kx < SX; ++kx)
                    atomicAdd(&w[base + kx + WX * (ky + WY * kz)], dep_term<float>(qc, wx[kx] * wy[ky] * wz[kz], qsc, 0.0));
    };
#pragma unroll
    for (int k = 0; k < TC_P; ++k)
        if (R.lc[k] >= 0) add(R.lc[k], R.u[0][k], R.u[1][k], R.u[2][k], R.q[k], R.v[0][k], R.v[1][k], R.v[2][k]);
    // the particles past the registers, two per lane and round
    for (int j0 = 64 * TC_P; j0 < n; j0 += 64 * 2) {
        int c[2];
        float u0[2], u1[2], u2[2], v0[2], v1[2], v2[2], q[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int j = j0 + lane + 64 * k;
            c[k] = -1;
            if (j < n) tc_load<KIND, PM>(A, a + j, c[k], u0[k], u1[k], u2[k], v0[k], v1[k], v2[k], q[k]);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
            if (c[k] >= 0) add(c[k], u0[k], u1[k], u2[k], q[k], v0[k], v1[k], v2[k]);
    }
    __syncthreads();
    // x halo planes between the tile's waves: the previous brick's x = 8 plane joins this
    // brick's x = 0 plane, the next brick's x = -1 plane this brick's x = 7 plane (each wave
    // writes only its own x = 0 and x = 7 planes and reads only halo planes)
    if constexpr (tc_nh(NGP, 0))
        if (wave > 0)
            for (int yz = lane; yz < PL; yz += 64) w[(0 - LX) + WX * yz] += win[wave - 1][(8 - LX) + WX * yz];
    if constexpr (tc_nl(MASK, 0))
        if (wave < TB - 1)
            for (int yz = lane; yz < PL; yz += 64) w[(7 - LX) + WX * yz] += win[wave + 1][(-1 - LX) + WX * yz];
    __syncthreads();
    const double inv = 1.0 / sc;
    auto cvt = [&](unsigned long long v) -> float { return (float)((double)(long long)v * inv); };
    const int X0 = tile * TX;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;  // this thread's x in the tile, and y
    // own nodes (x = tx, y = ty, z = 0 .. 7): kept until every component is done
    {
        const unsigned long long *src = &win[tx >> 3][((tx & 7) - LX) + WX * (ty - LY)];
#pragma unroll
        for (int z = 0; z < 8; ++z) own[COMP][z] = cvt(src[WX * WY * (z - LZ)]);
    }
    // shell rows: y halo layers (z over the window), then z halo layers (y = 0 .. 7)
    constexpr int RY = (tc_nl(MASK, 1) + tc_nh(NGP, 1)) * WZ, RR = tc_rows(MASK, NGP);
    float *sh = shell + T.shell_off[COMP] + (int64_t)row * RR * N + X0 + tx;
    for (int r = ty; r < RR; r += 8) {
        int y, z;
        if (r < RY) {
            const int hy = r / WZ;
            z = LZ + (r - hy * WZ);
            y = (tc_nl(MASK, 1) && hy == 0) ? -1 : 8;
        } else {
            const int s = r - RY, hz = s >> 3;
            y = s & 7;
            z = (tc_nl(MASK, 2) && hz == 0) ? -1 : 8;
        }
        sh[(int64_t)r * N] = cvt(win[tx >> 3][((tx & 7) - LX) + WX * ((y - LY) + WY * (z - LZ))]);
    }
    // seams: the x = -1 plane of the tile's first brick, the x = 8 plane of its last
    constexpr int NS = tc_nl(MASK, 0) + tc_nh(NGP, 0);
    if constexpr (NS > 0) {
        float *se = shell + T.seam_off[COMP] + ((int64_t)row * ntile + tile) * (NS * PL);
        for (int i = threadIdx.x; i < NS * PL; i += 64 * TB) {
            const int side = i / PL, yz = i - side * PL;
            const bool lower = tc_nl(MASK, 0) && side == 0;
            se[i] = cvt(lower ? win[0][(-1 - LX) + WX * yz] : win[TB - 1][(8 - LX) + WX * yz]);
        }
    }
}

// grid: one workgroup per tile (TB bricks along x) of every row of bricks
template <int KIND, bool PM, int TB>
__global__ __launch_bounds__(64 * TB, 4) void deposit_tiles_c_kernel(DepArgs A, int dense_min, int set, TcArgs T,
                                                                  float *__restrict__ shell) {
    constexpr int C = KindC<KIND>::C, TX = 8 * TB;
    __shared__ unsigned long long win[TB][TC_WMAX];
    __shared__ float smax[TB][4];
    __shared__ int sn[TB];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int NB = A.g.NB, N = A.g.N, G = N + 2 * A.h, ntile = NB / TB;
    const int tile = blockIdx.x % ntile, row = blockIdx.x / ntile;
    const int64_t brick = (int64_t)row * NB + tile * TB + wave;  // bricks are numbered x fastest
    const int a = A.start[brick];
    int n = A.start[brick + 1] - a;
    if (n >= dense_min) n = 0;  // the item kernel's
    TcRegs R;
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < TC_P; ++k) {
        const int j = lane + 64 * k;
        R.lc[k] = -1;
        if (j < n) {
            tc_load<KIND, PM>(A, a + j, R.lc[k], R.u[0][k], R.u[1][k], R.u[2][k], R.v[0][k], R.v[1][k], R.v[2][k],
                              R.q[k]);
            m[0] = fmaxf(m[0], fabsf(R.q[k]));
            m[1] = fmaxf(m[1], fabsf(R.v[0][k]));
            m[2] = fmaxf(m[2], fabsf(R.v[1][k]));
            m[3] = fmaxf(m[3], fabsf(R.v[2][k]));
        }
    }
    for (int j = 64 * TC_P + lane; j < n; j += 64) {
        int c;
        float u0, u1, u2, v0, v1, v2, q;
        tc_load<KIND, PM>(A, a + j, c, u0, u1, u2, v0, v1, v2, q);
        m[0] = fmaxf(m[0], fabsf(q));
        m[1] = fmaxf(m[1], fabsf(v0));
        m[2] = fmaxf(m[2], fabsf(v1));
        m[3] = fmaxf(m[3], fabsf(v2));
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[k] = fmaxf(m[k], __shfl_xor(m[k], o));
    if (lane == 0) {
        smax[wave][0] = m[0];
        smax[wave][1] = m[1];
        smax[wave][2] = m[2];
        smax[wave][3] = m[3];
        sn[wave] = n;
    }
    __syncthreads();
    float mq = 0.0f, mv[3] = {0.0f, 0.0f, 0.0f};
    int nm = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b) {
        mq = fmaxf(mq, smax[b][0]);
        mv[0] = fmaxf(mv[0], smax[b][1]);
        mv[1] = fmaxf(mv[1], smax[b][2]);
        mv[2] = fmaxf(mv[2], smax[b][3]);
        nm = max(nm, sn[b]);
    }
    float own[C][8];
#define LF2MI_TC_PASS(I) \
    tc_pass<KIND, I, PM, TB>(A, T, shell, win, R, a, n, tc_scale<KIND, I>(nm, mq, mv), own, tile, row, ntile);
    LF2MI_TC_PASS(0)
    if constexpr (C > 1) {
        LF2MI_TC_PASS(1)
        LF2MI_TC_PASS(2)
    }
    if constexpr (C > 3) {
        LF2MI_TC_PASS(3)
        LF2MI_TC_PASS(4)
        LF2MI_TC_PASS(5)
    }
#undef LF2MI_TC_PASS
    // the own nodes, all components of a site together (whole lines of the interleaved field)
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int by = row % NB, bz = row / NB;
    const int64_t i0 = site_index(N, A.h, tile * TX + tx, by * 8 + ty, bz * 8);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        float *f = A.field + (i0 + (int64_t)z * G * G) * C;
        if (set) {
#pragma unroll
            for (int c = 0; c < C; ++c) f[c] = own[c][z];
        } else {
            float o[C];
#pragma unroll
            for (int c = 0; c < C; ++c) o[c] = f[c];
#pragma unroll
            for (int c = 0; c < C; ++c) f[c] = o[c] + own[c][z];
        }
    }
}

// What the fold kernels need of each component's shells and seams.  The fold visits each
// node that receives shell or seam values once, sums every component's contributions and
// adds them to the field without atomics (nothing else writes the field meanwhile).  A node
// receives from the windows of neighbouring bricks: in y, a brick row's y = 8 layer lands on
// the next row's y = 0 layer and its y = -1 layer on the previous row's y = 7 layer (link
// axes), likewise in z; in x, a tile's outer planes land on the neighbouring tiles' edge
// nodes.  Receiving layers (union over the components): sy = y layers in the brick (0: a
// component has an upper y halo, 7: a lower one), cy the other y layers, likewise z, and sx
// the x positions in the tile (0 and TX - 1).
struct TcFold {
    int64_t shell_off[6], seam_off[6];  // as in TcArgs
    int rows[6], ns[6];                 // shell rows per brick row, seam planes per tile
    int lo[6][3], hi[6][3], w[6][3];    // window per axis
    int C, TX, ntile;
    int ny, nz, nx;                     // receiving layers
    int sy[2], sz[2], sx[2], cy[8], cz[8];
};

// the y and z shell rows of component c that land on the row (Y, Z): their offsets in the
// buffer (the value for x = X at offset + X)
__device__ __forceinline__ int tc_row_sources(const TcFold &F, int N, int NB, int c, int Y, int Z, int64_t (&off)[8]) {
    const int by = Y >> 3, ly = Y & 7, bz = Z >> 3, lz = Z & 7;
    const int LY = F.lo[c][1], HY = F.hi[c][1], LZ = F.lo[c][2], HZ = F.hi[c][2];
    const int WZ = F.w[c][2], RR = F.rows[c], RY = (F.w[c][1] - 8) * WZ;
    int n = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int zp = lz - 8 * dz;
        if (zp < LZ || zp > HZ) continue;
        const int bzs = dz < 0 ? (bz == 0 ? NB - 1 : bz - 1) : (dz > 0 ? (bz + 1 == NB ? 0 : bz + 1) : bz);
        for (int dy = -1; dy <= 1; ++dy) {
            const int yp = ly - 8 * dy;
            if (yp < LY || yp > HY || (dy == 0 && dz == 0)) continue;
            const int bys = dy < 0 ? (by == 0 ? NB - 1 : by - 1) : (dy > 0 ? (by + 1 == NB ? 0 : by + 1) : by);
            const int64_t srow = bys + (int64_t)NB * bzs;
            const int r = dy != 0 ? (yp == 8 ? -LY : 0) * WZ + (zp - LZ) : RY + (zp == 8 ? -LZ : 0) * 8 + yp;
            off[n++] = F.shell_off[c] + (srow * RR + r) * N;
        }
    }
    return n;
}

// the seam values of component c that land on node (X, Y, Z): from the previous tile's upper
// plane on the tile's first x, from the next tile's lower plane on its last x
__device__ __forceinline__ float tc_seam_sum(const TcFold &F, const float *__restrict__ shell, int NB, int c, int X,
                                             int Y, int Z) {
    const int tile = X / F.TX, lxt = X - tile * F.TX;
    int st, side;
    if (lxt == 0 && F.hi[c][0] == 8) {
        st = tile == 0 ? F.ntile - 1 : tile - 1;
        side = -F.lo[c][0];
    } else if (lxt == F.TX - 1 && F.lo[c][0] == -1) {
        st = tile + 1 == F.ntile ? 0 : tile + 1;
        side = 0;
    } else {
        return 0.0f;
    }
    const int by = Y >> 3, ly = Y & 7, bz = Z >> 3, lz = Z & 7;
    const int LY = F.lo[c][1], HY = F.hi[c][1], LZ = F.lo[c][2], HZ = F.hi[c][2], WY = F.w[c][1];
    const int PL = WY * F.w[c][2];
    float s = 0.0f;
    for (int dz = -1; dz <= 1; ++dz) {
        const int zp = lz - 8 * dz;
        if (zp < LZ || zp > HZ) continue;
        const int bzs = dz < 0 ? (bz == 0 ? NB - 1 : bz - 1) : (dz > 0 ? (bz + 1 == NB ? 0 : bz + 1) : bz);
        for (int dy = -1; dy <= 1; ++dy) {
            const int yp = ly - 8 * dy;
            if (yp < LY || yp > HY) continue;
            const int bys = dy < 0 ? (by == 0 ? NB - 1 : by - 1) : (dy > 0 ? (by + 1 == NB ? 0 : by + 1) : by);
            const int64_t srow = bys + (int64_t)NB * bzs;
            s += shell[F.seam_off[c] + ((srow * F.ntile + st) * F.ns[c] + side) * PL + (yp - LY) + WY * (zp - LZ)];
        }
    }
    return s;
}

// the receiving nodes of whole rows along x: the sy layers (every z), then the cy layers
// with the sz layers; one workgroup per row.  The row's shell sources are found once and
// summed for every x (coalesced); then the row's tile-edge nodes add their seam values, one
// thread each (after the barrier, which orders the two updates of those nodes).
template <int C>
__global__ __launch_bounds__(256) void fold_rows_kernel(float *__restrict__ field, int N, int h, int NB, TcFold F,
                                                        const float *__restrict__ shell) {
    __shared__ int64_t soff[C][8];
    __shared__ int scnt[C];
    int64_t b = blockIdx.x;
    const int64_t rowsY = (int64_t)F.ny * NB * N;
    int Y, Z;
    if (b < rowsY) {
        const int yi = (int)(b % (F.ny * NB));
        Z = (int)(b / (F.ny * NB));
        Y = (yi / F.ny) * 8 + F.sy[yi % F.ny];
    } else {
        b -= rowsY;
        const int my = (8 - F.ny) * NB;
        const int yi = (int)(b % my), zi = (int)(b / my);
        Y = (yi / (8 - F.ny)) * 8 + F.cy[yi % (8 - F.ny)];
        Z = (zi / F.nz) * 8 + F.sz[zi % F.nz];
    }
    if ((int)threadIdx.x < C) {
        int64_t off[8];
        const int c = threadIdx.x, n = tc_row_sources(F, N, NB, c, Y, Z, off);
        for (int k = 0; k < n; ++k) soff[c][k] = off[k];
        scnt[c] = n;
    }
    __syncthreads();
    float *f0 = field + site_index(N, h, 0, Y, Z) * C;
    for (int X = threadIdx.x; X < N; X += 256) {
        float v[C];
        bool any = false;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float s = 0.0f;
            for (int k = 0; k < scnt[c]; ++k) s += shell[soff[c][k] + X];
            v[c] = s;
            any |= s != 0.0f;
        }
        if (!any) continue;
        float *f = f0 + (int64_t)X * C;
#pragma unroll
        for (int c = 0; c < C; ++c) f[c] += v[c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < F.nx * F.ntile; e += 256) {
        const int X = (e / F.nx) * F.TX + F.sx[e % F.nx];
        float v[C];
        bool any = false;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            v[c] = tc_seam_sum(F, shell, NB, c, X, Y, Z);
            any |= v[c] != 0.0f;
        }
        if (!any) continue;
        float *f = f0 + (int64_t)X * C;
#pragma unroll
        for (int c = 0; c < C; ++c) f[c] += v[c];
    }
}

// the remaining receiving nodes: tile edges (sx) in the cy layers and cz layers
template <int C>
__global__ __launch_bounds__(256) void fold_edges_kernel(float *__restrict__ field, int N, int h, int NB, TcFold F,
                                                         const float *__restrict__ shell, int64_t total) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int nxt = F.nx * F.ntile, my = (8 - F.ny) * NB;
    const int xi = (int)(t % nxt);
    const int64_t r = t / nxt;
    const int yi = (int)(r % my), zi = (int)(r / my);
    const int X = (xi / F.nx) * F.TX + F.sx[xi % F.nx];
    const int Y = (yi / (8 - F.ny)) * 8 + F.cy[yi % (8 - F.ny)];
    const int Z = (zi / (8 - F.nz)) * 8 + F.cz[zi % (8 - F.nz)];
    // no shell row lands on these y and z layers: the seams only
    float v[C];
    bool any = false;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        v[c] = tc_seam_sum(F, shell, NB, c, X, Y, Z);
        any |= v[c] != 0.0f;
    }
    if (!any) return;
    float *f = field + site_index(N, h, X, Y, Z) * C;
#pragma unroll
    for (int c = 0; c < C; ++c) f[c] += v[c];
}

// maxima over the particles of |q| (out[0]) and |v_d| (out[1 + d]) as the bits of
// non-negative floats: the bound of the vector and tensor fixed-point deposits
// (grid-stride, one atomic per workgroup and quantity: the four addresses see few atomics)
static constexpr int DMAX_BLOCKS = 1024;
__global__ __launch_bounds__(256) void dep_max_kernel(int64_t np, Soa32 s, float mass_nv,
                                                      unsigned int *__restrict__ out) {
    __shared__ float wm[4][4];
    float m[4] = {0.f, 0.f, 0.f, 0.f};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
        m[0] = fmaxf(m[0], fabsf(s.mass ? s.mass[i] * mass_nv : mass_nv));
        m[1] = fmaxf(m[1], fabsf(s.vel[0][i]));
        m[2] = fmaxf(m[2], fabsf(s.vel[1][i]));
        m[3] = fmaxf(m[3], fabsf(s.vel[2][i]));
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m[k] = fmaxf(m[k], __shfl_xor(m[k], o));
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int k = 0; k < 4; ++k) wm[wave][k] = m[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        float v = 0.f;
        for (int w = 0; w < 4; ++w) v = fmaxf(v, wm[w][threadIdx.x]);
        atomicMax(&out[threadIdx.x], __float_as_uint(v));
    }
}

struct FxScales {
    double inv[6];
};

// component c of site i at fx[i C + c]: f += q / scale_c
__global__ void fixed_to_field_c_kernel(const unsigned long long *__restrict__ fx, float *__restrict__ f, int64_t n,
                                        int C, FxScales S) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const long long q = (long long)fx[i];
        if (q) f[i] += (float)((double)q * S.inv[i % C]);
    }
}

__global__ void fill_u64_kernel(unsigned long long *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

__global__ void fixed_to_field_kernel(const unsigned long long *__restrict__ fx, float *__restrict__ f, int64_t n,
                                      double scale) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const long long q = (long long)fx[i];
        if (q) f[i] += (float)((double)q * scale);
    }
}

// ---------------------------------------------------------------- gather + kick + drift (F32)
struct GkdArgs {
    Soa32 s;
    const int *start;
    const int *istart;
    const int *item_bucket;
    Geo g;
    const float *phi;
    int h;
    float dt_kick;
    double dt_drift;
    int *next_counts;       // histogram of the new buckets
    unsigned int *vmax2;    // bits of max |v|^2 (non-negative float)
    int chunked;            // bricks below GKD_DIRECT_MAX are left to gkd_chunk_kernel
};

// kick with the gathered acceleration, drift in FP64 cell units, re-anchor and store:
// X = (c + u) + fl(dt v) N, then the sequential periodic wrap of Particles.hpp:2028-2032
// and re-binning to (cell, FP32 offset).  Returns the new cell.
__device__ __forceinline__ int kick_drift_store(const GkdArgs &A, int p, int cx, int cy, int cz, const float u[3],
                                                const float vin[3], const float a[3], float &vmax) {
    const int N = A.g.N;
    float v[3];
    int c3[3] = {cx, cy, cz};
    float o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v[d] = vin[d] + A.dt_kick * a[d];
        double X = add_rn((double)c3[d], (double)u[d]);
        X = add_rn(X, mul_rn(A.dt_drift, (double)v[d]) * (double)N);
        if (X < 0.0) X = add_rn(X, (double)N);
        if (X >= (double)N) X = sub_rn(X, (double)N);
        // beyond the reference's range (|dt v| >= box) wrap fully so the cell stays valid
        if (X < 0.0 || X >= (double)N) X -= (double)N * floor(X / (double)N);
        const double fl = floor(X);
        int cn = (int)fl;
        if (cn >= N) cn = N - 1;
        if (cn < 0) cn = 0;
        float on = (float)(X - fl);
        if (!(on < 1.0f)) on = 0x1.fffffep-1f;
        c3[d] = cn;
        o[d] = on;
    }
    const int newcell = c3[0] + N * (c3[1] + N * c3[2]);
    A.s.cell[p] = newcell;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.s.off[d][p] = o[d];
        A.s.vel[d][p] = v[d];
    }
    vmax = fmaxf(vmax, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    return newcell;
}

// CIC interpolation of F = -grad phi (centred differences, N/2 per unit of phi) at a
// particle with offsets u in the cell whose stencil coordinates per axis are
// xs = {x-1, x, x+1, x+2} (likewise ys, zs): the 32 values around the cell from P(x, y, z)
template <typename PHI>
__device__ __forceinline__ void cic_force(const PHI &P, const int xs[4], const int ys[4], const int zs[4], float hn,
                                          const float u[3], float a[3]) {
    float core[2][2][2], xw[2][2][2], yw[2][2][2], zw[2][2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                core[k][j][i] = P(xs[1 + i], ys[1 + j], zs[1 + k]);
                xw[k][j][i] = P(xs[i ? 3 : 0], ys[1 + j], zs[1 + k]);  // i selects x-1 / x+2
                yw[k][j][i] = P(xs[1 + i], ys[j ? 3 : 0], zs[1 + k]);  // j selects y-1 / y+2
                zw[k][j][i] = P(xs[1 + i], ys[1 + j], zs[k ? 3 : 0]);  // k selects z-1 / z+2
            }
    a[0] = a[1] = a[2] = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const float w = (i ? u[0] : 1.0f - u[0]) * (j ? u[1] : 1.0f - u[1]) * (k ? u[2] : 1.0f - u[2]);
                // neighbours of node (i,j,k) along each axis, from the core or the wings
                const float xp = i ? xw[k][j][1] : core[k][j][1], xm = i ? core[k][j][0] : xw[k][j][0];
                const float yp = j ? yw[k][1][i] : core[k][1][i], ym = j ? core[k][0][i] : yw[k][0][i];
                const float zp = k ? zw[1][j][i] : core[1][j][i], zm = k ? core[0][j][i] : zw[0][j][i];
                a[0] += w * (-(xp - xm) * hn);
                a[1] += w * (-(yp - ym) * hn);
                a[2] += w * (-(zp - zm) * hn);
            }
}

// the stencil straight from phi in HBM (periodic wrap)
__device__ __forceinline__ void gather_direct(const GkdArgs &A, int cx, int cy, int cz, const float u[3], float a[3]) {
    const int N = A.g.N, h = A.h;
    const int xs[4] = {wrapi(cx - 1, N), cx, wrapi(cx + 1, N), wrapi(cx + 2, N)};
    const int ys[4] = {wrapi(cy - 1, N), cy, wrapi(cy + 1, N), wrapi(cy + 2, N)};
    const int zs[4] = {wrapi(cz - 1, N), cz, wrapi(cz + 1, N), wrapi(cz + 2, N)};
    const float *phi = A.phi;
    cic_force([=](int x, int y, int z) { return phi[site_index(N, h, x, y, z)]; }, xs, ys, zs, 0.5f * (float)N, u, a);
}

// bricks with fewer particles go to the chunks of gkd_chunk_kernel (whose window serves
// several bricks); items below it that belong to fuller bricks (a brick's last item)
// gather phi straight from HBM instead of staging a window and force table for a few
// particles
static constexpr int GKD_DIRECT_MAX = 192;
static constexpr int GKD_MERGE = 4;

// histogram of the new buckets of one item's particles (bucket b): particles that stay in
// the brick are counted with one LDS atomic per wave, moves to the 26 neighbouring bricks
// in a 27-entry LDS table, farther moves straight in HBM; the table is flushed once per
// item (gkd_hist_flush), so a dense brick's many items do not serialise on its counter
__device__ __forceinline__ void gkd_hist_add(const GkdArgs &A, int b, int *h27, bool act, int newcell) {
    const int nb = act ? bucket_of(newcell, A.g) : b;
    const bool stay = act && nb == b;
    const unsigned long long st = __ballot(stay);
    if (lane_id() == 0 && st) atomicAdd(&h27[13], __popcll(st));
    if (!act || stay) return;
    const int NB = A.g.NB;
    const int bx = nb % NB, by = (nb / NB) % NB, bz = nb / (NB * NB);
    const int rx = b % NB, ry = (b / NB) % NB, rz = b / (NB * NB);
    auto delta = [NB](int d) {
        if (d == NB - 1 && d > 1) return -1;
        if (d == 1 - NB && d < -1) return 1;
        return d;
    };
    const int dx = delta(bx - rx), dy = delta(by - ry), dz = delta(bz - rz);
    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1)
        atomicAdd(&h27[13 + dx + 3 * dy + 9 * dz], 1);
    else
        atomicAdd(&A.next_counts[nb], 1);
}

// after a barrier: one HBM atomic per non-empty neighbour entry
__device__ __forceinline__ void gkd_hist_flush(const GkdArgs &A, int b, const int *h27) {
    const int t = threadIdx.x;
    if (t >= 27 || h27[t] == 0) return;
    const int NB = A.g.NB;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int bx = (b % NB + dx + NB) % NB, by = ((b / NB) % NB + dy + NB) % NB, bz = (b / (NB * NB) + dz + NB) % NB;
    atomicAdd(&A.next_counts[bx + NB * (by + NB * bz)], h27[t]);
}

template <int BC>
__global__ __launch_bounds__(DEP_WG) void gkd_f32_kernel(GkdArgs A) {
    constexpr int PW = MAXB + 3;  // phi window [o-1, o+B+1]
    constexpr int FW = MAXB + 1;  // force nodes [o, o+B]
    __shared__ float phw[PW * PW * PW];
    __shared__ float frc[3][FW * FW * FW];
    __shared__ unsigned int red;
    __shared__ int h27[27];
    ItemRange r;
    if (!item_range(blockIdx.x, A.start, A.istart, A.item_bucket, A.g, r)) return;
    if (A.chunked && A.start[r.b + 1] - A.start[r.b] < GKD_DIRECT_MAX) return;
    if (A.phi) {
        // one workgroup per GKD_MERGE consecutive items of a brick: one phi window and force
        // table serve them all (the group's other items exit)
        const int bs = A.start[r.b], be = A.start[r.b + 1];
        if (((r.p0 - bs) / ITEM_P) % GKD_MERGE != 0) return;
        r.p1 = min(be, r.p0 + GKD_MERGE * ITEM_P);
    }
    const int N = A.g.N, B = BC ? BC : A.g.B;
    if (threadIdx.x == 0) red = 0u;
    if (threadIdx.x < 27) h27[threadIdx.x] = 0;
    __syncthreads();
    float vmax = 0.0f;
    if (!A.phi || r.p1 - r.p0 < GKD_DIRECT_MAX) {
        // sparse item (no window) or drift only (no potential: a = 0)
        for (int base = r.p0; base < r.p1; base += DEP_WG) {
            const int p = base + threadIdx.x;
            const bool act = p < r.p1;
            int newcell = 0;
            if (act) {
                int cx, cy, cz;
                decode_cell(A.s.cell[p], A.g, cx, cy, cz);
                const float u[3] = {A.s.off[0][p], A.s.off[1][p], A.s.off[2][p]};
                const float vin[3] = {A.s.vel[0][p], A.s.vel[1][p], A.s.vel[2][p]};
                float a[3] = {0.f, 0.f, 0.f};
                if (A.phi) gather_direct(A, cx, cy, cz, u, a);
                newcell = kick_drift_store(A, p, cx, cy, cz, u, vin, a, vmax);
            }
            gkd_hist_add(A, r.b, h27, act, newcell);
        }
    } else {
        const int pw = B + 3, fw = B + 1;
        // first batch of particle data in flight while the window loads
        int p = r.p0 + threadIdx.x;
        bool act = p < r.p1;
        int cell = 0;
        float u[3] = {0.f, 0.f, 0.f}, vin[3] = {0.f, 0.f, 0.f};
        if (act) {
            cell = A.s.cell[p];
            u[0] = A.s.off[0][p]; u[1] = A.s.off[1][p]; u[2] = A.s.off[2][p];
            vin[0] = A.s.vel[0][p]; vin[1] = A.s.vel[1][p]; vin[2] = A.s.vel[2][p];
        }
        {
            // every window load issued before the LDS stores
            constexpr int WK = ((MAXB + 3) * (MAXB + 3) * (MAXB + 3) + DEP_WG - 1) / DEP_WG;
            const int nwin = pw * pw * pw;
            float wv[WK];
#pragma unroll
            for (int k = 0; k < WK; ++k) {
                const int i = threadIdx.x + k * DEP_WG;
                wv[k] = 0.0f;
                if (i < nwin) {
                    const int gx = wrapi(r.ox - 1 + i % pw, N);
                    const int gy = wrapi(r.oy - 1 + (i / pw) % pw, N);
                    const int gz = wrapi(r.oz - 1 + i / (pw * pw), N);
                    wv[k] = A.phi[site_index(N, A.h, gx, gy, gz)];
                }
            }
#pragma unroll
            for (int k = 0; k < WK; ++k) {
                const int i = threadIdx.x + k * DEP_WG;
                if (i < nwin) phw[i] = wv[k];
            }
        }
        __syncthreads();
        const float hn = 0.5f * (float)N;
        for (int i = threadIdx.x; i < fw * fw * fw; i += DEP_WG) {
            const int x = i % fw + 1, y = (i / fw) % fw + 1, z = i / (fw * fw) + 1;
            const int c = x + pw * (y + pw * z);
            frc[0][i] = -(phw[c + 1] - phw[c - 1]) * hn;
            frc[1][i] = -(phw[c + pw] - phw[c - pw]) * hn;
            frc[2][i] = -(phw[c + pw * pw] - phw[c - pw * pw]) * hn;
        }
        __syncthreads();
        for (int base = r.p0; base < r.p1; base += DEP_WG) {
            if (base != r.p0) {
                p = base + threadIdx.x;
                act = p < r.p1;
                if (act) {
                    cell = A.s.cell[p];
                    u[0] = A.s.off[0][p]; u[1] = A.s.off[1][p]; u[2] = A.s.off[2][p];
                    vin[0] = A.s.vel[0][p]; vin[1] = A.s.vel[1][p]; vin[2] = A.s.vel[2][p];
                }
            }
            int newcell = 0;
            if (act) {
                int cx, cy, cz;
                decode_cell(cell, A.g, cx, cy, cz);
                const int lx = cx - r.ox, ly = cy - r.oy, lz = cz - r.oz;
                float a[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 2; ++k)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            const float w = (i ? u[0] : 1.0f - u[0]) * (j ? u[1] : 1.0f - u[1]) * (k ? u[2] : 1.0f - u[2]);
                            const int n = (lx + i) + fw * ((ly + j) + fw * (lz + k));
                            a[0] += w * frc[0][n];
                            a[1] += w * frc[1][n];
                            a[2] += w * frc[2][n];
                        }
                newcell = kick_drift_store(A, p, cx, cy, cz, u, vin, a, vmax);
            }
            gkd_hist_add(A, r.b, h27, act, newcell);
        }
    }
    if (A.vmax2) atomicMax(&red, __float_as_uint(vmax));
    __syncthreads();
    gkd_hist_flush(A, r.b, h27);
    if (A.vmax2 && threadIdx.x == 0) atomicMax(A.vmax2, red);
}

// ---- sparse bricks: chunks along x
// Runs of consecutive non-empty bricks of one brick row (same y, z) whose counts are below
// GKD_DIRECT_MAX form chunks of at most GKD_CHUNK_B bricks and ITEM_P particles (their
// particles are contiguous in the bucket order).  A chunk stages one phi window of
// (8 nbx + 3) x 11 x 11 nodes in LDS, so the phi lines its particles need are fetched once
// per chunk rather than once per particle stencil (the L1 gathers of the per-particle
// path are bound by the number of lines they touch).
static constexpr int GKD_CHUNK_B = 8;
// window rows a wave keeps in flight: with 4 waves x 32 >= (MAXB + 3)^2 rows the whole window
// is requested at once (a chunk is a chain of dependent latency phases, so each batch of
// rows costs a full HBM round trip: 4 batches of 8 rows took 0.70 ms, one batch 0.63 ms)
static constexpr int GKD_STAGE = 32;
static_assert((DEP_WG / 64) * GKD_STAGE >= (MAXB + 3) * (MAXB + 3), "one batch of window rows");

// walk one brick row, calling emit(p0, p1, bx0, nbx) per chunk
template <typename F>
__device__ __forceinline__ void gkd_walk_row(const int *__restrict__ start, int64_t b_row, int NB, F emit) {
    int bx0 = -1, p0 = 0, p1 = 0;
    for (int bx = 0; bx < NB; ++bx) {
        const int s0 = start[b_row + bx], s1 = start[b_row + bx + 1];
        const int c = s1 - s0;
        const bool take = c > 0 && c < GKD_DIRECT_MAX;
        if (bx0 >= 0 && (!take || bx - bx0 >= GKD_CHUNK_B || p1 - p0 + c > ITEM_P)) {
            emit(p0, p1, bx0, bx - bx0);
            bx0 = -1;
        }
        if (!take) continue;
        if (bx0 < 0) {
            bx0 = bx;
            p0 = s0;
        }
        p1 = s1;
    }
    if (bx0 >= 0) emit(p0, p1, bx0, NB - bx0);
}

// one thread per brick row: count the row's chunks, reserve room, write them
// (int4: particle range, first bucket, number of bricks)
__global__ void gkd_chunks_kernel(const int *__restrict__ start, Geo g, int4 *__restrict__ chunks,
                                  int *__restrict__ nchunk) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (int64_t)g.NB * g.NB) return;
    const int64_t b_row = row * g.NB;
    int n = 0;
    gkd_walk_row(start, b_row, g.NB, [&](int, int, int, int) { ++n; });
    if (n == 0) return;
    int k = atomicAdd(nchunk, n);
    gkd_walk_row(start, b_row, g.NB, [&](int p0, int p1, int bx0, int nbx) {
        chunks[k++] = make_int4(p0, p1, (int)(b_row + bx0), nbx);
    });
}

// persistent over the chunk list; every workgroup leaves once the list is done
template <int BC>
__global__ __launch_bounds__(DEP_WG) void gkd_chunk_kernel(GkdArgs A, const int4 *__restrict__ chunks,
                                                           int *__restrict__ nchunk) {
    constexpr int WXM = GKD_CHUNK_B * MAXB + 3, WYM = MAXB + 3;
    constexpr int NHM = (GKD_CHUNK_B + 2) * 9;
    __shared__ float win[WXM * WYM * WYM];
    __shared__ int hist[NHM];
    __shared__ unsigned int red;
    __shared__ int claim;
    const int N = A.g.N, B = BC ? BC : A.g.B, NB = A.g.NB;
    const int total = nchunk[0];
    const float hn = 0.5f * (float)N;
    float vmax = 0.0f;
    if (threadIdx.x == 0) red = 0u;
    // chunks are claimed one at a time from nchunk[1] (they hold 1 to GKD_CHUNK_B bricks)
    // the following chunk is claimed while this one's particles are processed
    if (threadIdx.x == 0) claim = atomicAdd(&nchunk[1], 1);
    __syncthreads();
    int ci = claim;
    for (;;) {
        if (ci >= total) break;
        const int4 ch = chunks[ci];
        const int b0 = ch.z, nbx = ch.w;
        const int bx0 = b0 % NB, by = (b0 / NB) % NB, bz = b0 / (NB * NB);
        const int ox = bx0 * B, oy = by * B, oz = bz * B;
        const int wx = nbx * B + 3, wy = B + 3, nh = (nbx + 2) * 9;
        for (int i = threadIdx.x; i < nh; i += DEP_WG) hist[i] = 0;
        // every particle of the chunk (at most ITEM_P) in flight while the window loads
        constexpr int CP = ITEM_P / DEP_WG;
        int pcell[CP];
        float pu[CP][3], pv[CP][3];
#pragma unroll
        for (int k = 0; k < CP; ++k) {
            const int q = ch.x + threadIdx.x + k * DEP_WG;
            pcell[k] = 0;
            if (q < ch.y) {
                pcell[k] = A.s.cell[q];
                pu[k][0] = A.s.off[0][q]; pu[k][1] = A.s.off[1][q]; pu[k][2] = A.s.off[2][q];
                pv[k][0] = A.s.vel[0][q]; pv[k][1] = A.s.vel[1][q]; pv[k][2] = A.s.vel[2][q];
            }
        }
        // rows (y, z) of the window, one wave per row: the row's base index is wave-uniform,
        // a lane holds x = lane and x = lane + 64; every row's loads are issued before the
        // LDS stores (a load-store loop would wait out one HBM round trip per row)
        {
            const int nrows = wy * wy;
            const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int lane = threadIdx.x & 63;
            const int64_t G = (int64_t)N + 2 * A.h;
            const bool has0 = lane < wx, has1 = lane + 64 < wx;
            const int gx0 = wrapi(ox - 1 + lane, N) + A.h;
            const int gx1 = has1 ? wrapi(ox - 1 + 64 + lane, N) + A.h : 0;
            constexpr int NWV = DEP_WG / 64;
            for (int r0 = wave; r0 < nrows; r0 += NWV * GKD_STAGE) {
                float v0[GKD_STAGE], v1[GKD_STAGE];
#pragma unroll
                for (int k = 0; k < GKD_STAGE; ++k) {
                    const int row = r0 + NWV * k;
                    v0[k] = 0.0f;
                    v1[k] = 0.0f;
                    if (row < nrows) {
                        const int y = row % wy, z = row / wy;
                        const int64_t rb = G * ((int64_t)(wrapi(oy - 1 + y, N) + A.h) + G * (int64_t)(wrapi(oz - 1 + z, N) + A.h));
                        if (has0) v0[k] = A.phi[rb + gx0];
                        if (has1) v1[k] = A.phi[rb + gx1];
                    }
                }
#pragma unroll
                for (int k = 0; k < GKD_STAGE; ++k) {
                    const int row = r0 + NWV * k;
                    if (row < nrows) {
                        if (has0) win[lane + wx * row] = v0[k];
                        if (has1) win[lane + 64 + wx * row] = v1[k];
                    }
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) claim = atomicAdd(&nchunk[1], 1);  // every lane has read ci
#pragma unroll
        for (int k = 0; k < CP; ++k) {
            const int p = ch.x + threadIdx.x + k * DEP_WG;
            if (p >= ch.y) break;
            const int cell = pcell[k];
            float u[3] = {pu[k][0], pu[k][1], pu[k][2]}, vin[3] = {pv[k][0], pv[k][1], pv[k][2]};
            int cx, cy, cz;
            decode_cell(cell, A.g, cx, cy, cz);
            // window coordinates: cell c sits at c - o + 1
            const int lx = cx - ox + 1, ly = cy - oy + 1, lz = cz - oz + 1;
            const int xs[4] = {lx - 1, lx, lx + 1, lx + 2};
            const int ys[4] = {ly - 1, ly, ly + 1, ly + 2};
            const int zs[4] = {lz - 1, lz, lz + 1, lz + 2};
            float a[3];
            cic_force([&](int x, int y, int z) { return win[x + wx * (y + wy * z)]; }, xs, ys, zs, hn, u, a);
            const int newcell = kick_drift_store(A, p, cx, cy, cz, u, vin, a, vmax);
            // new bucket relative to the chunk: x in [-1, nbx], y and z in [-1, 1] (periodic)
            int nx, ny, nz;
            decode_cell(newcell, A.g, nx, ny, nz);
            nx /= B;
            ny /= B;
            nz /= B;
            int dx = nx - bx0, dy = ny - by, dz = nz - bz;
            dx += dx < 0 ? NB : 0;
            dy += dy < 0 ? NB : 0;
            dz += dz < 0 ? NB : 0;
            if (dx > nbx) dx -= NB;
            if (dy > 1) dy -= NB;
            if (dz > 1) dz -= NB;
            if (dx >= -1 && dy >= -1 && dz >= -1 && dy <= 1 && dz <= 1)
                atomicAdd(&hist[(dx + 1) + (nbx + 2) * ((dy + 1) + 3 * (dz + 1))], 1);
            else
                atomicAdd(&A.next_counts[nx + NB * (ny + NB * nz)], 1);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nh; i += DEP_WG) {
            const int c = hist[i];
            if (c == 0) continue;
            const int ex = i % (nbx + 2) - 1, ey = (i / (nbx + 2)) % 3 - 1, ez = i / (3 * (nbx + 2)) - 1;
            const int bx = (bx0 + ex + NB) % NB, bby = (by + ey + NB) % NB, bbz = (bz + ez + NB) % NB;
            atomicAdd(&A.next_counts[bx + NB * (bby + NB * bbz)], c);
        }
        __syncthreads();
        ci = claim;
    }
    if (A.vmax2) {
        atomicMax(&red, __float_as_uint(vmax));
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(A.vmax2, red);
    }
}

// ---------------------------------------------------------------- F64 plain path
struct Dep64Args {
    Soa64 s;
    int64_t np;
    int N;
    double mass_nv;
    double *field;
    int h, tgt;
};

// NX, NY, NZ: nodes per axis (1 on a nearest-grid-point axis, whose weight is wx[0] = 1)
template <int NX = 2, int NY = 2, int NZ = 2>
__device__ __forceinline__ void add8_64(double *f, int C, int comp, int N, int h, int tgt, int cx, int cy, int cz,
                                        int lox, int loy, int loz, const double wx[2], const double wy[2],
                                        const double wz[2], double q) {
#pragma unroll
    for (int k = 0; k < NZ; ++k)
#pragma unroll
        for (int j = 0; j < NY; ++j)
#pragma unroll
            for (int i = 0; i < NX; ++i) {
                int x = cx + lox + i, y = cy + loy + j, z = cz + loz + k;
                if (tgt == LF2MI_DEP_PERIODIC) {
                    x = wrapi(x, N);
                    y = wrapi(y, N);
                    z = wrapi(z, N);
                }
                atomicAdd(&f[site_index(N, h, x, y, z) * C + comp], q * wx[i] * wy[j] * wz[k]);
            }
}

template <int KIND>
__global__ void deposit_f64_kernel(Dep64Args A) {
    constexpr int C = KindC<KIND>::C;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const int N = A.N;
    int c[3];
    double vw[3][2], lw[3][2];
    int llo[3];
    for (int d = 0; d < 3; ++d) {
        const double t = A.s.pos[d][p] * (double)N;
        c[d] = (int)floor(t) % N;
        const double u = t - floor(t);
        vw[d][0] = 1.0 - u;
        vw[d][1] = u;
        const double s = u - 0.5;
        if (s < 0) {
            llo[d] = -1;
            lw[d][0] = -s;
            lw[d][1] = 1.0 - lw[d][0];
        } else {
            llo[d] = 0;
            lw[d][1] = s;
            lw[d][0] = 1.0 - lw[d][1];
        }
    }
    const double m = A.s.mass ? A.s.mass[p] * A.mass_nv : A.mass_nv;
    if constexpr (KIND == 0) {
        add8_64(A.field, C, 0, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m);
    } else {
        const double v[3] = {A.s.vel[0][p], A.s.vel[1][p], A.s.vel[2][p]};
        if constexpr (KIND == 1) {
            add8_64(A.field, C, 0, N, A.h, A.tgt, c[0], c[1], c[2], llo[0], 0, 0, lw[0], vw[1], vw[2], m * v[0]);
            add8_64(A.field, C, 1, N, A.h, A.tgt, c[0], c[1], c[2], 0, llo[1], 0, vw[0], lw[1], vw[2], m * v[1]);
            add8_64(A.field, C, 2, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, llo[2], vw[0], vw[1], lw[2], m * v[2]);
        } else if constexpr (KIND == 3) {
            const double one[2] = {1.0, 0.0};
            add8_64<1, 2, 2>(A.field, C, 0, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, one, vw[1], vw[2], m * v[0]);
            add8_64<2, 1, 2>(A.field, C, 1, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], one, vw[2], m * v[1]);
            add8_64<2, 2, 1>(A.field, C, 2, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], one, m * v[2]);
        } else if constexpr (KIND == 4) {
            const double one[2] = {1.0, 0.0};
            add8_64(A.field, C, 0, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[0] * v[0]);
            add8_64(A.field, C, 3, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[1] * v[1]);
            add8_64(A.field, C, 5, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[2] * v[2]);
            add8_64<1, 1, 2>(A.field, C, 1, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, one, one, vw[2], m * v[0] * v[1]);
            add8_64<1, 2, 1>(A.field, C, 2, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, one, vw[1], one, m * v[0] * v[2]);
            add8_64<2, 1, 1>(A.field, C, 4, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], one, one, m * v[1] * v[2]);
        } else {
            add8_64(A.field, C, 0, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[0] * v[0]);
            add8_64(A.field, C, 3, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[1] * v[1]);
            add8_64(A.field, C, 5, N, A.h, A.tgt, c[0], c[1], c[2], 0, 0, 0, vw[0], vw[1], vw[2], m * v[2] * v[2]);
            add8_64(A.field, C, 1, N, A.h, A.tgt, c[0], c[1], c[2], llo[0], llo[1], 0, lw[0], lw[1], vw[2], m * v[0] * v[1]);
            add8_64(A.field, C, 2, N, A.h, A.tgt, c[0], c[1], c[2], llo[0], 0, llo[2], lw[0], vw[1], lw[2], m * v[0] * v[2]);
            add8_64(A.field, C, 4, N, A.h, A.tgt, c[0], c[1], c[2], 0, llo[1], llo[2], vw[0], lw[1], lw[2], m * v[1] * v[2]);
        }
    }
}

struct Gkd64Args {
    Soa64 s;
    int64_t np;
    int N, h;
    const double *phi;
    double dt_kick, dt_drift;
    unsigned long long *vmax2;  // bits of max |v|^2 (non-negative double)
};

__global__ void gkd_f64_kernel(Gkd64Args A) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const int N = A.N;
    int c[3];
    double u[3];
    for (int d = 0; d < 3; ++d) {
        const double t = A.s.pos[d][p] * (double)N;
        c[d] = (int)floor(t) % N;
        u[d] = t - (double)(int)floor(t);
    }
    const double hn = 0.5 * N;
    double a[3] = {0.0, 0.0, 0.0};
    if (A.phi)
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const double w = (i ? u[0] : 1.0 - u[0]) * (j ? u[1] : 1.0 - u[1]) * (k ? u[2] : 1.0 - u[2]);
                const int x = c[0] + i, y = c[1] + j, z = c[2] + k;
                auto P = [&](int xx, int yy, int zz) {
                    return A.phi[site_index(N, A.h, wrapi(xx, N), wrapi(yy, N), wrapi(zz, N))];
                };
                a[0] += w * (-(P(x + 1, y, z) - P(x - 1, y, z)) * hn);
                a[1] += w * (-(P(x, y + 1, z) - P(x, y - 1, z)) * hn);
                a[2] += w * (-(P(x, y, z + 1) - P(x, y, z - 1)) * hn);
            }
    double v2 = 0.0;
    int cn[3];
    for (int d = 0; d < 3; ++d) {
        const double v = add_rn(A.s.vel[d][p], mul_rn(A.dt_kick, a[d]));
        A.s.vel[d][p] = v;
        v2 += v * v;
        double x = add_rn(A.s.pos[d][p], mul_rn(A.dt_drift, v));
        if (x < 0.0) x = add_rn(x, 1.0);
        if (x >= 1.0) x = sub_rn(x, 1.0);
        // beyond the reference's range (|dt v| >= box) wrap fully so the cell stays valid
        if (x < 0.0 || x >= 1.0) x -= floor(x);
        A.s.pos[d][p] = x;
        cn[d] = (int)floor(x * (double)N) % N;
    }
    A.s.cell[p] = cn[0] + N * (cn[1] + N * cn[2]);
    if (A.vmax2) atomicMax(A.vmax2, (unsigned long long)__double_as_longlong(v2));
}

// ---------------------------------------------------------------- download / counts
__global__ void download_f32_kernel(int64_t np, Soa32 s, int N, double *__restrict__ pos, double *__restrict__ vel,
                                    double *__restrict__ mass, double mass_global) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int cell = s.cell[i];
    const int c[3] = {cell % N, (cell / N) % N, cell / (N * N)};
    for (int d = 0; d < 3; ++d) {
        if (pos) pos[3 * i + d] = add_rn((double)c[d], (double)s.off[d][i]) / (double)N;
        if (vel) vel[3 * i + d] = (double)s.vel[d][i];
    }
    if (mass) mass[i] = s.mass ? (double)s.mass[i] : mass_global;
}

__global__ void download_f64_kernel(int64_t np, Soa64 s, double *__restrict__ pos, double *__restrict__ vel) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    for (int d = 0; d < 3; ++d) {
        if (pos) pos[3 * i + d] = s.pos[d][i];
        if (vel) vel[3 * i + d] = s.vel[d][i];
    }
}

// particles whose cell lies in the box [lo, hi) per axis, compacted through a cursor
struct BoxArgs {
    int lo[3], hi[3];
    int N;
    int64_t np, max_out;
    int *cursor;
    double *pos, *vel, *mass;
    double mass_global;
};
__global__ void download_box_f32_kernel(BoxArgs A, Soa32 s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    int c[3] = {0, 0, 0};
    if (i < A.np) {
        const int cell = s.cell[i];
        if (cell >= 0) {
            c[0] = cell % A.N;
            c[1] = (cell / A.N) % A.N;
            c[2] = cell / (A.N * A.N);
            in = true;
            for (int d = 0; d < 3; ++d) in = in && c[d] >= A.lo[d] && c[d] < A.hi[d];
        }
    }
    const int j = agg_inc(A.cursor, 0, in);
    if (!in || j >= A.max_out) return;
    for (int d = 0; d < 3; ++d) {
        A.pos[3 * (int64_t)j + d] = add_rn((double)c[d], (double)s.off[d][i]) / (double)A.N;
        A.vel[3 * (int64_t)j + d] = (double)s.vel[d][i];
    }
    A.mass[j] = s.mass ? (double)s.mass[i] : A.mass_global;
}
__global__ void download_box_f64_kernel(BoxArgs A, Soa64 s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (i < A.np) {
        const int cell = s.cell[i];
        const int c[3] = {cell % A.N, (cell / A.N) % A.N, cell / (A.N * A.N)};
        in = true;
        for (int d = 0; d < 3; ++d) in = in && c[d] >= A.lo[d] && c[d] < A.hi[d];
    }
    const int j = agg_inc(A.cursor, 0, in);
    if (!in || j >= A.max_out) return;
    for (int d = 0; d < 3; ++d) {
        A.pos[3 * (int64_t)j + d] = s.pos[d][i];
        A.vel[3 * (int64_t)j + d] = s.vel[d][i];
    }
    A.mass[j] = s.mass ? s.mass[i] : A.mass_global;
}

__global__ void cell_count_kernel(int64_t np, const int32_t *__restrict__ cell, int32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np && cell[i] >= 0) atomicAdd(&counts[cell[i]], 1);
}

// keys of FP64 positions (AoS x,y,z) without a container; counts of the valid keys
__global__ void bin_keys_kernel(int64_t np, const double *__restrict__ pos, int N, int32_t *__restrict__ keys,
                                int32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    int cell;
    float off[3];
    bin_one(p, N, cell, off);
    if (keys) keys[i] = cell;
    if (counts && cell >= 0) atomicAdd(&counts[cell], 1);
}

__global__ void fill_int_kernel(int *__restrict__ p, int64_t n, int v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- particle statistics
// The output[] reductions of updateVel / moveParticles (particles/LATfield2_Particles.hpp:
// 788-807, 846-885) in FP64.  Each workgroup folds a grid-stride share of the particles and
// then its lanes by a fixed tree; one thread per quantity folds the workgroups' partials in
// index order.  The summation order depends only on np, so results repeat exactly.
static constexpr int RED_Q = 8;        // quantities per call
static constexpr int RED_WG = 256;
static constexpr int RED_BLOCKS = 1024;
enum { RED_NONE = -1, RED_OP_SUM = 0, RED_OP_MIN = 1, RED_OP_MAX = 2 };

struct RedArgs {
    int64_t np;
    const float *vel32[3];
    const float *mass32;   // per-particle masses (F32 container) or null
    const double *vel64[3];
    const double *mass64;  // per-particle masses (F64 container) or null
    double mass;           // the global mass, when no per-particle one is used
    int nq;
    int q[RED_Q];
    int op[RED_Q];
};

__device__ __forceinline__ double red_fold(int op, double a, double b) {
    return op == RED_OP_SUM ? a + b : (op == RED_OP_MIN ? fmin(a, b) : fmax(a, b));
}
__device__ __forceinline__ double red_identity(int op) {
    return op == RED_OP_SUM ? 0.0 : (op == RED_OP_MIN ? __builtin_inf() : -__builtin_inf());
}

template <bool F64>
__global__ __launch_bounds__(RED_WG) void parts_reduce_kernel(RedArgs A, double *__restrict__ partial) {
    double acc[RED_Q];
#pragma unroll
    for (int k = 0; k < RED_Q; ++k) acc[k] = red_identity(A.op[k]);
    const int64_t stride = (int64_t)gridDim.x * RED_WG;
    for (int64_t i = (int64_t)blockIdx.x * RED_WG + threadIdx.x; i < A.np; i += stride) {
        double v[3], m;
        if (F64) {
            for (int d = 0; d < 3; ++d) v[d] = A.vel64[d][i];
            m = A.mass64 ? A.mass64[i] : A.mass;
        } else {
            for (int d = 0; d < 3; ++d) v[d] = (double)A.vel32[d][i];
            m = A.mass32 ? (double)A.mass32[i] : A.mass;
        }
        const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
#pragma unroll
        for (int k = 0; k < RED_Q; ++k) {
            if (k >= A.nq || A.op[k] == RED_NONE) continue;
            double x;
            switch (A.q[k]) {
                case LF2MI_PQ_V2: x = v2; break;
                case LF2MI_PQ_MVX: x = m * v[0]; break;
                case LF2MI_PQ_MVY: x = m * v[1]; break;
                case LF2MI_PQ_MVZ: x = m * v[2]; break;
                case LF2MI_PQ_MASS: x = m; break;
                default: x = 0.5 * m * v2; break;
            }
            acc[k] = red_fold(A.op[k], acc[k], x);
        }
    }
    __shared__ double lds[RED_WG];
    for (int k = 0; k < A.nq; ++k) {
        lds[threadIdx.x] = acc[k];
        __syncthreads();
        for (int h = RED_WG / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) lds[threadIdx.x] = red_fold(A.op[k], lds[threadIdx.x], lds[threadIdx.x + h]);
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * RED_Q + k] = lds[0];
        __syncthreads();
    }
}

// out[k] = the reference's starting value folded with every workgroup's partial in order
__global__ void parts_reduce_final_kernel(RedArgs A, const double *__restrict__ partial, int nblocks,
                                          double *__restrict__ out) {
    const int k = threadIdx.x;
    if (k >= A.nq || A.op[k] == RED_NONE) return;
    const int op = A.op[k];
    double r = op == RED_OP_SUM ? 0.0 : (op == RED_OP_MIN ? 9223372036854775807.0 : -9223372036854775807.0);
    for (int b = 0; b < nblocks; ++b) r = red_fold(op, r, partial[(int64_t)b * RED_Q + k]);
    out[k] = r;
}

}  // namespace lf2mi

// ---------------------------------------------------------------- container
struct lf2mi_parts_s {
    int N, dtype, flags, device;
    lf2mi::Geo g;
    int64_t cap, np;
    double mass_global;
    double total_mass;   // sum of masses (PM mean density with box volume 1)
    lf2mi::Soa32 b32[2];
    int cur;
    lf2mi::Soa64 b64;
    int *bstart, *bcount, *istart, *bsum, *item_bucket;
    int4 *chunks;  // gather chunks of sparse bricks (at most one per brick)
    int *nchunk;   // [number of chunks, claim counter]
    int2 *dgroups; // work list of the dense deposit kernel (at most cap / group + bricks)
    int *ndgroup;  // [number of entries, unused]
    int *ditems;   // work list of the item kernel after deposit_tiles_c_kernel (items of dense bricks)
    int *nditem;   // [number of entries, unused]
    unsigned long long *fx;
    int64_t fx_n;
    void *shell;          // shell and seam buffers of the tiles deposit kernel (grown on demand)
    int64_t shell_bytes;
    unsigned int *vmax_f32;
    unsigned long long *vmax_f64;
    double *red;   // statistics: RED_BLOCKS x RED_Q partials, then RED_Q results
    std::vector<void *> allocs;
    int64_t bytes;
    int plw;       // payload words (8 bytes) per particle, 0 without a payload
    int mass_src;  // deposits use: 0 the global mass, 1 the per-particle masses
    int64_t vec_off;  // vector deposits use: -1 the velocity, else the Real[3] at this payload offset
    int vec_real;     // that vector's Real type
    int64_t aos_vel;  // offset of vel in the uploaded structs (-1: none uploaded)
    void *vsrc;       // 3 planes of cap values (container precision) for the gathered vector
};

namespace lf2mi {

static int dalloc(lf2mi_parts_s *p, void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) return LF2MI_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(particles)", __FILE__, __LINE__);
    p->allocs.push_back(*ptr);
    p->bytes += (int64_t)bytes;
    return LF2MI_OK;
}

static unsigned grid1(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

static int fill_int(int *ptr, int64_t n, int v, hipStream_t st) {
    if (n <= 0) return LF2MI_OK;
    LF2MI_PROF("fill", st);
    fill_int_kernel<<<(unsigned)std::min<int64_t>(grid1(n), 65536), 256, 0, st>>>(ptr, n, v);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

// exclusive scans of bcount -> bstart, items -> istart
static int scan_buckets(lf2mi_parts_s *p, hipStream_t st) {
    const int64_t n = p->g.nbk;
    const int nblocks = (int)((n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    if (nblocks > 4096) return fail(LF2MI_ERANGE, "too many buckets (%lld)", (long long)n);
    LF2MI_PROF("rebucket_scan", st);
    scan_local_kernel<<<nblocks, SCAN_T, 0, st>>>(p->bcount, n, p->bstart, p->istart, p->bsum);
    LF2MI_LAUNCHED();
    scan_sums_kernel<<<1, 1024, 0, st>>>(p->bsum, nblocks, p->bstart, p->istart, n);
    LF2MI_LAUNCHED();
    scan_add_kernel<<<nblocks, SCAN_T, 0, st>>>(p->bstart, p->istart, p->bsum, n);
    LF2MI_LAUNCHED();
    item_bucket_kernel<<<grid1(n), 256, 0, st>>>(p->istart, n, p->item_bucket);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

// counting sort of buffer `src` (counts already in bcount) into buffer 1 - src
static int sort_into_other(lf2mi_parts_s *p, int src, int64_t nsrc, hipStream_t st) {
    if (int rc = scan_buckets(p, st)) return rc;
    if (int rc = fill_int(p->bcount, p->g.nbk, 0, st)) return rc;
    LF2MI_PROF("rebucket_scatter", st);
    scatter_kernel<<<grid1(nsrc, 256 * SCAT_PER), 256, 0, st>>>(nsrc, p->b32[src], p->b32[1 - src], p->g, p->bstart,
                                                                  p->bcount, p->plw);
    LF2MI_LAUNCHED();
    p->cur = 1 - src;
    return LF2MI_OK;
}

static int read_total(lf2mi_parts_s *p, hipStream_t st, int64_t *out) {
    int tot = 0;
    LF2MI_HIP(hipMemcpyAsync(&tot, p->bstart + p->g.nbk, sizeof(int), hipMemcpyDeviceToHost, st));
    LF2MI_HIP(hipStreamSynchronize(st));
    *out = tot;
    return LF2MI_OK;
}

static int64_t item_grid(const lf2mi_parts_s *p) { return p->g.nbk + p->np / ITEM_P + 1; }

// deposit kernel for brick edge 8 or any, with or without per-particle masses
template <int KIND, typename ACC>
static void launch_deposit(const DepArgs &A, bool b8, unsigned grid, hipStream_t st) {
    const bool pm = A.s.mass != nullptr;
    if (KIND == 0 && !pm) {
        if (b8) deposit_merged_kernel<ACC, 8><<<grid, DEP_WG, 0, st>>>(A);
        else deposit_merged_kernel<ACC, 0><<<grid, DEP_WG, 0, st>>>(A);
        return;
    }
    if (b8) {
        if (pm) deposit_f32_kernel<KIND, ACC, 8, true><<<grid, DEP_WG, 0, st>>>(A);
        else deposit_f32_kernel<KIND, ACC, 8, false><<<grid, DEP_WG, 0, st>>>(A);
    } else {
        if (pm) deposit_f32_kernel<KIND, ACC, 0, true><<<grid, DEP_WG, 0, st>>>(A);
        else deposit_f32_kernel<KIND, ACC, 0, false><<<grid, DEP_WG, 0, st>>>(A);
    }
}

// bricks with fewer particles are deposited by deposit_tiles_kernel, the others by
// deposit_dense_kernel
static constexpr int DENSE_MIN = 1024;
// workgroups of the dense deposit kernel, each taking a static share of the work list (512^3
// blobs, about 13 K groups: 2048 / 4096 / 8192 / 16384 / 32768 workers 1.10 / 1.08 / 1.06 /
// 1.05 / 1.05 ms for the step's deposit)
static constexpr int DENSE_WORKERS = 16384;
// workgroups of the dense deposit's launch that add the tiles kernel's shells meanwhile
static constexpr int FOLD_BLOCKS = 1024;

// the scalar deposit of a global mass into a periodic target with 8^3 bricks
static bool tiles_path(const lf2mi_parts_s *p, int kind, int target, bool pm) {
    return kind == LF2MI_PROJ_SCALAR && !pm && target == LF2MI_DEP_PERIODIC && p->g.B == 8 &&
           p->g.NB % TILE_B == 0 && p->N <= 1024;
}

// the shell and seam buffers of the tiles kernel (grown on demand): N^3 (9 + 8) / 64 values
// and 81 per tile
static int shell_buffers(lf2mi_parts_s *p, size_t esz, void **ysh, void **zsh, void **xsh) {
    const int64_t N = p->N, NB = p->g.NB;
    const int64_t ny = N * N * N / 64 * 9, nz = N * N * N / 64 * 8, nx = NB * NB * (NB / TILE_B) * 81;
    const int64_t bytes = (ny + nz + nx) * (int64_t)esz;
    if (p->shell_bytes < bytes) {
        if (p->shell) hipFree(p->shell);
        p->shell = nullptr;
        p->shell_bytes = 0;
        LF2MI_HIP(hipMalloc(&p->shell, (size_t)bytes));
        p->shell_bytes = bytes;
    }
    *ysh = p->shell;
    *zsh = (char *)p->shell + ny * (int64_t)esz;
    *xsh = (char *)p->shell + (ny + nz) * (int64_t)esz;
    return LF2MI_OK;
}

template <typename ACC>
static int launch_tiles_dense(lf2mi_parts_s *p, const DepArgs &A, bool overwrite, hipStream_t st) {
    void *ysh = nullptr, *zsh = nullptr, *xsh = nullptr;
    if (int rc = shell_buffers(p, sizeof(ACC), &ysh, &zsh, &xsh)) return rc;
    const int64_t NB = p->g.NB, N = p->N;
    deposit_tiles_kernel<ACC><<<(unsigned)(NB * NB * (NB / TILE_B)), 64 * TILE_B, 0, st>>>(
        A, DENSE_MIN, overwrite ? 1 : 0, (ACC *)ysh, (ACC *)zsh, (ACC *)xsh);
    LF2MI_LAUNCHED();
    if (int rc = fill_int(p->ndgroup, 2, 0, st)) return rc;
    dense_groups_kernel<<<grid1(p->g.nbk, DG_WG), DG_WG, 0, st>>>(p->bstart, p->g.nbk, DENSE_MIN, p->dgroups, p->ndgroup);
    LF2MI_LAUNCHED();
    deposit_dense_fold_kernel<ACC><<<FOLD_BLOCKS + DENSE_WORKERS, DEP_WG, 0, st>>>(
        A, p->dgroups, p->ndgroup, (const ACC *)ysh, (const ACC *)zsh, (const ACC *)xsh, FOLD_BLOCKS);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

// the other kinds (and the scalar kind with per-particle masses) into a periodic target
// with 8^3 bricks
static bool tiles_c_path(const lf2mi_parts_s *p, int kind, int target, bool pm) {
    return (kind != LF2MI_PROJ_SCALAR || pm) && target == LF2MI_DEP_PERIODIC && p->g.B == 8 &&
           p->g.NB % TC_TB == 0 && p->N <= 1024;
}

// the shell buffer of deposit_tiles_c_kernel (grown on demand), its layout, the fold and the
// item deposit of the dense bricks
template <int KIND>
static int launch_tiles_c(lf2mi_parts_s *p, DepArgs A, bool pm, bool overwrite, hipStream_t st) {
    constexpr int C = KindC<KIND>::C, TB = TC_TB;
    const int64_t N = p->N, NB = p->g.NB, ntile = NB / TB;
    TcArgs T{};
    TcFold F{};
    F.C = C;
    F.TX = 8 * TB;
    F.ntile = (int)ntile;
    bool ry[2] = {false, false}, rz[2] = {false, false}, rx[2] = {false, false};  // layers 0 / 7
    int64_t off = 0;
    for (int c = 0; c < C; ++c) {
        const int mask = TC_MASK[KIND][c], ngp = TC_NGP[KIND][c];
        T.shell_off[c] = F.shell_off[c] = off;
        F.rows[c] = tc_rows(mask, ngp);
        F.ns[c] = tc_nl(mask, 0) + tc_nh(ngp, 0);
        for (int d = 0; d < 3; ++d) {
            F.lo[c][d] = -tc_nl(mask, d);
            F.hi[c][d] = 7 + tc_nh(ngp, d);
            F.w[c][d] = tc_w(mask, ngp, d);
        }
        ry[0] |= tc_nh(ngp, 1) != 0;
        ry[1] |= tc_nl(mask, 1) != 0;
        rz[0] |= tc_nh(ngp, 2) != 0;
        rz[1] |= tc_nl(mask, 2) != 0;
        rx[0] |= tc_nh(ngp, 0) != 0;
        rx[1] |= tc_nl(mask, 0) != 0;
        off += NB * NB * F.rows[c] * N;
    }
    int64_t seams = 0;
    for (int c = 0; c < C; ++c) {
        T.seam_off[c] = F.seam_off[c] = off + seams;
        seams += NB * NB * ntile * tc_seam(TC_MASK[KIND][c], TC_NGP[KIND][c]);
    }
    auto layers = [](const bool (&r)[2], int last, int *s, int *cpl, int &ns) {
        ns = 0;
        int nc = 0;
        for (int l = 0; l <= 7; ++l) {
            const bool recv = (l == 0 && r[0]) || (l == 7 && r[1]);
            if (recv) s[ns++] = l == 7 ? last : 0;
            else if (cpl) cpl[nc++] = l;
        }
    };
    layers(ry, 7, F.sy, F.cy, F.ny);
    layers(rz, 7, F.sz, F.cz, F.nz);
    layers(rx, F.TX - 1, F.sx, nullptr, F.nx);
    const int64_t bytes = (off + seams) * (int64_t)sizeof(float);
    if (p->shell_bytes < bytes) {
        if (p->shell) hipFree(p->shell);
        p->shell = nullptr;
        p->shell_bytes = 0;
        LF2MI_HIP(hipMalloc(&p->shell, (size_t)bytes));
        p->shell_bytes = bytes;
    }
    float *shell = (float *)p->shell;
    const unsigned grid = (unsigned)(NB * NB * ntile);
    if (pm) deposit_tiles_c_kernel<KIND, true, TB><<<grid, 64 * TB, 0, st>>>(A, DENSE_MIN, overwrite ? 1 : 0, T, shell);
    else deposit_tiles_c_kernel<KIND, false, TB><<<grid, 64 * TB, 0, st>>>(A, DENSE_MIN, overwrite ? 1 : 0, T, shell);
    LF2MI_LAUNCHED();
    const int64_t rows = (int64_t)F.ny * NB * N + (int64_t)(8 - F.ny) * NB * F.nz * NB;
    if (rows > 0) {
        fold_rows_kernel<C><<<(unsigned)rows, 256, 0, st>>>(A.field, (int)N, A.h, (int)NB, F, shell);
        LF2MI_LAUNCHED();
    }
    const int64_t edges = (int64_t)F.nx * ntile * (8 - F.ny) * NB * (8 - F.nz) * NB;
    if (edges > 0) {
        fold_edges_kernel<C><<<grid1(edges), 256, 0, st>>>(A.field, (int)N, A.h, (int)NB, F, shell, edges);
        LF2MI_LAUNCHED();
    }
    if (int rc = fill_int(p->nditem, 2, 0, st)) return rc;
    dense_items_kernel<<<grid1(p->g.nbk, DG_WG), DG_WG, 0, st>>>(p->bstart, p->istart, p->g.nbk, DENSE_MIN, p->ditems,
                                                                 p->nditem);
    LF2MI_LAUNCHED();
    A.dlist = p->ditems;
    A.dcount = p->nditem;
    launch_deposit<KIND, float>(A, true, (unsigned)std::min<int64_t>(item_grid(p), DENSE_WORKERS), st);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

static int ic_common(lf2mi_parts_s *p, IcArgs &A, hipStream_t st) {
    if (A.np > p->cap) return fail(LF2MI_ERANGE, "%lld particles exceed the capacity %lld", (long long)A.np, (long long)p->cap);
    if (p->dtype == LF2MI_F32) {
        if (int rc = fill_int(p->bcount, p->g.nbk, 0, st)) return rc;
        const int src = 1 - p->cur;
        ic_f32_kernel<<<grid1(A.np), 256, 0, st>>>(A, p->b32[src], p->g, p->bcount);
        LF2MI_LAUNCHED();
        if (p->plw && A.np > 0)
            LF2MI_HIP(hipMemsetAsync(p->b32[src].pl, 0, (size_t)A.np * p->plw * sizeof(unsigned long long), st));
        if (int rc = sort_into_other(p, src, A.np, st)) return rc;
        if (int rc = read_total(p, st, &p->np)) return rc;
    } else {
        ic_f64_kernel<<<grid1(A.np), 256, 0, st>>>(A, p->b64, p->N);
        LF2MI_LAUNCHED();
        if (p->plw && A.np > 0)
            LF2MI_HIP(hipMemsetAsync(p->b64.pl, 0, (size_t)A.np * p->plw * sizeof(unsigned long long), st));
        LF2MI_HIP(hipStreamSynchronize(st));
        p->np = A.np;
    }
    // synthetic particles: per-particle masses (when kept) are 1
    p->total_mass = (p->mass_src ? 1.0 : p->mass_global) * (double)p->np;
    return LF2MI_OK;
}

}  // namespace lf2mi

using namespace lf2mi;

namespace lf2mi {

// sum of the stored per-particle masses: per-workgroup partial sums (grid-stride, a tree in
// LDS), then mass_total_kernel adds the partials in order, so the total is the same from
// run to run
template <typename T>
__global__ __launch_bounds__(RED_WG) void mass_partial_kernel(int64_t np, const T *__restrict__ mass,
                                                              double *__restrict__ partial) {
    __shared__ double lds[RED_WG];
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * RED_WG;
    for (int64_t i = (int64_t)blockIdx.x * RED_WG + threadIdx.x; i < np; i += stride) m += (double)mass[i];
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int h = RED_WG / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}
__global__ void mass_total_kernel(const double *__restrict__ partial, int nblocks, double *__restrict__ total) {
    double t = 0.0;
    for (int b = 0; b < nblocks; ++b) t += partial[b];
    *total = t;
}

// per-particle masses from the value of type S (double, or float) at byte `off` of each
// payload
template <typename T, typename S>
__global__ void payload_mass_kernel(int64_t np, const unsigned long long *__restrict__ pl, int plw, int64_t off,
                                    T *__restrict__ mass) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned char *b = reinterpret_cast<const unsigned char *>(pl + i * plw);
    S v;
    memcpy(&v, b + off, sizeof(S));
    mass[i] = (T)v;
}

// the Real[3] at byte `off` of each payload into three planes of `cap` values
template <typename T, typename R>
__global__ void payload_vec_kernel(int64_t np, const unsigned long long *__restrict__ pl, int plw, int64_t off,
                                   T *__restrict__ out, int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned char *b = reinterpret_cast<const unsigned char *>(pl + i * plw) + off;
    R v[3];
    memcpy(v, b, sizeof v);
    out[i] = (T)v[0];
    out[cap + i] = (T)v[1];
    out[2 * cap + i] = (T)v[2];
}

// point the deposit's velocity planes at the vector selected by lf2mi_parts_vector_from
// (gathered from the payload on this stream)
template <typename T>
static int vector_source(lf2mi_parts_s *p, const unsigned long long *pl, T *(&vel)[3], hipStream_t st) {
    if (p->vec_off < 0) return LF2MI_OK;
    const int64_t cap = std::max<int64_t>(p->cap, 1);
    if (!p->vsrc)
        if (int rc = dalloc(p, &p->vsrc, (size_t)cap * 3 * sizeof(T))) return rc;
    T *out = static_cast<T *>(p->vsrc);
    if (p->vec_real == LF2MI_F32)
        payload_vec_kernel<T, float><<<grid1(p->np), 256, 0, st>>>(p->np, pl, p->plw, p->vec_off, out, cap);
    else
        payload_vec_kernel<T, double><<<grid1(p->np), 256, 0, st>>>(p->np, pl, p->plw, p->vec_off, out, cap);
    LF2MI_LAUNCHED();
    for (int d = 0; d < 3; ++d) vel[d] = out + d * cap;
    return LF2MI_OK;
}

}  // namespace lf2mi

extern "C" {

int lf2mi_parts_create(int N, int64_t capacity, int dtype, int flags, lf2mi_parts *out) {
    if (!out) return fail(LF2MI_EINVAL, "lf2mi_parts_create: null output");
    *out = nullptr;
    if (N < 2 || N > 1290) return fail(LF2MI_ERANGE, "lattice size N=%d outside [2, 1290]", N);
    if (capacity < 0 || capacity > 2000000000LL) return fail(LF2MI_ERANGE, "capacity %lld outside [0, 2e9]", (long long)capacity);
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    lf2mi_parts_s *p = new (std::nothrow) lf2mi_parts_s();
    if (!p) return fail(LF2MI_ENOMEM, "host allocation failed");
    p->N = N;
    p->dtype = dtype;
    p->flags = flags;
    hipGetDevice(&p->device);
    p->cap = capacity;
    p->np = 0;
    p->mass_global = 1.0;
    p->cur = 0;
    p->plw = 0;
    p->mass_src = (flags & LF2MI_PARTS_MASS) ? 1 : 0;
    p->vec_off = -1;
    p->vec_real = LF2MI_F32;
    p->aos_vel = -1;
    p->vsrc = nullptr;
    int B = 1;
    for (int b = MAXB; b >= 1; --b)
        if (N % b == 0) { B = b; break; }
    p->g.N = N;
    p->g.B = B;
    p->g.NB = N / B;
    p->g.nbk = (int64_t)p->g.NB * p->g.NB * p->g.NB;
    p->g.lgN = is_pow2(N) ? __builtin_ctz((unsigned)N) : -1;
    p->g.lgB = is_pow2(B) ? __builtin_ctz((unsigned)B) : -1;
    const size_t cap = (size_t)std::max<int64_t>(capacity, 1);
    int rc = LF2MI_OK;
    auto A = [&](void **ptr, size_t bytes) {
        if (!rc) rc = dalloc(p, ptr, bytes);
    };
    if (dtype == LF2MI_F32) {
        for (int k = 0; k < 2; ++k) {
            Soa32 &s = p->b32[k];
            A((void **)&s.cell, cap * sizeof(int32_t));
            for (int d = 0; d < 3; ++d) {
                A((void **)&s.off[d], cap * sizeof(float));
                A((void **)&s.vel[d], cap * sizeof(float));
            }
            s.mass = nullptr;
            s.id = nullptr;
            s.pl = nullptr;
            if (flags & LF2MI_PARTS_MASS) A((void **)&s.mass, cap * sizeof(float));
            if (flags & LF2MI_PARTS_ID) A((void **)&s.id, cap * sizeof(int64_t));
        }
        const size_t nb = (size_t)p->g.nbk + 1;
        A((void **)&p->bstart, nb * sizeof(int));
        A((void **)&p->istart, nb * sizeof(int));
        A((void **)&p->bcount, nb * sizeof(int));
        A((void **)&p->bsum, 2 * 4096 * sizeof(int));
        A((void **)&p->item_bucket, (size_t)(p->g.nbk + cap / ITEM_P + 1) * sizeof(int));
        A((void **)&p->vmax_f32, sizeof(unsigned int));
        A((void **)&p->chunks, (size_t)p->g.nbk * sizeof(int4));
        A((void **)&p->nchunk, 2 * sizeof(int));
        A((void **)&p->dgroups, (size_t)(cap / (DEP_MERGE * ITEM_P) + p->g.nbk + 1) * sizeof(int2));
        A((void **)&p->ndgroup, 2 * sizeof(int));
        A((void **)&p->ditems, (size_t)(p->g.nbk + cap / ITEM_P + 1) * sizeof(int));
        A((void **)&p->nditem, 2 * sizeof(int));
    } else {
        Soa64 &s = p->b64;
        A((void **)&s.cell, cap * sizeof(int32_t));
        for (int d = 0; d < 3; ++d) {
            A((void **)&s.pos[d], cap * sizeof(double));
            A((void **)&s.vel[d], cap * sizeof(double));
        }
        s.mass = nullptr;
        s.id = nullptr;
        s.pl = nullptr;
        if (flags & LF2MI_PARTS_MASS) A((void **)&s.mass, cap * sizeof(double));
        if (flags & LF2MI_PARTS_ID) A((void **)&s.id, cap * sizeof(int64_t));
        A((void **)&p->vmax_f64, sizeof(unsigned long long));
    }
    A((void **)&p->red, (size_t)(RED_BLOCKS + 1) * RED_Q * sizeof(double));
    if (rc) {
        for (void *q : p->allocs) hipFree(q);
        delete p;
        return rc;
    }
    if (dtype == LF2MI_F32) {
        hipMemset(p->bstart, 0, ((size_t)p->g.nbk + 1) * sizeof(int));
        hipMemset(p->istart, 0, ((size_t)p->g.nbk + 1) * sizeof(int));
    }
    *out = p;
    return LF2MI_OK;
}

int lf2mi_parts_destroy(lf2mi_parts p) {
    if (!p) return LF2MI_OK;
    for (void *q : p->allocs) hipFree(q);
    if (p->fx) hipFree(p->fx);
    if (p->shell) hipFree(p->shell);
    delete p;
    return LF2MI_OK;
}

int lf2mi_parts_set_mass(lf2mi_parts p, double m) {
    if (!p) return fail(LF2MI_EINVAL, "null container");
    if (p->total_mass != 0.0 && !(p->flags & LF2MI_PARTS_MASS)) p->total_mass = m * (double)p->np;
    p->mass_global = m;
    if (!(p->flags & LF2MI_PARTS_MASS)) p->total_mass = m * (double)p->np;
    return LF2MI_OK;
}

int lf2mi_parts_count(lf2mi_parts p, int64_t *np) {
    if (!p || !np) return fail(LF2MI_EINVAL, "null argument");
    *np = p->np;
    return LF2MI_OK;
}

int lf2mi_parts_bytes(lf2mi_parts p, int64_t *bytes) {
    if (!p || !bytes) return fail(LF2MI_EINVAL, "null argument");
    *bytes = p->bytes;
    return LF2MI_OK;
}

// the upload; pl (may be null) holds each input particle's payload (p->plw words, in input
// order), zeros when the container keeps payloads and none is given
static int upload_impl(lf2mi_parts p, int64_t np, const double *pos, const double *vel, const double *mass,
                       const int64_t *id, const unsigned char *pl, int64_t *kept, void *stream) {
    if (!p || np < 0 || (np > 0 && (!pos || !vel))) return fail(LF2MI_EINVAL, "lf2mi_parts_upload: bad argument");
    if (np > p->cap) return fail(LF2MI_ERANGE, "%lld particles exceed the capacity %lld", (long long)np, (long long)p->cap);
    if ((p->flags & LF2MI_PARTS_MASS) && np > 0 && !mass) return fail(LF2MI_EINVAL, "per-particle masses required");
    hipStream_t st = as_stream(stream);
    // drop what the reference's setCoord rejects (negative cell coordinates) on the host,
    // so the device arrays hold only valid particles
    std::vector<int64_t> keep;
    keep.reserve((size_t)np);
    for (int64_t i = 0; i < np; ++i) {
        bool ok = true;
        for (int d = 0; d < 3; ++d)
            if ((int)floor(pos[3 * i + d] * (double)p->N) % p->N < 0) ok = false;
        if (ok) keep.push_back(i);
    }
    const int64_t nk = (int64_t)keep.size();
    std::vector<double> hp((size_t)nk * 3), hv((size_t)nk * 3), hm(mass ? (size_t)nk : 0);
    std::vector<int64_t> hid((size_t)nk);
    double msum = 0.0;
    for (int64_t k = 0; k < nk; ++k) {
        const int64_t i = keep[(size_t)k];
        for (int d = 0; d < 3; ++d) {
            hp[3 * k + d] = pos[3 * i + d];
            hv[3 * k + d] = vel[3 * i + d];
        }
        if (mass) {
            hm[(size_t)k] = mass[i];
            msum += mass[i];
        }
        hid[(size_t)k] = id ? id[i] : i;
    }
    const size_t plb = (size_t)p->plw * sizeof(unsigned long long);
    std::vector<unsigned char> hpl(plb * (size_t)nk, 0);
    if (pl)
        for (int64_t k = 0; k < nk; ++k) memcpy(hpl.data() + plb * (size_t)k, pl + plb * (size_t)keep[(size_t)k], plb);
    double *dpos = nullptr, *dvel = nullptr, *dmass = nullptr;
    int64_t *did = nullptr;
    unsigned long long *dpl = nullptr;
    if (plb) LF2MI_HIP(hipMalloc((void **)&dpl, plb * (size_t)std::max<int64_t>(nk, 1)));
    const size_t n3 = (size_t)std::max<int64_t>(nk, 1) * 3;
    LF2MI_HIP(hipMalloc((void **)&dpos, n3 * sizeof(double)));
    LF2MI_HIP(hipMalloc((void **)&dvel, n3 * sizeof(double)));
    LF2MI_HIP(hipMalloc((void **)&did, (n3 / 3) * sizeof(int64_t)));
    if (mass) LF2MI_HIP(hipMalloc((void **)&dmass, (n3 / 3) * sizeof(double)));
    int rc = LF2MI_OK;
    hipError_t e = hipSuccess;
    if (nk > 0) {
        e = hipMemcpyAsync(dpos, hp.data(), nk * 3 * sizeof(double), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dvel, hv.data(), nk * 3 * sizeof(double), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(did, hid.data(), nk * sizeof(int64_t), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && mass) e = hipMemcpyAsync(dmass, hm.data(), nk * sizeof(double), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && plb) e = hipMemcpyAsync(dpl, hpl.data(), plb * (size_t)nk, hipMemcpyHostToDevice, st);
    }
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync(upload)", __FILE__, __LINE__);
    if (!rc) {
        if (p->dtype == LF2MI_F32) {
            rc = fill_int(p->bcount, p->g.nbk, 0, st);
            const int src = 1 - p->cur;
            if (!rc && nk > 0) {
                bin_upload_kernel<<<grid1(nk), 256, 0, st>>>(nk, dpos, dvel, dmass, did, dpl, p->plw, p->b32[src], p->g,
                                                             p->bcount);
                if ((e = hipGetLastError()) != hipSuccess) rc = hip_fail(e, "bin_upload_kernel", __FILE__, __LINE__);
            }
            if (!rc) rc = sort_into_other(p, src, nk, st);
            if (!rc) rc = read_total(p, st, &p->np);
        } else {
            if (nk > 0) {
                upload_f64_kernel<<<grid1(nk), 256, 0, st>>>(nk, dpos, dvel, dmass, did, p->b64, p->N);
                if ((e = hipGetLastError()) != hipSuccess) rc = hip_fail(e, "upload_f64_kernel", __FILE__, __LINE__);
                if (!rc && plb && (e = hipMemcpyAsync(p->b64.pl, dpl, plb * (size_t)nk, hipMemcpyDeviceToDevice, st)) !=
                                      hipSuccess)
                    rc = hip_fail(e, "payload", __FILE__, __LINE__);
            }
            if (!rc && (e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_fail(e, "sync", __FILE__, __LINE__);
            p->np = nk;
        }
    }
    hipStreamSynchronize(st);
    hipFree(dpos);
    hipFree(dvel);
    hipFree(did);
    if (dmass) hipFree(dmass);
    if (dpl) hipFree(dpl);
    if (rc) return rc;
    p->total_mass = mass && p->mass_src ? msum : p->mass_global * (double)p->np;
    if (kept) *kept = p->np;
    return LF2MI_OK;
}

int lf2mi_parts_upload(lf2mi_parts p, int64_t np, const double *pos, const double *vel, const double *mass,
                       const int64_t *id, int64_t *kept, void *stream) {
    return upload_impl(p, np, pos, vel, mass, id, nullptr, kept, stream);
}

int lf2mi_parts_download(lf2mi_parts p, double *pos, double *vel, double *mass, int64_t *id, int32_t *cell,
                         void *stream) {
    if (!p) return fail(LF2MI_EINVAL, "null container");
    hipStream_t st = as_stream(stream);
    const int64_t np = p->np;
    if (np == 0) return LF2MI_OK;
    double *dpos = nullptr, *dvel = nullptr, *dmass = nullptr;
    LF2MI_HIP(hipMalloc((void **)&dpos, np * 3 * sizeof(double)));
    LF2MI_HIP(hipMalloc((void **)&dvel, np * 3 * sizeof(double)));
    LF2MI_HIP(hipMalloc((void **)&dmass, np * sizeof(double)));
    int rc = LF2MI_OK;
    if (p->dtype == LF2MI_F32) {
        Soa32 s = p->b32[p->cur];
        if (!p->mass_src) s.mass = nullptr;  // the global mass is the one in use
        download_f32_kernel<<<grid1(np), 256, 0, st>>>(np, s, p->N, dpos, dvel, dmass, p->mass_global);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && pos) e = hipMemcpyAsync(pos, dpos, np * 3 * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && vel) e = hipMemcpyAsync(vel, dvel, np * 3 * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && mass) e = hipMemcpyAsync(mass, dmass, np * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && id) {
            if (s.id) e = hipMemcpyAsync(id, s.id, np * sizeof(int64_t), hipMemcpyDeviceToHost, st);
            else rc = fail(LF2MI_EINVAL, "container keeps no IDs");
        }
        if (e == hipSuccess && cell) e = hipMemcpyAsync(cell, s.cell, np * sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) rc = hip_fail(e, "download", __FILE__, __LINE__);
    } else {
        const Soa64 &s = p->b64;
        download_f64_kernel<<<grid1(np), 256, 0, st>>>(np, s, dpos, dvel);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && pos) e = hipMemcpyAsync(pos, dpos, np * 3 * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && vel) e = hipMemcpyAsync(vel, dvel, np * 3 * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && mass) {
            if (s.mass && p->mass_src) e = hipMemcpyAsync(mass, s.mass, np * sizeof(double), hipMemcpyDeviceToHost, st);
            else for (int64_t i = 0; i < np; ++i) mass[i] = p->mass_global;
        }
        if (e == hipSuccess && id) {
            if (s.id) e = hipMemcpyAsync(id, s.id, np * sizeof(int64_t), hipMemcpyDeviceToHost, st);
            else rc = fail(LF2MI_EINVAL, "container keeps no IDs");
        }
        if (e == hipSuccess && cell) e = hipMemcpyAsync(cell, s.cell, np * sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) rc = hip_fail(e, "download", __FILE__, __LINE__);
    }
    hipStreamSynchronize(st);
    hipFree(dpos);
    hipFree(dvel);
    hipFree(dmass);
    return rc;
}

// ---------------------------------------------------------------- user structs
int lf2mi_parts_set_payload(lf2mi_parts p, int bytes) {
    if (!p) return fail(LF2MI_EINVAL, "null container");
    if (bytes < 0 || bytes > 4096 || bytes % 8) return fail(LF2MI_EINVAL, "payload of %d bytes: a multiple of 8 up to 4096", bytes);
    if (p->np > 0 || p->plw) return fail(LF2MI_EINVAL, "the payload is set once, before the first upload");
    if (bytes == 0) return LF2MI_OK;
    const size_t cap = (size_t)std::max<int64_t>(p->cap, 1);
    int rc = LF2MI_OK;
    if (p->dtype == LF2MI_F32) {
        for (int k = 0; k < 2 && !rc; ++k) {
            Soa32 &s = p->b32[k];
            rc = dalloc(p, (void **)&s.pl, cap * (size_t)bytes);
            if (!rc && !s.mass) rc = dalloc(p, (void **)&s.mass, cap * sizeof(float));
        }
    } else {
        rc = dalloc(p, (void **)&p->b64.pl, cap * (size_t)bytes);
        if (!rc && !p->b64.mass) rc = dalloc(p, (void **)&p->b64.mass, cap * sizeof(double));
    }
    if (rc) return rc;
    p->plw = bytes / 8;
    return LF2MI_OK;
}

static bool real_field_ok(int64_t off, int64_t size, int64_t stride) { return off < 0 || off + size <= stride; }

int lf2mi_parts_upload_aos(lf2mi_parts p, int64_t n, const void *base, int64_t stride, int real_dtype,
                           int64_t off_pos, int64_t off_vel, int64_t off_id, int64_t off_mass, int64_t *kept,
                           void *stream) {
    if (!p || n < 0 || (n > 0 && !base) || stride <= 0) return fail(LF2MI_EINVAL, "lf2mi_parts_upload_aos: bad argument");
    if (real_dtype != LF2MI_F32 && real_dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad real dtype %d", real_dtype);
    const int64_t rs = real_dtype == LF2MI_F32 ? 4 : 8;
    if (off_pos < 0 || off_vel < 0 || !real_field_ok(off_pos, 3 * rs, stride) || !real_field_ok(off_vel, 3 * rs, stride) ||
        !real_field_ok(off_id, 8, stride) || !real_field_ok(off_mass, 8, stride))
        return fail(LF2MI_EINVAL, "field offsets outside the %lld-byte struct", (long long)stride);
    if (p->plw && stride != (int64_t)p->plw * 8)
        return fail(LF2MI_EINVAL, "struct of %lld bytes against a payload of %d", (long long)stride, p->plw * 8);
    if (off_mass >= 0 && !(p->dtype == LF2MI_F32 ? (void *)p->b32[0].mass : (void *)p->b64.mass))
        return fail(LF2MI_EINVAL, "per-particle masses need a container with masses or a payload");
    const unsigned char *b = static_cast<const unsigned char *>(base);
    std::vector<double> pos((size_t)n * 3), vel((size_t)n * 3), mass(off_mass >= 0 ? (size_t)n : 0);
    std::vector<int64_t> id(off_id >= 0 ? (size_t)n : 0);
    for (int64_t i = 0; i < n; ++i) {
        const unsigned char *q = b + i * stride;
        for (int d = 0; d < 3; ++d) {
            if (rs == 4) {
                float x, v;
                memcpy(&x, q + off_pos + 4 * d, 4);
                memcpy(&v, q + off_vel + 4 * d, 4);
                pos[3 * i + d] = x;
                vel[3 * i + d] = v;
            } else {
                memcpy(&pos[3 * i + d], q + off_pos + 8 * d, 8);
                memcpy(&vel[3 * i + d], q + off_vel + 8 * d, 8);
            }
        }
        if (off_id >= 0) memcpy(&id[(size_t)i], q + off_id, 8);
        if (off_mass >= 0) memcpy(&mass[(size_t)i], q + off_mass, 8);
    }
    if (p->plw) p->aos_vel = off_vel;
    // a container with per-particle masses whose structs carry none readable as a double (a
    // float member, gathered afterwards by lf2mi_parts_mass_from) starts at mass 0
    const bool zero_mass = off_mass < 0 && (p->flags & LF2MI_PARTS_MASS) && p->plw;
    if (zero_mass) mass.assign((size_t)n, 0.0);
    return upload_impl(p, n, pos.data(), vel.data(), (off_mass >= 0 || zero_mass) ? mass.data() : nullptr,
                       off_id >= 0 ? id.data() : nullptr, p->plw ? b : nullptr, kept, stream);
}

int lf2mi_parts_download_aos(lf2mi_parts p, void *base, int64_t n_room, int64_t stride, int real_dtype,
                             int64_t off_pos, int64_t off_vel, int64_t off_id, int64_t *n_out, void *stream) {
    if (!p || stride <= 0 || n_room < 0) return fail(LF2MI_EINVAL, "lf2mi_parts_download_aos: bad argument");
    if (real_dtype != LF2MI_F32 && real_dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad real dtype %d", real_dtype);
    const int64_t rs = real_dtype == LF2MI_F32 ? 4 : 8;
    if (!real_field_ok(off_pos, 3 * rs, stride) || !real_field_ok(off_vel, 3 * rs, stride) || !real_field_ok(off_id, 8, stride))
        return fail(LF2MI_EINVAL, "field offsets outside the %lld-byte struct", (long long)stride);
    if (p->plw && stride != (int64_t)p->plw * 8)
        return fail(LF2MI_EINVAL, "struct of %lld bytes against a payload of %d", (long long)stride, p->plw * 8);
    const int64_t np = p->np;
    if (n_out) *n_out = np;
    if (np > n_room) return fail(LF2MI_ERANGE, "%lld particles exceed the room for %lld", (long long)np, (long long)n_room);
    if (np == 0) return LF2MI_OK;
    if (!base) return fail(LF2MI_EINVAL, "lf2mi_parts_download_aos: null buffer");
    hipStream_t st = as_stream(stream);
    std::vector<double> pos((size_t)np * 3), vel((size_t)np * 3);
    std::vector<int64_t> id(off_id >= 0 ? (size_t)np : 0);
    if (int rc = lf2mi_parts_download(p, pos.data(), vel.data(), nullptr, off_id >= 0 ? id.data() : nullptr, nullptr,
                                      stream))
        return rc;
    unsigned char *b = static_cast<unsigned char *>(base);
    if (p->plw) {
        const unsigned long long *pl = p->dtype == LF2MI_F32 ? p->b32[p->cur].pl : p->b64.pl;
        LF2MI_HIP(hipMemcpyAsync(b, pl, (size_t)np * p->plw * 8, hipMemcpyDeviceToHost, st));
        LF2MI_HIP(hipStreamSynchronize(st));
    }
    for (int64_t i = 0; i < np; ++i) {
        unsigned char *q = b + i * stride;
        for (int d = 0; d < 3; ++d) {
            if (off_pos >= 0) {
                if (rs == 4) {
                    const float x = (float)pos[3 * i + d];
                    memcpy(q + off_pos + 4 * d, &x, 4);
                } else {
                    memcpy(q + off_pos + 8 * d, &pos[3 * i + d], 8);
                }
            }
            if (off_vel >= 0) {
                if (rs == 4) {
                    const float v = (float)vel[3 * i + d];
                    memcpy(q + off_vel + 4 * d, &v, 4);
                } else {
                    memcpy(q + off_vel + 8 * d, &vel[3 * i + d], 8);
                }
            }
        }
        if (off_id >= 0) memcpy(q + off_id, &id[(size_t)i], 8);
    }
    return LF2MI_OK;
}

int lf2mi_parts_mass_from(lf2mi_parts p, int where, int64_t offset, void *stream) {
    if (!p) return fail(LF2MI_EINVAL, "null container");
    if (where != LF2MI_MASS_INFO && where != LF2MI_MASS_PART && where != LF2MI_MASS_PART_F32)
        return fail(LF2MI_EINVAL, "bad mass source %d", where);
    if (where == LF2MI_MASS_INFO) {
        p->mass_src = 0;
        p->total_mass = p->mass_global * (double)p->np;
        return LF2MI_OK;
    }
    const bool has_mass = p->dtype == LF2MI_F32 ? p->b32[0].mass != nullptr : p->b64.mass != nullptr;
    if (!has_mass) return fail(LF2MI_EINVAL, "per-particle masses need a container with masses or a payload");
    const int64_t msize = where == LF2MI_MASS_PART_F32 ? 4 : 8;
    if (offset >= 0 && (!p->plw || offset + msize > (int64_t)p->plw * 8)) {
        return fail(LF2MI_EINVAL, "mass offset %lld outside the payload", (long long)offset);
    }
    hipStream_t st = as_stream(stream);
    double *res = p->red + (int64_t)RED_BLOCKS * RED_Q;  // the partials before it
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (p->np + RED_WG - 1) / RED_WG));
    if (offset >= 0 && p->np > 0) {
        const bool f4 = where == LF2MI_MASS_PART_F32;
        const unsigned g = grid1(p->np);
        if (p->dtype == LF2MI_F32) {
            Soa32 &s = p->b32[p->cur];
            if (f4) payload_mass_kernel<float, float><<<g, 256, 0, st>>>(p->np, s.pl, p->plw, offset, s.mass);
            else payload_mass_kernel<float, double><<<g, 256, 0, st>>>(p->np, s.pl, p->plw, offset, s.mass);
        } else {
            Soa64 &s = p->b64;
            if (f4) payload_mass_kernel<double, float><<<g, 256, 0, st>>>(p->np, s.pl, p->plw, offset, s.mass);
            else payload_mass_kernel<double, double><<<g, 256, 0, st>>>(p->np, s.pl, p->plw, offset, s.mass);
        }
        LF2MI_LAUNCHED();
    }
    if (p->dtype == LF2MI_F32) mass_partial_kernel<float><<<nb, RED_WG, 0, st>>>(p->np, p->b32[p->cur].mass, p->red);
    else mass_partial_kernel<double><<<nb, RED_WG, 0, st>>>(p->np, p->b64.mass, p->red);
    LF2MI_LAUNCHED();
    mass_total_kernel<<<1, 1, 0, st>>>(p->red, nb, res);
    LF2MI_LAUNCHED();
    double total = 0.0;
    LF2MI_HIP(hipMemcpyAsync(&total, res, sizeof(double), hipMemcpyDeviceToHost, st));
    LF2MI_HIP(hipStreamSynchronize(st));
    p->mass_src = 1;
    p->total_mass = total;
    return LF2MI_OK;
}

int lf2mi_parts_vector_from(lf2mi_parts p, int64_t offset, int real_dtype) {
    if (!p) return fail(LF2MI_EINVAL, "null container");
    // the payload's copy of vel is the uploaded one: the kicked velocity lives in the container
    if (offset < 0 || offset == p->aos_vel) {
        p->vec_off = -1;
        return LF2MI_OK;
    }
    if (real_dtype != LF2MI_F32 && real_dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad real dtype %d", real_dtype);
    const int64_t rs = real_dtype == LF2MI_F32 ? 4 : 8;
    if (!p->plw || offset + 3 * rs > (int64_t)p->plw * 8)
        return fail(LF2MI_EINVAL, "vector offset %lld outside the payload", (long long)offset);
    p->vec_off = offset;
    p->vec_real = real_dtype;
    return LF2MI_OK;
}

int lf2mi_parts_download_box(lf2mi_parts p, const int *lo, const int *hi, int64_t max_out, double *pos, double *vel,
                             double *mass, int64_t *n_out, void *stream) {
    if (!p || !lo || !hi || !n_out || max_out < 0 || (max_out > 0 && (!pos || !vel || !mass)))
        return fail(LF2MI_EINVAL, "lf2mi_parts_download_box: bad argument");
    BoxArgs A{};
    for (int d = 0; d < 3; ++d) {
        if (lo[d] < 0 || hi[d] > p->N || lo[d] >= hi[d]) return fail(LF2MI_ERANGE, "box [%d, %d) outside [0, %d)", lo[d], hi[d], p->N);
        A.lo[d] = lo[d];
        A.hi[d] = hi[d];
    }
    *n_out = 0;
    if (p->np == 0) return LF2MI_OK;
    hipStream_t st = as_stream(stream);
    const int64_t m = std::max<int64_t>(max_out, 1);
    double *d_pos = nullptr, *d_vel = nullptr, *d_mass = nullptr;
    int *d_cur = nullptr;
    hipError_t e = hipMalloc((void **)&d_pos, m * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_vel, m * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_mass, m * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&d_cur, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(d_cur, 0, sizeof(int), st);
    int cnt = 0;
    if (e == hipSuccess) {
        A.N = p->N;
        A.np = p->np;
        A.max_out = max_out;
        A.cursor = d_cur;
        A.pos = d_pos;
        A.vel = d_vel;
        A.mass = d_mass;
        A.mass_global = p->mass_global;
        Soa32 s32 = p->b32[p->cur];
        Soa64 s64 = p->b64;
        if (!p->mass_src) s32.mass = nullptr, s64.mass = nullptr;
        if (p->dtype == LF2MI_F32) download_box_f32_kernel<<<grid1(p->np), 256, 0, st>>>(A, s32);
        else download_box_f64_kernel<<<grid1(p->np), 256, 0, st>>>(A, s64);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d_cur, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const int64_t n = std::min<int64_t>(cnt, max_out);
    if (e == hipSuccess && n > 0) e = hipMemcpy(pos, d_pos, n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && n > 0) e = hipMemcpy(vel, d_vel, n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && n > 0) e = hipMemcpy(mass, d_mass, n * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d_pos);
    hipFree(d_vel);
    hipFree(d_mass);
    hipFree(d_cur);
    if (e != hipSuccess) return hip_fail(e, "lf2mi_parts_download_box", __FILE__, __LINE__);
    *n_out = cnt;
    if (cnt > max_out) return fail(LF2MI_ERANGE, "%d particles in the box exceed max_out %lld", cnt, (long long)max_out);
    return LF2MI_OK;
}

int lf2mi_bin_positions(const double *pos, int64_t np, int N, int32_t *keys, int32_t *counts, void *stream) {
    if (np < 0 || (np > 0 && !pos) || N < 1 || N > 1290) return fail(LF2MI_EINVAL, "lf2mi_bin_positions: bad argument");
    hipStream_t st = as_stream(stream);
    if (counts)
        if (int rc = fill_int(counts, (int64_t)N * N * N, 0, st)) return rc;
    if (np == 0 || (!keys && !counts)) return LF2MI_OK;
    LF2MI_PROF("bin_keys", st);
    bin_keys_kernel<<<grid1(np), 256, 0, st>>>(np, pos, N, keys, counts);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

int lf2mi_parts_cell_counts(lf2mi_parts p, int32_t *counts, void *stream) {
    if (!p || !counts) return fail(LF2MI_EINVAL, "null argument");
    hipStream_t st = as_stream(stream);
    const int64_t nc = (int64_t)p->N * p->N * p->N;
    if (int rc = fill_int(counts, nc, 0, st)) return rc;
    const int32_t *cell = p->dtype == LF2MI_F32 ? p->b32[p->cur].cell : p->b64.cell;
    if (p->np > 0) {
        cell_count_kernel<<<grid1(p->np), 256, 0, st>>>(p->np, cell, counts);
        LF2MI_LAUNCHED();
    }
    return LF2MI_OK;
}

int lf2mi_parts_ic_lattice(lf2mi_parts p, int ng, double sigma_cells, double sigma_v, uint64_t seed, void *stream) {
    if (!p || ng < 1) return fail(LF2MI_EINVAL, "bad argument");
    IcArgs A{};
    A.kind = 0;
    A.np = (int64_t)ng * ng * ng;
    A.ng = ng;
    A.sigma_pos = sigma_cells / p->N;
    A.sigma_v = sigma_v;
    A.seed = seed;
    return ic_common(p, A, as_stream(stream));
}

int lf2mi_parts_ic_uniform(lf2mi_parts p, int64_t np, double sigma_v, uint64_t seed, void *stream) {
    if (!p || np < 0) return fail(LF2MI_EINVAL, "bad argument");
    IcArgs A{};
    A.kind = 1;
    A.np = np;
    A.sigma_v = sigma_v;
    A.seed = seed;
    return ic_common(p, A, as_stream(stream));
}

int lf2mi_parts_ic_blobs(lf2mi_parts p, int64_t np, int nblobs, double alpha, double sigma_cells, double frac_uniform,
                         double sigma_v, uint64_t seed, void *stream) {
    if (!p || np < 0 || nblobs < 1 || !(alpha > 0) || frac_uniform < 0 || frac_uniform > 1)
        return fail(LF2MI_EINVAL, "bad argument");
    const int64_t ncl = np - (int64_t)llround(frac_uniform * (double)np);
    // Pareto(alpha) occupancies w = u^(-1/alpha), counts by largest remainder
    std::vector<double> w((size_t)nblobs), cen((size_t)nblobs * 3);
    double ws = 0.0;
    for (int b = 0; b < nblobs; ++b) {
        double u = urand(seed ^ 0xB10Bull, (uint64_t)b, 0);
        if (u < 1e-300) u = 1e-300;
        w[(size_t)b] = pow(u, -1.0 / alpha);
        ws += w[(size_t)b];
        for (int d = 0; d < 3; ++d) cen[(size_t)b * 3 + d] = urand(seed ^ 0xCE47ull, (uint64_t)b, (uint32_t)d);
    }
    std::vector<int64_t> cnt((size_t)nblobs);
    std::vector<std::pair<double, int>> rem((size_t)nblobs);
    int64_t assigned = 0;
    for (int b = 0; b < nblobs; ++b) {
        const double exact = w[(size_t)b] / ws * (double)ncl;
        cnt[(size_t)b] = (int64_t)floor(exact);
        assigned += cnt[(size_t)b];
        rem[(size_t)b] = {exact - floor(exact), b};
    }
    std::sort(rem.begin(), rem.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
        return a.first > b.first || (a.first == b.first && a.second < b.second);
    });
    for (int64_t k = 0; k < ncl - assigned; ++k) cnt[(size_t)rem[(size_t)(k % nblobs)].second]++;
    std::vector<int64_t> bst((size_t)nblobs + 1, 0);
    for (int b = 0; b < nblobs; ++b) bst[(size_t)b + 1] = bst[(size_t)b] + cnt[(size_t)b];
    hipStream_t st = as_stream(stream);
    double *dcen = nullptr;
    int64_t *dbst = nullptr;
    LF2MI_HIP(hipMalloc((void **)&dcen, cen.size() * sizeof(double)));
    LF2MI_HIP(hipMalloc((void **)&dbst, bst.size() * sizeof(int64_t)));
    LF2MI_HIP(hipMemcpy(dcen, cen.data(), cen.size() * sizeof(double), hipMemcpyHostToDevice));
    LF2MI_HIP(hipMemcpy(dbst, bst.data(), bst.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    IcArgs A{};
    A.kind = 2;
    A.np = np;
    A.sigma_pos = sigma_cells / p->N;
    A.sigma_v = sigma_v;
    A.seed = seed;
    A.nblobs = nblobs;
    A.nclustered = ncl;
    A.centres = dcen;
    A.bstart = dbst;
    const int rc = ic_common(p, A, st);
    hipStreamSynchronize(st);
    hipFree(dcen);
    hipFree(dbst);
    return rc;
}

}  // extern "C"

// overwrite: the native scalar deposit on the tiles path stores every interior node instead of
// adding to it (the PM step's deposit into a field it would otherwise clear first)
static int deposit_impl(lf2mi_parts p, int kind, void *field, int halo, int target, int acc, bool overwrite,
                        void *stream) {
    if (!p || !field) return fail(LF2MI_EINVAL, "lf2mi_deposit: null argument");
    if (kind < LF2MI_PROJ_SCALAR || kind > LF2MI_PROJ_SYMTENSOR_NGP)
        return fail(LF2MI_EINVAL, "bad projection kind %d", kind);
    if (halo < 1) return fail(LF2MI_EINVAL, "the field has to have at least a halo of 1");
    if (target != LF2MI_DEP_UNFOLDED && target != LF2MI_DEP_PERIODIC) return fail(LF2MI_EINVAL, "bad target %d", target);
    if (acc != LF2MI_ACC_NATIVE && acc != LF2MI_ACC_FIXED64) return fail(LF2MI_EINVAL, "bad accumulation mode %d", acc);
    if (overwrite && target != LF2MI_DEP_PERIODIC) return fail(LF2MI_EINVAL, "an overwriting deposit is periodic");
    hipStream_t st = as_stream(stream);
    const int N = p->N;
    const double nv = (double)N * N * N;
    if (overwrite) {
        // the field's previous contents are replaced: the FP32 tiles kernels store every
        // interior node, so only the ghosts need clearing; the other paths add into a
        // cleared field
        static const int comps[5] = {1, 3, 6, 3, 6};
        const int C = comps[kind];
        const int64_t G = N + 2 * halo;
        const bool f32 = p->dtype == LF2MI_F32;
        const bool pmass = p->mass_src && (f32 ? p->b32[p->cur].mass != nullptr : p->b64.mass != nullptr);
        const bool stores = f32 && acc == LF2MI_ACC_NATIVE && p->np > 0 &&
                            (tiles_path(p, kind, target, pmass) || tiles_c_path(p, kind, target, pmass));
        if (stores) {
            if (int rc = zero_ghosts(field, N, halo, C, p->dtype, st)) return rc;
        } else {
            if (int rc = lf2mi_zero(field, G * G * G * C, p->dtype, stream)) return rc;
            overwrite = false;
        }
    }
    if (p->np == 0) return LF2MI_OK;
    if (p->dtype == LF2MI_F64) {
        if (acc != LF2MI_ACC_NATIVE) return fail(LF2MI_EUNSUPPORTED, "fixed-point deposit is an F32 option");
        Dep64Args A{};
        A.s = p->b64;
        A.np = p->np;
        A.N = N;
        A.mass_nv = p->mass_global * nv;
        A.field = (double *)field;
        A.h = halo;
        A.tgt = target;
        if (p->mass_src && p->b64.mass) A.mass_nv = nv;
        else A.s.mass = nullptr;
        if (kind != LF2MI_PROJ_SCALAR)
            if (int rc = vector_source(p, p->b64.pl, A.s.vel, st)) return rc;
        const unsigned g = grid1(p->np);
        static const char *const names64[5] = {"deposit_f64_scalar", "deposit_f64_vector", "deposit_f64_vecvec",
                                               "deposit_f64_vector_ngp", "deposit_f64_symtensor_ngp"};
        LF2MI_PROF(names64[kind], st);
        switch (kind) {
            case 0: deposit_f64_kernel<0><<<g, 256, 0, st>>>(A); break;
            case 1: deposit_f64_kernel<1><<<g, 256, 0, st>>>(A); break;
            case 2: deposit_f64_kernel<2><<<g, 256, 0, st>>>(A); break;
            case 3: deposit_f64_kernel<3><<<g, 256, 0, st>>>(A); break;
            default: deposit_f64_kernel<4><<<g, 256, 0, st>>>(A); break;
        }
        LF2MI_LAUNCHED();
        return LF2MI_OK;
    }
    DepArgs A{};
    A.s = p->b32[p->cur];
    A.start = p->bstart;
    A.istart = p->istart;
    A.item_bucket = p->item_bucket;
    A.g = p->g;
    // per-particle masses only when chosen (lf2mi_parts_mass_from)
    const bool pm = p->mass_src && A.s.mass;
    if (!pm) A.s.mass = nullptr;
    A.mass_nv = (float)(pm ? nv : p->mass_global * nv);
    if (kind != LF2MI_PROJ_SCALAR)
        if (int rc = vector_source(p, A.s.pl, A.s.vel, st)) return rc;
    A.field = (float *)field;
    A.h = halo;
    A.tgt = target;
    const unsigned g = (unsigned)item_grid(p);
    if (acc == LF2MI_ACC_FIXED64) {
        static const int comps[5] = {1, 3, 6, 3, 6};
        const int C = comps[kind];
        const int64_t G = N + 2 * halo;
        const int64_t n = G * G * G * C;
        if (p->fx_n < n) {
            if (p->fx) hipFree(p->fx);
            p->fx = nullptr;
            p->fx_n = 0;
            LF2MI_HIP(hipMalloc((void **)&p->fx, n * sizeof(unsigned long long)));
            p->fx_n = n;
        }
        // cleared by a kernel (a memset node would misbehave in a captured graph)
        fill_u64_kernel<<<(unsigned)std::min<int64_t>(grid1(n), 65536), 256, 0, st>>>(p->fx, n);
        LF2MI_LAUNCHED();
        A.fx = p->fx;
        FxScales S{};
        if (kind == LF2MI_PROJ_SCALAR && !pm) {
            // one global mass: dimensionless weights in 2^-32 units
            const double mref = A.mass_nv;
            A.fxs[0] = 4294967296.0 / mref;
            S.inv[0] = mref / 4294967296.0;
        } else {
            // per component a power-of-two scale 2^(62 - e) with np max|q| max|v_a| max|v_b| < 2^e
            // (the scalar kind with per-particle masses: np max|q| < 2^e): no node (nor any
            // partial sum) can leave the int64 range, the quantum is 2^-62 of that bound, and
            // the integer sums do not depend on the order of the particles
            unsigned int *dmax = nullptr, hmax[4] = {0u, 0u, 0u, 0u};
            LF2MI_HIP(hipMallocAsync((void **)&dmax, 4 * sizeof(unsigned int), st));
            fill_int_kernel<<<1, 64, 0, st>>>(reinterpret_cast<int *>(dmax), 4, 0);
            dep_max_kernel<<<(unsigned)std::min<int64_t>(grid1(p->np), DMAX_BLOCKS), 256, 0, st>>>(p->np, A.s, A.mass_nv,
                                                                                          dmax);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(hmax, dmax, sizeof hmax, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            hipFreeAsync(dmax, st);
            if (e != hipSuccess) return hip_fail(e, "fixed-point bound", __FILE__, __LINE__);
            float mx[4];
            memcpy(mx, hmax, sizeof mx);
            static const int va[5][6] = {{-1}, {0, 1, 2}, {0, 0, 0, 1, 1, 2}, {0, 1, 2}, {0, 0, 0, 1, 1, 2}};
            static const int vb[5][6] = {{-1}, {-1, -1, -1}, {0, 1, 2, 1, 2, 2}, {-1, -1, -1}, {0, 1, 2, 1, 2, 2}};
            for (int c = 0; c < C; ++c) {
                double bound = (double)p->np * (double)mx[0];
                if (va[kind][c] >= 0) bound *= (double)mx[1 + va[kind][c]];
                if (vb[kind][c] >= 0) bound *= (double)mx[1 + vb[kind][c]];
                double sc = 1.0;
                if (bound > 0.0 && bound < 1e300) {
                    int ex;
                    frexp(bound, &ex);
                    sc = ldexp(1.0, 62 - ex);
                }
                A.fxs[c] = sc;
                S.inv[c] = 1.0 / sc;
            }
        }
        {
            static const char *const namesfx[5] = {"deposit_f32_fixed64", "deposit_f32_vector_fixed64",
                                                   "deposit_f32_vecvec_fixed64", "deposit_f32_vector_ngp_fixed64",
                                                   "deposit_f32_symtensor_ngp_fixed64"};
            LF2MI_PROF(namesfx[kind], st);
            const bool b8 = p->g.B == 8;
            if (tiles_path(p, kind, target, pm)) {
                // the fixed-point buffer starts cleared, so its interior can be stored
                if (int rc = launch_tiles_dense<unsigned long long>(p, A, true, st)) return rc;
            } else {
                switch (kind) {
                    case 0: launch_deposit<0, unsigned long long>(A, b8, g, st); break;
                    case 1: launch_deposit<1, unsigned long long>(A, b8, g, st); break;
                    case 2: launch_deposit<2, unsigned long long>(A, b8, g, st); break;
                    case 3: launch_deposit<3, unsigned long long>(A, b8, g, st); break;
                    default: launch_deposit<4, unsigned long long>(A, b8, g, st); break;
                }
            }
            LF2MI_LAUNCHED();
        }
        {
            LF2MI_PROF("fixed_to_field", st);
            if (C == 1)
                fixed_to_field_kernel<<<(unsigned)std::min<int64_t>(grid1(n), 65536), 256, 0, st>>>(
                    p->fx, (float *)field, n, S.inv[0]);
            else
                fixed_to_field_c_kernel<<<(unsigned)std::min<int64_t>(grid1(n), 65536), 256, 0, st>>>(
                    p->fx, (float *)field, n, C, S);
            LF2MI_LAUNCHED();
        }
        return LF2MI_OK;
    }
    static const char *const names32[5] = {"deposit_f32_scalar", "deposit_f32_vector", "deposit_f32_vecvec",
                                           "deposit_f32_vector_ngp", "deposit_f32_symtensor_ngp"};
    LF2MI_PROF(names32[kind], st);
    const bool b8 = p->g.B == 8;
    if (tiles_path(p, kind, target, pm)) return launch_tiles_dense<float>(p, A, overwrite, st);
    if (tiles_c_path(p, kind, target, pm)) {
        switch (kind) {
            case 0: return launch_tiles_c<0>(p, A, pm, overwrite, st);
            case 1: return launch_tiles_c<1>(p, A, pm, overwrite, st);
            case 2: return launch_tiles_c<2>(p, A, pm, overwrite, st);
            case 3: return launch_tiles_c<3>(p, A, pm, overwrite, st);
            default: return launch_tiles_c<4>(p, A, pm, overwrite, st);
        }
    }
    switch (kind) {
        case 0: launch_deposit<0, float>(A, b8, g, st); break;
        case 1: launch_deposit<1, float>(A, b8, g, st); break;
        case 2: launch_deposit<2, float>(A, b8, g, st); break;
        case 3: launch_deposit<3, float>(A, b8, g, st); break;
        default: launch_deposit<4, float>(A, b8, g, st); break;
    }
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

extern "C" {

int lf2mi_deposit(lf2mi_parts p, int kind, void *field, int halo, int target, int acc, void *stream) {
    if (target == LF2MI_DEP_PERIODIC_SET)
        return deposit_impl(p, kind, field, halo, LF2MI_DEP_PERIODIC, acc, true, stream);
    return deposit_impl(p, kind, field, halo, target, acc, false, stream);
}

int lf2mi_gather_kick_drift(lf2mi_parts p, const void *phi, int halo, double dt_kick, double dt_drift, double *vmax2,
                            void *stream) {
    if (!p || halo < 0) return fail(LF2MI_EINVAL, "lf2mi_gather_kick_drift: bad argument");
    if (!phi && dt_kick != 0.0) return fail(LF2MI_EINVAL, "lf2mi_gather_kick_drift: a kick needs the potential field");
    hipStream_t st = as_stream(stream);
    if (p->np == 0) {
        if (vmax2) *vmax2 = 0.0;
        return LF2MI_OK;
    }
    if (p->dtype == LF2MI_F64) {
        Gkd64Args A{};
        A.s = p->b64;
        A.np = p->np;
        A.N = p->N;
        A.h = halo;
        A.phi = (const double *)phi;
        A.dt_kick = dt_kick;
        A.dt_drift = dt_drift;
        A.vmax2 = vmax2 ? p->vmax_f64 : nullptr;
        if (vmax2) LF2MI_HIP(hipMemsetAsync(p->vmax_f64, 0, sizeof(unsigned long long), st));
        LF2MI_PROF("gkd_f64", st);
        gkd_f64_kernel<<<grid1(p->np), 256, 0, st>>>(A);
        LF2MI_LAUNCHED();
        if (vmax2) {
            unsigned long long bits = 0;
            LF2MI_HIP(hipMemcpyAsync(&bits, p->vmax_f64, sizeof(bits), hipMemcpyDeviceToHost, st));
            LF2MI_HIP(hipStreamSynchronize(st));
            double v;
            memcpy(&v, &bits, sizeof(v));
            *vmax2 = v;
        }
        return LF2MI_OK;
    }
    if (int rc = fill_int(p->bcount, p->g.nbk, 0, st)) return rc;
    GkdArgs A{};
    A.s = p->b32[p->cur];
    A.start = p->bstart;
    A.istart = p->istart;
    A.item_bucket = p->item_bucket;
    A.g = p->g;
    A.phi = (const float *)phi;
    A.h = halo;
    A.dt_kick = (float)dt_kick;
    A.dt_drift = dt_drift;
    A.next_counts = p->bcount;
    A.vmax2 = vmax2 ? p->vmax_f32 : nullptr;
    A.chunked = phi != nullptr;
    if (vmax2) LF2MI_HIP(hipMemsetAsync(p->vmax_f32, 0, sizeof(unsigned int), st));
    const bool b8 = p->g.B == 8;
    if (A.chunked) {
        if (int rc = fill_int(p->nchunk, 2, 0, st)) return rc;
        LF2MI_PROF("gkd_chunk_list", st);
        const int64_t rows = (int64_t)p->g.NB * p->g.NB;
        gkd_chunks_kernel<<<grid1(rows, 64), 64, 0, st>>>(p->bstart, p->g, p->chunks, p->nchunk);
        LF2MI_LAUNCHED();
    }
    {
    LF2MI_PROF("gkd_f32", st);
    if (b8) gkd_f32_kernel<8><<<(unsigned)item_grid(p), DEP_WG, 0, st>>>(A);
    else gkd_f32_kernel<0><<<(unsigned)item_grid(p), DEP_WG, 0, st>>>(A);
    LF2MI_LAUNCHED();
    }
    if (A.chunked) {
        // persistent: a few workgroups per CU walk the chunk list
        LF2MI_PROF("gkd_f32_chunks", st);
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(p->g.nbk, 2048));
        if (b8) gkd_chunk_kernel<8><<<grid, DEP_WG, 0, st>>>(A, p->chunks, p->nchunk);
        else gkd_chunk_kernel<0><<<grid, DEP_WG, 0, st>>>(A, p->chunks, p->nchunk);
        LF2MI_LAUNCHED();
    }
    if (int rc = sort_into_other(p, p->cur, p->np, st)) return rc;
    if (vmax2) {
        unsigned int bits = 0;
        LF2MI_HIP(hipMemcpyAsync(&bits, p->vmax_f32, sizeof(bits), hipMemcpyDeviceToHost, st));
        LF2MI_HIP(hipStreamSynchronize(st));
        float v;
        memcpy(&v, &bits, sizeof(v));
        *vmax2 = v;
    }
    return LF2MI_OK;
}

int lf2mi_parts_reduce(lf2mi_parts p, int n, const int *quantity, const int *reduce_type, double *out, void *stream) {
    if (!p || n < 0 || n > RED_Q || (n > 0 && (!quantity || !reduce_type || !out)))
        return fail(LF2MI_EINVAL, "lf2mi_parts_reduce: bad argument (at most %d quantities)", RED_Q);
    if (n == 0) return LF2MI_OK;
    RedArgs A{};
    A.np = p->np;
    A.nq = n;
    for (int k = 0; k < RED_Q; ++k) {
        A.q[k] = 0;
        A.op[k] = RED_NONE;
    }
    for (int k = 0; k < n; ++k) {
        if (quantity[k] < LF2MI_PQ_V2 || quantity[k] > LF2MI_PQ_EKIN)
            return fail(LF2MI_EINVAL, "lf2mi_parts_reduce: unknown quantity %d", quantity[k]);
        const int rt = reduce_type[k];
        A.q[k] = quantity[k];
        A.op[k] = (rt & (LF2MI_RED_SUM | LF2MI_RED_SUM_LOCAL))   ? RED_OP_SUM
                  : (rt & (LF2MI_RED_MIN | LF2MI_RED_MIN_LOCAL)) ? RED_OP_MIN
                  : (rt & (LF2MI_RED_MAX | LF2MI_RED_MAX_LOCAL)) ? RED_OP_MAX
                                                                 : RED_NONE;
    }
    A.mass = p->mass_global;
    hipStream_t st = as_stream(stream);
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (p->np + RED_WG - 1) / RED_WG));
    double *res = p->red + (int64_t)RED_BLOCKS * RED_Q;
    LF2MI_PROF("parts_reduce", st);
    if (p->dtype == LF2MI_F32) {
        const Soa32 &s = p->b32[p->cur];
        for (int d = 0; d < 3; ++d) A.vel32[d] = s.vel[d];
        A.mass32 = p->mass_src ? s.mass : nullptr;
        parts_reduce_kernel<false><<<nb, RED_WG, 0, st>>>(A, p->red);
    } else {
        for (int d = 0; d < 3; ++d) A.vel64[d] = p->b64.vel[d];
        A.mass64 = p->mass_src ? p->b64.mass : nullptr;
        parts_reduce_kernel<true><<<nb, RED_WG, 0, st>>>(A, p->red);
    }
    LF2MI_LAUNCHED();
    parts_reduce_final_kernel<<<1, RED_Q, 0, st>>>(A, p->red, nb, res);
    LF2MI_LAUNCHED();
    double host[RED_Q];
    LF2MI_HIP(hipMemcpyAsync(host, res, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    LF2MI_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < n; ++k)
        if (A.op[k] != RED_NONE) out[k] = host[k];
    return LF2MI_OK;
}

int lf2mi_pm_step(lf2mi_parts p, lf2mi_fft_plan plan, void *rho, void *phi, int halo, double dt_kick, double dt_drift,
                  int acc, void *stream) {
    if (!p || !plan || !rho || !phi) return fail(LF2MI_EINVAL, "lf2mi_pm_step: null argument");
    const int N = p->N;
    // rho replaced by the deposit (the FP32 tiles kernels store its interior)
    if (int rc = deposit_impl(p, LF2MI_PROJ_SCALAR, rho, halo, LF2MI_DEP_PERIODIC, acc, true, stream)) return rc;
    // box volume 1: the mean density is the total mass
    const double scale = 1.0 / ((double)N * N * N);
    if (int rc = lf2mi_poisson_solve(plan, rho, phi, p->total_mass, scale, stream)) return rc;
    return lf2mi_gather_kick_drift(p, phi, halo, dt_kick, dt_drift, nullptr, stream);
}

}  // extern "C"
