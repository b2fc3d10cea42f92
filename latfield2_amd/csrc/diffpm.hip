// Differentiable PM on one gfx950 device: nsteps of a kick-drift leapfrog of the PM
// step, with hand-written adjoint kernels.  No counterpart in the reference; it
// differentiates the reference's forward pieces:
//   deposit  rho = sum_p q W(X - X_p), q = m N^3   (scalarProjectionCIC_project,
//            particles/projections.hpp:76-300, periodic)
//   Poisson  phi = G rho, G = -K^-2, zero mode 0   (examples/poissonSolver.cpp:99-111)
//   force    F_d = -D_d phi, D_d the central difference (N/2)(f(n+e_d) - f(n-e_d))
//   gather   a_p = sum_corners W F(corner)          (updateVel, Particles.hpp:825-828)
//   leapfrog v += dt a(x);  x += dt v, wrapped into the box (moveParticles)
//   loss     L = sum_cells (rho(x_final) - target)^2
// Adjoint (x in box units, X = x N in cell units; lambda_x = dL/dx, lambda_v = dL/dv):
//   final state          lambda_x = q N grad_X interp(2 (rho - target))(X_p)
//   per step, backwards  lambda_v += dt lambda_x;  mu = dt lambda_v
//                        lambda_x += N sum_d mu_d grad_X interp(F_d)(X_p)   (gather weights)
//                        F-bar_d = sum_p mu_{p,d} W(X - X_p)  (the gather's transpose is a deposit)
//                        phi-bar = sum_d D_d F-bar_d          (D_d is antisymmetric)
//                        rho-bar = G phi-bar                  (G is symmetric: the same FFT solve)
//                        lambda_x += q N grad_X interp(rho-bar)(X_p)   (deposit weights)
// Particles keep their order (no bucketing) so gradients stay attached to them.  The
// starting state of every step is checkpointed in HBM and the fields are recomputed
// on the way back.  Positions are kept as (cell, offset in cell units) and re-anchored
// in FP64 after each drift, as in the PM step.  T = float or double selects the
// precision of the whole path (fields, FFTs, offsets, velocities, adjoints).
#include <algorithm>
#include <cmath>
#include <new>

#include "common.hpp"

namespace lf2mi {

namespace {

template <typename T> struct DpState {
    int N;
    int64_t np;
    int32_t *cell;
    T *off[3];
    T *vel[3];
};

constexpr int DP_WG = 256;

inline unsigned dp_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + DP_WG - 1) / DP_WG, 1 << 16)); }

__device__ __forceinline__ int wrapn(int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); }

__device__ __forceinline__ void decode(int cell, int N, int c[3]) {
    c[0] = cell % N;
    c[1] = (cell / N) % N;
    c[2] = cell / (N * N);
}

__device__ __forceinline__ double add_nc(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double mul_nc(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

template <typename T> __device__ __forceinline__ T below_one();
template <> __device__ __forceinline__ float below_one<float>() { return 0x1.fffffep-1f; }
template <> __device__ __forceinline__ double below_one<double>() { return 0x1.fffffffffffffp-1; }

// X (cell units, FP64) -> (cell, offset) with periodic wrap
template <typename T> __device__ __forceinline__ void anchor(double X, int N, int &c, T &o) {
    if (X < 0.0 || X >= (double)N) X -= (double)N * floor(X / (double)N);
    const double fl = floor(X);
    c = (int)fl;
    if (c >= N) c = N - 1;
    if (c < 0) c = 0;
    o = (T)(X - fl);
    if (!(o < (T)1)) o = below_one<T>();
}

template <typename T> __device__ __forceinline__ T wsel(int i, T u) { return i ? u : (T)1 - u; }

template <typename T, typename I>
__global__ void dp_init_kernel(DpState<T> s, const I *__restrict__ pos, const I *__restrict__ vel) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            T o;
            anchor<T>(mul_nc((double)pos[3 * p + d], (double)s.N), s.N, c[d], o);
            s.off[d][p] = o;
            s.vel[d][p] = vel ? (T)vel[3 * p + d] : (T)0;
        }
        s.cell[p] = c[0] + s.N * (c[1] + s.N * c[2]);
    }
}

template <typename T, typename I> __global__ void dp_state_out_kernel(DpState<T> s, I *__restrict__ pos, I *__restrict__ vel) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        decode(s.cell[p], s.N, c);
        for (int d = 0; d < 3; ++d) {
            if (pos) pos[3 * p + d] = (I)(add_nc((double)c[d], (double)s.off[d][p]) / (double)s.N);
            if (vel) vel[3 * p + d] = (I)s.vel[d][p];
        }
    }
}

template <typename T> __global__ void dp_density_kernel(DpState<T> s, T *__restrict__ rho, T q) {
    const int N = s.N;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        decode(s.cell[p], N, c);
        const T u[3] = {s.off[0][p], s.off[1][p], s.off[2][p]};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const T w = wsel(i, u[0]) * wsel(j, u[1]) * wsel(k, u[2]);
                    atomicAdd(&rho[site_index(N, 1, wrapn(c[0] + i, N), wrapn(c[1] + j, N), wrapn(c[2] + k, N))], q * w);
                }
    }
}

// F = -D phi at the 8 corners of a cell from the 32 surrounding values: Fc[d][k][j][i]
template <typename T>
__device__ __forceinline__ void corner_forces(const T *__restrict__ phi, int N, const int c[3], T Fc[3][2][2][2]) {
    int xs[4], ys[4], zs[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        xs[a] = wrapn(c[0] - 1 + a, N);
        ys[a] = wrapn(c[1] - 1 + a, N);
        zs[a] = wrapn(c[2] - 1 + a, N);
    }
    const T hn = (T)0.5 * (T)N;
    auto P = [&](int a, int b, int e) { return phi[site_index(N, 1, xs[a], ys[b], zs[e])]; };
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                Fc[0][k][j][i] = -(P(i + 2, j + 1, k + 1) - P(i, j + 1, k + 1)) * hn;
                Fc[1][k][j][i] = -(P(i + 1, j + 2, k + 1) - P(i + 1, j, k + 1)) * hn;
                Fc[2][k][j][i] = -(P(i + 1, j + 1, k + 2) - P(i + 1, j + 1, k)) * hn;
            }
}

// v += dt a(X); X += dt N v; re-anchor
template <typename T> __global__ void dp_step_kernel(DpState<T> s, const T *__restrict__ phi, T dt) {
    const int N = s.N;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        decode(s.cell[p], N, c);
        const T u[3] = {s.off[0][p], s.off[1][p], s.off[2][p]};
        T Fc[3][2][2][2];
        corner_forces(phi, N, c, Fc);
        T a[3] = {(T)0, (T)0, (T)0};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const T w = wsel(i, u[0]) * wsel(j, u[1]) * wsel(k, u[2]);
#pragma unroll
                    for (int d = 0; d < 3; ++d) a[d] += w * Fc[d][k][j][i];
                }
        int cn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T v = s.vel[d][p] + dt * a[d];
            s.vel[d][p] = v;
            const double X = add_nc(add_nc((double)c[d], (double)u[d]), mul_nc(mul_nc((double)dt, (double)v), (double)N));
            T o;
            anchor<T>(X, N, cn[d], o);
            s.off[d][p] = o;
        }
        s.cell[p] = cn[0] + N * (cn[1] + N * cn[2]);
    }
}

// resid = 2 (rho - target) on the interior; loss += sum (rho - target)^2 in FP64;
// rho_out (x fastest, may be null) receives the interior of rho
template <typename T, typename I>
__global__ void dp_resid_kernel(const T *__restrict__ rho, const I *__restrict__ target, T *__restrict__ resid,
                                I *__restrict__ rho_out, int N, double *__restrict__ loss) {
    __shared__ double red[DP_WG];
    const int64_t n = (int64_t)N * N * N;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(t % N), y = (int)((t / N) % N), z = (int)(t / ((int64_t)N * N));
        const int64_t i = site_index(N, 1, x, y, z);
        const T r = rho[i] - (target ? (T)target[t] : (T)0);
        resid[i] = (T)2 * r;
        if (rho_out) rho_out[t] = (I)rho[i];
        acc += (double)r * (double)r;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = DP_WG / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// gradient (per cell unit) of the CIC interpolant of g at (c, u)
template <typename T>
__device__ __forceinline__ void interp_grad(const T *__restrict__ g, int N, const int c[3], const T u[3], T gr[3]) {
    gr[0] = gr[1] = gr[2] = (T)0;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const T v = g[site_index(N, 1, wrapn(c[0] + i, N), wrapn(c[1] + j, N), wrapn(c[2] + k, N))];
                const T wx = wsel(i, u[0]), wy = wsel(j, u[1]), wz = wsel(k, u[2]);
                gr[0] += (i ? v : -v) * wy * wz;
                gr[1] += wx * (j ? v : -v) * wz;
                gr[2] += wx * wy * (k ? v : -v);
            }
}

// lambda_x (AoS, np*3) = (init ? 0 : lambda_x) + qN grad_X interp(g)
template <typename T>
__global__ void dp_pos_grad_kernel(DpState<T> s, const T *__restrict__ g, T qN, T *__restrict__ lx, int init) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        decode(s.cell[p], s.N, c);
        const T u[3] = {s.off[0][p], s.off[1][p], s.off[2][p]};
        T gr[3];
        interp_grad(g, s.N, c, u, gr);
#pragma unroll
        for (int d = 0; d < 3; ++d) lx[3 * p + d] = (init ? (T)0 : lx[3 * p + d]) + qN * gr[d];
    }
}

// particle part of one backward step; fbar is a 3-component AoS field
template <typename T>
__global__ void dp_adj_particles_kernel(DpState<T> s, const T *__restrict__ phi, T dt, T *__restrict__ lx,
                                        T *__restrict__ lv, T *__restrict__ fbar) {
    const int N = s.N;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.np; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3];
        decode(s.cell[p], N, c);
        const T u[3] = {s.off[0][p], s.off[1][p], s.off[2][p]};
        T mu[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T v = lv[3 * p + d] + dt * lx[3 * p + d];
            lv[3 * p + d] = v;
            mu[d] = dt * v;
        }
        T Fc[3][2][2][2];
        corner_forces(phi, N, c, Fc);
        T add[3] = {(T)0, (T)0, (T)0};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const T wx = wsel(i, u[0]), wy = wsel(j, u[1]), wz = wsel(k, u[2]);
                    const T mf = mu[0] * Fc[0][k][j][i] + mu[1] * Fc[1][k][j][i] + mu[2] * Fc[2][k][j][i];
                    add[0] += (i ? mf : -mf) * wy * wz;
                    add[1] += wx * (j ? mf : -mf) * wz;
                    add[2] += wx * wy * (k ? mf : -mf);
                    const int64_t node = site_index(N, 1, wrapn(c[0] + i, N), wrapn(c[1] + j, N), wrapn(c[2] + k, N));
                    const T w = wx * wy * wz;
#pragma unroll
                    for (int d = 0; d < 3; ++d) atomicAdd(&fbar[3 * node + d], mu[d] * w);
                }
#pragma unroll
        for (int d = 0; d < 3; ++d) lx[3 * p + d] += (T)N * add[d];
    }
}

// phi-bar = sum_d D_d F-bar_d on the interior, periodic
template <typename T> __global__ void dp_divergence_kernel(const T *__restrict__ fbar, T *__restrict__ phib, int N) {
    const int64_t n = (int64_t)N * N * N;
    const T hn = (T)0.5 * (T)N;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(t % N), y = (int)((t / N) % N), z = (int)(t / ((int64_t)N * N));
        auto F = [&](int xx, int yy, int zz, int d) { return fbar[3 * site_index(N, 1, wrapn(xx, N), wrapn(yy, N), wrapn(zz, N)) + d]; };
        phib[site_index(N, 1, x, y, z)] = (F(x + 1, y, z, 0) - F(x - 1, y, z, 0)) * hn +
                                          (F(x, y + 1, z, 1) - F(x, y - 1, z, 1)) * hn +
                                          (F(x, y, z + 1, 2) - F(x, y, z - 1, 2)) * hn;
    }
}

template <typename T> __global__ void dp_fill_kernel(T *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (T)0;
}

template <typename T, typename I> __global__ void dp_convert_kernel(const T *__restrict__ src, I *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = (I)src[i];
}

}  // namespace
}  // namespace lf2mi

struct lf2mi_diffpm_s {
    int N, nsteps, dtype, io_dtype;
    int64_t np, G3;
    double mass, dt;
    lf2mi_fft_plan plan;
    size_t tsize;   // sizeof(T)
    void *ck_mem;   // nsteps + 1 particle states (cell, off[3], vel[3])
    void *rho, *phi, *resid, *fbar, *phib, *rhob;  // (N+2)^3 fields of T (fbar: 3 components)
    void *lx, *lv;  // adjoints, np*3 of T
    double *loss_dev;
    bool forward_done;
};

namespace lf2mi {
namespace {

size_t dp_state_bytes(int64_t np, size_t tsize) { return (((size_t)np * 4 + 255) & ~(size_t)255) + (size_t)np * 6 * tsize; }

template <typename T> DpState<T> dp_view(const lf2mi_diffpm_s *m, int k) {
    char *base = (char *)m->ck_mem + dp_state_bytes(m->np, sizeof(T)) * (size_t)k;
    DpState<T> s;
    s.N = m->N;
    s.np = m->np;
    s.cell = (int32_t *)base;
    T *f = (T *)(base + (((size_t)m->np * 4 + 255) & ~(size_t)255));
    for (int d = 0; d < 3; ++d) s.off[d] = f + d * m->np;
    for (int d = 0; d < 3; ++d) s.vel[d] = f + (3 + d) * m->np;
    return s;
}

int dp_malloc(void **p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc(diffpm)", __FILE__, __LINE__);
    }
    return LF2MI_OK;
}

template <typename T> int dp_forces(lf2mi_diffpm_s *m, const DpState<T> &s, hipStream_t st) {
    dp_fill_kernel<T><<<dp_grid(m->G3), DP_WG, 0, st>>>((T *)m->rho, m->G3);
    LF2MI_LAUNCHED();
    const double n3 = (double)m->N * m->N * m->N;
    dp_density_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(s, (T *)m->rho, (T)(m->mass * n3));
    LF2MI_LAUNCHED();
    return lf2mi_poisson_solve(m->plan, m->rho, m->phi, m->mass * (double)m->np, 1.0 / n3, st);
}

template <typename T, typename I>
int dp_forward(lf2mi_diffpm_s *m, const I *pos0, const I *vel0, const I *target, I *rho_out, double *loss, hipStream_t st) {
    const size_t sb = dp_state_bytes(m->np, sizeof(T));
    dp_init_kernel<T, I><<<dp_grid(m->np), DP_WG, 0, st>>>(dp_view<T>(m, 0), pos0, vel0);
    LF2MI_LAUNCHED();
    for (int s = 0; s < m->nsteps; ++s) {
        const DpState<T> nx = dp_view<T>(m, s + 1);
        LF2MI_HIP(hipMemcpyAsync(nx.cell, dp_view<T>(m, s).cell, sb, hipMemcpyDeviceToDevice, st));
        if (int rc = dp_forces<T>(m, nx, st)) return rc;
        dp_step_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(nx, (const T *)m->phi, (T)m->dt);
        LF2MI_LAUNCHED();
    }
    const DpState<T> fin = dp_view<T>(m, m->nsteps);
    dp_fill_kernel<T><<<dp_grid(m->G3), DP_WG, 0, st>>>((T *)m->rho, m->G3);
    dp_fill_kernel<double><<<1, DP_WG, 0, st>>>(m->loss_dev, 1);
    LF2MI_LAUNCHED();
    const double n3 = (double)m->N * m->N * m->N;
    dp_density_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(fin, (T *)m->rho, (T)(m->mass * n3));
    LF2MI_LAUNCHED();
    dp_resid_kernel<T, I><<<std::min<unsigned>(dp_grid((int64_t)n3), 1024), DP_WG, 0, st>>>(
        (const T *)m->rho, target, (T *)m->resid, rho_out, m->N, m->loss_dev);
    LF2MI_LAUNCHED();
    if (loss) {
        LF2MI_HIP(hipMemcpyAsync(loss, m->loss_dev, sizeof(double), hipMemcpyDeviceToHost, st));
        LF2MI_HIP(hipStreamSynchronize(st));
    }
    return LF2MI_OK;
}

template <typename T, typename I> int dp_backward(lf2mi_diffpm_s *m, I *grad_pos, I *grad_vel, hipStream_t st) {
    const double n3 = (double)m->N * m->N * m->N;
    const T qN = (T)(m->mass * n3 * m->N);
    T *lx = (T *)m->lx, *lv = (T *)m->lv;
    dp_fill_kernel<T><<<dp_grid(3 * m->np), DP_WG, 0, st>>>(lv, 3 * m->np);
    LF2MI_LAUNCHED();
    dp_pos_grad_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(dp_view<T>(m, m->nsteps), (const T *)m->resid, qN, lx, 1);
    LF2MI_LAUNCHED();
    for (int s = m->nsteps - 1; s >= 0; --s) {
        const DpState<T> xs = dp_view<T>(m, s);
        if (int rc = dp_forces<T>(m, xs, st)) return rc;  // recompute phi of step s
        dp_fill_kernel<T><<<dp_grid(3 * m->G3), DP_WG, 0, st>>>((T *)m->fbar, 3 * m->G3);
        LF2MI_LAUNCHED();
        dp_adj_particles_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(xs, (const T *)m->phi, (T)m->dt, lx, lv, (T *)m->fbar);
        LF2MI_LAUNCHED();
        dp_divergence_kernel<T><<<dp_grid((int64_t)n3), DP_WG, 0, st>>>((const T *)m->fbar, (T *)m->phib, m->N);
        LF2MI_LAUNCHED();
        if (int rc = lf2mi_poisson_solve(m->plan, m->phib, m->rhob, 0.0, 1.0 / n3, st)) return rc;
        dp_pos_grad_kernel<T><<<dp_grid(m->np), DP_WG, 0, st>>>(xs, (const T *)m->rhob, qN, lx, 0);
        LF2MI_LAUNCHED();
    }
    dp_convert_kernel<T, I><<<dp_grid(3 * m->np), DP_WG, 0, st>>>(lx, grad_pos, 3 * m->np);
    LF2MI_LAUNCHED();
    if (grad_vel) {
        dp_convert_kernel<T, I><<<dp_grid(3 * m->np), DP_WG, 0, st>>>(lv, grad_vel, 3 * m->np);
        LF2MI_LAUNCHED();
    }
    return LF2MI_OK;
}

template <typename T, typename I> int dp_state(lf2mi_diffpm_s *m, int step, I *pos, I *vel, hipStream_t st) {
    dp_state_out_kernel<T, I><<<dp_grid(m->np), DP_WG, 0, st>>>(dp_view<T>(m, step), pos, vel);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

}  // namespace
}  // namespace lf2mi

using namespace lf2mi;

extern "C" {

int lf2mi_diffpm_bytes(lf2mi_diffpm m, int64_t *bytes) {
    if (!m || !bytes) return fail(LF2MI_EINVAL, "lf2mi_diffpm_bytes: null argument");
    int64_t fft = 0;
    if (int rc = lf2mi_fft_plan_bytes(m->plan, &fft)) return rc;
    *bytes = (int64_t)(dp_state_bytes(m->np, m->tsize) * (size_t)(m->nsteps + 1) + 8 * (size_t)m->G3 * m->tsize +
                       6 * (size_t)m->np * m->tsize + sizeof(double)) + fft;
    return LF2MI_OK;
}

int lf2mi_diffpm_create(int N, int64_t np, int nsteps, double mass, double dt, int dtype, lf2mi_diffpm *out) {
    if (!out) return fail(LF2MI_EINVAL, "lf2mi_diffpm_create: null output");
    *out = nullptr;
    if (N < 8 || !is_pow2(N) || np < 1 || np > INT32_MAX || nsteps < 0 || !(mass > 0) || !std::isfinite(dt))
        return fail(LF2MI_EINVAL, "lf2mi_diffpm_create: bad argument (N=%d np=%lld nsteps=%d)", N, (long long)np, nsteps);
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "lf2mi_diffpm_create: dtype must be F32 or F64");
    lf2mi_diffpm_s *m = new (std::nothrow) lf2mi_diffpm_s();
    if (!m) return fail(LF2MI_ENOMEM, "host allocation failed");
    m->N = N;
    m->np = np;
    m->nsteps = nsteps;
    m->mass = mass;
    m->dt = dt;
    m->dtype = dtype;
    m->tsize = dtype == LF2MI_F32 ? 4 : 8;
    m->G3 = (int64_t)(N + 2) * (N + 2) * (N + 2);
    const size_t fb = (size_t)m->G3 * m->tsize;
    int rc = lf2mi_fft_plan_create(N, 1, 1, 0, dtype, &m->plan);
    if (!rc) rc = dp_malloc(&m->ck_mem, dp_state_bytes(np, m->tsize) * (size_t)(nsteps + 1));
    if (!rc) rc = dp_malloc(&m->rho, fb);
    if (!rc) rc = dp_malloc(&m->phi, fb);
    if (!rc) rc = dp_malloc(&m->resid, fb);
    if (!rc) rc = dp_malloc(&m->fbar, 3 * fb);
    if (!rc) rc = dp_malloc(&m->phib, fb);
    if (!rc) rc = dp_malloc(&m->rhob, fb);
    if (!rc) rc = dp_malloc(&m->lx, (size_t)np * 3 * m->tsize);
    if (!rc) rc = dp_malloc(&m->lv, (size_t)np * 3 * m->tsize);
    if (!rc) rc = dp_malloc((void **)&m->loss_dev, sizeof(double));
    if (!rc) {
        // the ghosts of these fields are never written by the kernels here; keep them defined
        for (void *p : {m->rho, m->phi, m->resid, m->phib, m->rhob}) {
            hipError_t e = hipMemset(p, 0, fb);
            if (e != hipSuccess) {
                rc = hip_fail(e, "hipMemset(diffpm)", __FILE__, __LINE__);
                break;
            }
        }
    }
    if (rc) {
        lf2mi_diffpm_destroy(m);
        return rc;
    }
    *out = m;
    return LF2MI_OK;
}

int lf2mi_diffpm_destroy(lf2mi_diffpm m) {
    if (!m) return LF2MI_OK;
    if (m->plan) lf2mi_fft_plan_destroy(m->plan);
    for (void *p : {m->ck_mem, m->rho, m->phi, m->resid, m->fbar, m->phib, m->rhob, m->lx, m->lv, (void *)m->loss_dev})
        if (p) (void)hipFree(p);
    delete m;
    return LF2MI_OK;
}

int lf2mi_diffpm_forward(lf2mi_diffpm m, const void *pos0, const void *vel0, const void *target, int io_dtype,
                         void *rho_out, double *loss, void *stream) {
    if (!m || !pos0) return fail(LF2MI_EINVAL, "lf2mi_diffpm_forward: null argument");
    if (io_dtype != LF2MI_F32 && io_dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "lf2mi_diffpm_forward: io dtype must be F32 or F64");
    hipStream_t st = as_stream(stream);
    LF2MI_PROF("diffpm_forward", st);
    m->forward_done = false;
    int rc;
    if (m->dtype == LF2MI_F32) {
        if (io_dtype == LF2MI_F32)
            rc = dp_forward<float, float>(m, (const float *)pos0, (const float *)vel0, (const float *)target, (float *)rho_out, loss, st);
        else
            rc = dp_forward<float, double>(m, (const double *)pos0, (const double *)vel0, (const double *)target, (double *)rho_out, loss, st);
    } else {
        if (io_dtype == LF2MI_F32)
            rc = dp_forward<double, float>(m, (const float *)pos0, (const float *)vel0, (const float *)target, (float *)rho_out, loss, st);
        else
            rc = dp_forward<double, double>(m, (const double *)pos0, (const double *)vel0, (const double *)target, (double *)rho_out, loss, st);
    }
    if (rc) return rc;
    m->io_dtype = io_dtype;
    m->forward_done = true;
    return LF2MI_OK;
}

int lf2mi_diffpm_backward(lf2mi_diffpm m, void *grad_pos, void *grad_vel, void *stream) {
    if (!m || !grad_pos) return fail(LF2MI_EINVAL, "lf2mi_diffpm_backward: null argument");
    if (!m->forward_done) return fail(LF2MI_EINVAL, "lf2mi_diffpm_backward: no forward pass to differentiate");
    hipStream_t st = as_stream(stream);
    LF2MI_PROF("diffpm_backward", st);
    if (m->dtype == LF2MI_F32)
        return m->io_dtype == LF2MI_F32 ? dp_backward<float, float>(m, (float *)grad_pos, (float *)grad_vel, st)
                                        : dp_backward<float, double>(m, (double *)grad_pos, (double *)grad_vel, st);
    return m->io_dtype == LF2MI_F32 ? dp_backward<double, float>(m, (float *)grad_pos, (float *)grad_vel, st)
                                    : dp_backward<double, double>(m, (double *)grad_pos, (double *)grad_vel, st);
}

int lf2mi_diffpm_state(lf2mi_diffpm m, int step, void *pos, void *vel, int io_dtype, void *stream) {
    if (!m || (!pos && !vel)) return fail(LF2MI_EINVAL, "lf2mi_diffpm_state: null argument");
    if (!m->forward_done) return fail(LF2MI_EINVAL, "lf2mi_diffpm_state: no forward pass");
    if (step < 0 || step > m->nsteps) return fail(LF2MI_ERANGE, "lf2mi_diffpm_state: step %d outside [0, %d]", step, m->nsteps);
    if (io_dtype != LF2MI_F32 && io_dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "lf2mi_diffpm_state: io dtype must be F32 or F64");
    hipStream_t st = as_stream(stream);
    if (m->dtype == LF2MI_F32)
        return io_dtype == LF2MI_F32 ? dp_state<float, float>(m, step, (float *)pos, (float *)vel, st)
                                     : dp_state<float, double>(m, step, (double *)pos, (double *)vel, st);
    return io_dtype == LF2MI_F32 ? dp_state<double, float>(m, step, (float *)pos, (float *)vel, st)
                                 : dp_state<double, double>(m, step, (double *)pos, (double *)vel, st);
}

}  // extern "C"
