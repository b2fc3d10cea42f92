// Field kernels of the lf2mi library: projection_init, Field::updateHalo, the CIC ghost
// folds (*_comm) and the driver-side reductions.  All are HBM-bound copies/adds.
#include <cstdarg>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"

namespace lf2mi {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    return fail(e == hipErrorOutOfMemory ? LF2MI_ENOMEM : LF2MI_EHIP, "%s failed: %s (%s:%d)", what,
                hipGetErrorString(e), file, line);
}

// ---------------------------------------------------------------- profiling registry
struct ProfSlot {
    std::string name;
    double total_ms = 0.0;
    int64_t count = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    hipEvent_t open_start = nullptr;
};
static bool g_prof_on = false;
static std::vector<ProfSlot> g_prof;

bool profiling() { return g_prof_on; }

int prof_begin(const char *name, hipStream_t st) {
    int slot = -1;
    for (size_t i = 0; i < g_prof.size(); ++i)
        if (g_prof[i].name == name) slot = (int)i;
    if (slot < 0) {
        g_prof.emplace_back();
        g_prof.back().name = name;
        slot = (int)g_prof.size() - 1;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    if (hipEventRecord(e, st) != hipSuccess) {
        (void)hipEventDestroy(e);
        return -1;
    }
    g_prof[(size_t)slot].open_start = e;
    return slot;
}

void prof_end(int slot, hipStream_t st) {
    ProfSlot &s = g_prof[(size_t)slot];
    hipEvent_t e;
    if (!s.open_start || hipEventCreate(&e) != hipSuccess) return;
    if (hipEventRecord(e, st) != hipSuccess) {
        (void)hipEventDestroy(e);
        return;
    }
    s.pending.emplace_back(s.open_start, e);
    s.open_start = nullptr;
}

static void prof_collect() {
    for (ProfSlot &s : g_prof) {
        for (auto &pr : s.pending) {
            float ms = 0.f;
            if (hipEventSynchronize(pr.second) == hipSuccess && hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                s.total_ms += ms;
                s.count += 1;
            }
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        s.pending.clear();
    }
}

// ---------------------------------------------------------------- zero
template <typename T>
__global__ void zero_kernel(T *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T(0);
}

// ---------------------------------------------------------------- ghost fill
// One thread per (ghost site, component).  Ghost sites are enumerated as the six
// slabs of the gross lattice outside the interior, split so that each site appears
// once: z-slabs (all x, y), then y-slabs (z interior, all x), then x-slabs (y, z
// interior).  Each copies from its periodic image, so the result does not depend on
// the order (LATfield2_Field_decl.hpp:905-926 computes the same source index).
// ZERO: the ghost sites are cleared instead (a field whose interior is then overwritten).
template <typename T, bool ZERO = false>
__global__ void ghost_fill_kernel(T *__restrict__ f, int nx, int ny, int nz, int h, int C) {
    const int Gx = nx + 2 * h, Gy = ny + 2 * h;
    const int64_t cz = 2LL * h * Gx * Gy;        // z-slabs
    const int64_t cy = 2LL * h * nz * Gx;        // y-slabs, z interior
    const int64_t cx = 2LL * h * ny * nz;        // x-slabs, y,z interior
    const int64_t total = (cz + cy + cx) * C;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int c = (int)(t % C);
        int64_t s = t / C;
        int x, y, z;  // interior coordinates in [-h, n+h)
        if (s < cz) {
            x = (int)(s % Gx) - h;
            s /= Gx;
            y = (int)(s % Gy) - h;
            s /= Gy;
            const int k = (int)s;  // [0, 2h)
            z = k < h ? k - h : nz + (k - h);
        } else if ((s -= cz) < cy) {
            x = (int)(s % Gx) - h;
            s /= Gx;
            const int k = (int)(s % (2 * h));
            s /= (2 * h);
            z = (int)s;
            y = k < h ? k - h : ny + (k - h);
        } else {
            s -= cy;
            const int k = (int)(s % (2 * h));
            s /= (2 * h);
            y = (int)(s % ny);
            z = (int)(s / ny);
            x = k < h ? k - h : nx + (k - h);
        }
        const int sx = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        const int sy = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        const int sz = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        const int64_t dst = (int64_t)(x + h) + Gx * ((int64_t)(y + h) + (int64_t)Gy * (z + h));
        const int64_t src = (int64_t)(sx + h) + Gx * ((int64_t)(sy + h) + (int64_t)Gy * (sz + h));
        f[dst * C + c] = ZERO ? T(0) : f[src * C + c];
    }
}

// ---------------------------------------------------------------- folds
// One launch per stage; stages run in the reference's order on the stream.
//   stage 0 (x): (0,j,k) += (N,j,k) for up comps; (N-1,j,k) += (-1,j,k) for down comps;
//                j,k in [-1, N]
//   stage 1 (y): (i,0,k) += (i,N,k) up; (i,N-1,k) += (i,-1,k) down; i in [0,N),
//                k in [klo, N]
//   stage 2 (z): (i,j,0) += (i,j,N) up; (i,j,N-1) += (i,j,-1) down; i,j in [0,N)
// `down` is a bit mask over components.
template <typename T>
__global__ void fold_kernel(T *__restrict__ f, int N, int h, int C, int stage, unsigned down, int klo) {
    int na, nb, a0, b0;
    if (stage == 0) { na = N + 2; nb = N + 2; a0 = -1; b0 = -1; }
    else if (stage == 1) { na = N; nb = N + 1 - klo; a0 = 0; b0 = klo; }
    else { na = N; nb = N; a0 = 0; b0 = 0; }
    const int64_t total = (int64_t)na * nb * C;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int c = (int)(t % C);
        const int64_t s = t / C;
        const int a = (int)(s % na) + a0;
        const int b = (int)(s / na) + b0;
        int64_t dst_up, src_up, dst_dn, src_dn;
        if (stage == 0) {
            dst_up = site_index(N, h, 0, a, b);
            src_up = site_index(N, h, N, a, b);
            dst_dn = site_index(N, h, N - 1, a, b);
            src_dn = site_index(N, h, -1, a, b);
        } else if (stage == 1) {
            dst_up = site_index(N, h, a, 0, b);
            src_up = site_index(N, h, a, N, b);
            dst_dn = site_index(N, h, a, N - 1, b);
            src_dn = site_index(N, h, a, -1, b);
        } else {
            dst_up = site_index(N, h, a, b, 0);
            src_up = site_index(N, h, a, b, N);
            dst_dn = site_index(N, h, a, b, N - 1);
            src_dn = site_index(N, h, a, b, -1);
        }
        // read both sources before writing: matches the pack-then-unpack order of
        // the reference (sources are ghost rows, destinations interior rows)
        const T up = f[src_up * C + c];
        const T dn = ((down >> c) & 1u) ? f[src_dn * C + c] : T(0);
        f[dst_up * C + c] += up;
        if ((down >> c) & 1u) f[dst_dn * C + c] += dn;
    }
}

// ---------------------------------------------------------------- reductions
template <typename T>
__global__ void interior_sum_kernel(const T *__restrict__ f, int N, int h, int C, int comp,
                                    double *__restrict__ out) {
    __shared__ double red[256];
    const int64_t n = (int64_t)N * N * N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int x = (int)(i % N);
        const int y = (int)((i / N) % N);
        const int z = (int)(i / ((int64_t)N * N));
        acc += (double)f[site_index(N, h, x, y, z) * C + comp];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <typename T>
__global__ void interior_affine_kernel(T *__restrict__ f, int N, int h, int C, int comp, double mul,
                                       double add) {
    const int64_t n = (int64_t)N * N * N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int x = (int)(i % N);
        const int y = (int)((i / N) % N);
        const int z = (int)(i / ((int64_t)N * N));
        T &v = f[site_index(N, h, x, y, z) * C + comp];
        v = (T)((double)v * mul + add);
    }
}

static unsigned grid_cap(int64_t n, int threads) {
    const int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

static int check_field(int N, int halo, int C, int dtype) {
    if (N < 1 || halo < 0 || halo > N || C < 1) return fail(LF2MI_EINVAL, "bad field geometry N=%d halo=%d C=%d", N, halo, C);
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    return LF2MI_OK;
}

// clears the ghost sites of a cubic field of ghost width h (the interior is left alone)
int zero_ghosts(void *field, int N, int h, int C, int dtype, hipStream_t st) {
    if (h == 0) return LF2MI_OK;
    const int64_t G = N + 2 * h;
    const int64_t total = (G * G * G - (int64_t)N * N * N) * C;
    const unsigned g = grid_cap(total, 256);
    if (dtype == LF2MI_F32)
        ghost_fill_kernel<float, true><<<g, 256, 0, st>>>((float *)field, N, N, N, h, C);
    else
        ghost_fill_kernel<double, true><<<g, 256, 0, st>>>((double *)field, N, N, N, h, C);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}
}  // namespace lf2mi

using namespace lf2mi;

extern "C" {

int lf2mi_version(void) { return 1; }

const char *lf2mi_last_error(void) { return g_last_error.c_str(); }

int lf2mi_device_init(int device) {
    LF2MI_HIP(hipSetDevice(device));
    LF2MI_HIP(hipFree(nullptr));
    return green_scratch_init();
}

int lf2mi_synchronize(void) {
    LF2MI_HIP(hipDeviceSynchronize());
    return LF2MI_OK;
}

int lf2mi_device_alloc(void **ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return fail(LF2MI_EINVAL, "lf2mi_device_alloc: bad argument");
    *ptr = nullptr;
    if (bytes == 0) return LF2MI_OK;
    LF2MI_HIP(hipMalloc(ptr, (size_t)bytes));
    return LF2MI_OK;
}

int lf2mi_device_free(void *ptr) {
    if (ptr) LF2MI_HIP(hipFree(ptr));
    return LF2MI_OK;
}

int lf2mi_copy(void *dst, const void *src, int64_t bytes, int kind, void *stream) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(LF2MI_EINVAL, "lf2mi_copy: bad argument");
    if (bytes == 0) return LF2MI_OK;
    hipMemcpyKind k;
    switch (kind) {
        case LF2MI_COPY_H2D: k = hipMemcpyHostToDevice; break;
        case LF2MI_COPY_D2H: k = hipMemcpyDeviceToHost; break;
        case LF2MI_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
        default: return fail(LF2MI_EINVAL, "lf2mi_copy: bad kind %d", kind);
    }
    LF2MI_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, as_stream(stream)));
    if (kind != LF2MI_COPY_D2D) LF2MI_HIP(hipStreamSynchronize(as_stream(stream)));
    return LF2MI_OK;
}

int lf2mi_profile_enable(int on) {
    g_prof_on = on != 0;
    return LF2MI_OK;
}

int lf2mi_profile_reset(void) {
    prof_collect();
    g_prof.clear();
    return LF2MI_OK;
}

int lf2mi_profile_get(int index, char *name, int name_cap, double *total_ms, int64_t *count) {
    prof_collect();
    if (index < 0 || index >= (int)g_prof.size()) return fail(LF2MI_ERANGE, "profile slot %d out of range", index);
    const ProfSlot &s = g_prof[(size_t)index];
    if (name && name_cap > 0) {
        strncpy(name, s.name.c_str(), (size_t)name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (total_ms) *total_ms = s.total_ms;
    if (count) *count = s.count;
    return LF2MI_OK;
}

int lf2mi_zero(void *ptr, int64_t count, int dtype, void *stream) {
    if (count < 0 || (count > 0 && !ptr)) return fail(LF2MI_EINVAL, "lf2mi_zero: bad pointer/count");
    if (count == 0) return LF2MI_OK;
    const unsigned g = grid_cap(count, 256);
    LF2MI_PROF("zero", as_stream(stream));
    if (dtype == LF2MI_F32)
        zero_kernel<float><<<g, 256, 0, as_stream(stream)>>>((float *)ptr, count);
    else if (dtype == LF2MI_F64)
        zero_kernel<double><<<g, 256, 0, as_stream(stream)>>>((double *)ptr, count);
    else
        return fail(LF2MI_EINVAL, "lf2mi_zero: bad dtype %d", dtype);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

int lf2mi_ghost_fill(void *field, int N, int halo, int C, int dtype, void *stream) {
    return lf2mi_ghost_fill_box(field, N, N, N, halo, C, dtype, stream);
}

int lf2mi_ghost_fill_box(void *field, int nx, int ny, int nz, int halo, int C, int dtype, void *stream) {
    const int nmin = nx < ny ? (nx < nz ? nx : nz) : (ny < nz ? ny : nz);
    if (nmin < 1 || halo < 0 || halo > nmin || C < 1)
        return fail(LF2MI_EINVAL, "bad field geometry %dx%dx%d halo=%d C=%d", nx, ny, nz, halo, C);
    if (dtype != LF2MI_F32 && dtype != LF2MI_F64) return fail(LF2MI_EINVAL, "bad dtype %d", dtype);
    if (!field) return fail(LF2MI_EINVAL, "lf2mi_ghost_fill: null field");
    if (halo == 0) return LF2MI_OK;
    const int64_t gross = (int64_t)(nx + 2 * halo) * (ny + 2 * halo) * (nz + 2 * halo);
    const int64_t total = (gross - (int64_t)nx * ny * nz) * C;
    const unsigned g = grid_cap(total, 256);
    LF2MI_PROF("ghost_fill", as_stream(stream));
    if (dtype == LF2MI_F32)
        ghost_fill_kernel<float><<<g, 256, 0, as_stream(stream)>>>((float *)field, nx, ny, nz, halo, C);
    else
        ghost_fill_kernel<double><<<g, 256, 0, as_stream(stream)>>>((double *)field, nx, ny, nz, halo, C);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

int lf2mi_projection_comm(void *field, int N, int halo, int kind, int dtype, void *stream) {
    int C;
    unsigned down;
    int klo;
    switch (kind) {
        case LF2MI_PROJ_SCALAR: C = 1; down = 0u; klo = 0; break;     // y fold: z in [0, N]
        case LF2MI_PROJ_VECTOR: C = 3; down = 7u; klo = -1; break;    // y fold: z in [-1, N]
        case LF2MI_PROJ_VECVEC: C = 6; down = (1u << 1) | (1u << 2) | (1u << 4); klo = -1; break;
        // the CIC-NGP folds are the scalar fold on every component
        case LF2MI_PROJ_VECTOR_NGP: C = 3; down = 0u; klo = 0; break;
        case LF2MI_PROJ_SYMTENSOR_NGP: C = 6; down = 0u; klo = 0; break;
        default: return fail(LF2MI_EINVAL, "lf2mi_projection_comm: bad kind %d", kind);
    }
    if (int rc = check_field(N, halo, C, dtype)) return rc;
    if (halo < 1) return fail(LF2MI_EINVAL, "the field has to have at least a halo of 1");
    if (!field) return fail(LF2MI_EINVAL, "lf2mi_projection_comm: null field");
    LF2MI_PROF("projection_comm", as_stream(stream));
    for (int stage = 0; stage < 3; ++stage) {
        const int64_t total = (int64_t)(N + 2) * (N + 2) * C;
        const unsigned g = grid_cap(total, 256);
        if (dtype == LF2MI_F32)
            fold_kernel<float><<<g, 256, 0, as_stream(stream)>>>((float *)field, N, halo, C, stage, down, klo);
        else
            fold_kernel<double><<<g, 256, 0, as_stream(stream)>>>((double *)field, N, halo, C, stage, down, klo);
        LF2MI_LAUNCHED();
    }
    return LF2MI_OK;
}

int lf2mi_field_sum(const void *field, int N, int halo, int C, int comp, int dtype, double *out, void *stream) {
    if (int rc = check_field(N, halo, C, dtype)) return rc;
    if (!field || !out || comp < 0 || comp >= C) return fail(LF2MI_EINVAL, "lf2mi_field_sum: bad argument");
    const int64_t n = (int64_t)N * N * N;
    const unsigned g = grid_cap(n, 256) > 1024 ? 1024 : grid_cap(n, 256);
    double *part = nullptr;
    LF2MI_HIP(hipMallocAsync((void **)&part, sizeof(double) * g, as_stream(stream)));
    if (dtype == LF2MI_F32)
        interior_sum_kernel<float><<<g, 256, 0, as_stream(stream)>>>((const float *)field, N, halo, C, comp, part);
    else
        interior_sum_kernel<double><<<g, 256, 0, as_stream(stream)>>>((const double *)field, N, halo, C, comp, part);
    LF2MI_LAUNCHED();
    double host[1024];
    LF2MI_HIP(hipMemcpyAsync(host, part, sizeof(double) * g, hipMemcpyDeviceToHost, as_stream(stream)));
    LF2MI_HIP(hipFreeAsync(part, as_stream(stream)));
    LF2MI_HIP(hipStreamSynchronize(as_stream(stream)));
    double s = 0.0;
    for (unsigned i = 0; i < g; ++i) s += host[i];
    *out = s;
    return LF2MI_OK;
}

int lf2mi_field_affine(void *field, int N, int halo, int C, int comp, double mul, double add, int dtype,
                       void *stream) {
    if (int rc = check_field(N, halo, C, dtype)) return rc;
    if (!field || comp < 0 || comp >= C) return fail(LF2MI_EINVAL, "lf2mi_field_affine: bad argument");
    const unsigned g = grid_cap((int64_t)N * N * N, 256);
    if (dtype == LF2MI_F32)
        interior_affine_kernel<float><<<g, 256, 0, as_stream(stream)>>>((float *)field, N, halo, C, comp, mul, add);
    else
        interior_affine_kernel<double><<<g, 256, 0, as_stream(stream)>>>((double *)field, N, halo, C, comp, mul, add);
    LF2MI_LAUNCHED();
    return LF2MI_OK;
}

}  // extern "C"
