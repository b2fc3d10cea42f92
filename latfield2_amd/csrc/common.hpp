// Shared helpers of the lf2mi HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "lf2mi.h"

namespace lf2mi {

// records the message returned by lf2mi_last_error() and returns `code`
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define LF2MI_HIP(call)                                                             \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) return ::lf2mi::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define LF2MI_LAUNCHED() LF2MI_HIP(hipGetLastError())

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// looks up, once per device, the current device's address of the scratch table of
// lf2mi_poisson_green_noalloc (fft.hip)
int green_scratch_init();

// clears the ghost sites of a cubic field (field.hip)
int zero_ghosts(void *field, int N, int h, int C, int dtype, hipStream_t st);

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using cplx = typename vec2<T>::type;

template <typename T> __host__ __device__ inline cplx<T> mk(T x, T y) {
    cplx<T> r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// storage index of interior coordinates (x, y, z) in [-h, N+h) of a field of ghost width h
__host__ __device__ inline int64_t site_index(int N, int h, int x, int y, int z) {
    const int64_t G = (int64_t)N + 2 * h;
    return (int64_t)(x + h) + G * ((int64_t)(y + h) + G * (int64_t)(z + h));
}

inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Optional per-kernel timing with HIP events recorded on the launch stream
// (lf2mi_profile_*).  Disabled by default; when disabled the macros cost a branch.
bool profiling();
int prof_begin(const char *name, hipStream_t st);  // returns slot or -1
void prof_end(int slot, hipStream_t st);

struct ProfScope {
    int slot;
    hipStream_t st;
    ProfScope(const char *name, hipStream_t s) : slot(profiling() ? prof_begin(name, s) : -1), st(s) {}
    ~ProfScope() {
        if (slot >= 0) prof_end(slot, st);
    }
};
#define LF2MI_PROF(name, st) ::lf2mi::ProfScope lf2mi_prof_scope_(name, st)

}  // namespace lf2mi
