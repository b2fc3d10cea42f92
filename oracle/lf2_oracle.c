/*
 * lf2_oracle.c -- FP64 CPU restatement of the LATfield2 particle-mesh path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each function cites the reference source it restates (paths relative to the
 * LATfield2 tree).  The arithmetic follows the reference's formulas in double
 * precision; loop structure is this file's own.  FFTs are a textbook radix-2
 * (power-of-two lengths) or direct DFT (other lengths) in double precision: the
 * reference delegates them to FFTW, whose exact rounding is not part of the
 * contract (the KATs of benchmarks_tests/wave_test*.cpp pin values and layout).
 */
#include "lf2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef orc_i64 i64;

#define ORC_EBADARG (-1)
#define ORC_ENOMEM (-2)

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    (void)n;
    return 0;
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* storage index of the site with interior coordinates (x,y,z), each in [-h, N+h) */
static inline i64 sidx(int N, int h, int x, int y, int z) {
    const i64 G = (i64)N + 2 * h;
    return (i64)(x + h) + G * ((i64)(y + h) + G * (i64)(z + h));
}

int orc_zero(double *f, i64 count) {
    if (!f || count < 0) return ORC_EBADARG;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < count; ++i) f[i] = 0.0;
    return 0;
}

/* Field::updateHalo (LATfield2_Field_decl.hpp:889-935): every ghost site copies all
 * components from the site shifted by +-N along each out-of-range axis. */
int orc_ghost_fill(double *f, int N, int h, int C) {
    if (!f || N < 1 || h < 0 || h > N || C < 1) return ORC_EBADARG;
    const int G = N + 2 * h;
#pragma omp parallel for schedule(static)
    for (int sz = 0; sz < G; ++sz) {
        for (int sy = 0; sy < G; ++sy) {
            for (int sx = 0; sx < G; ++sx) {
                int c3[3] = {sx - h, sy - h, sz - h};
                int ghost = 0;
                for (int d = 0; d < 3; ++d) {
                    if (c3[d] < 0) { c3[d] += N; ghost = 1; }
                    else if (c3[d] >= N) { c3[d] -= N; ghost = 1; }
                }
                if (!ghost) continue;
                const i64 dst = (i64)sx + (i64)G * ((i64)sy + (i64)G * sz);
                const i64 src = sidx(N, h, c3[0], c3[1], c3[2]);
                for (int c = 0; c < C; ++c) f[dst * C + c] = f[src * C + c];
            }
        }
    }
    return 0;
}

/* Particles::getPartCoord (particles/LATfield2_Particles.hpp:464-467) */
static inline int cell_coord(double p, int N) { return (int)floor(p * N) % N; }

int orc_bin(const double *pos, i64 np, int N, int *key) {
    if (!pos || !key || np < 0 || N < 1) return ORC_EBADARG;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < np; ++i) {
        int c0 = cell_coord(pos[3 * i + 0], N);
        int c1 = cell_coord(pos[3 * i + 1], N);
        int c2 = cell_coord(pos[3 * i + 2], N);
        key[i] = (c0 < 0 || c1 < 0 || c2 < 0) ? -1 : c0 + N * (c1 + N * c2);
    }
    return 0;
}

int orc_cell_counts(const double *pos, i64 np, int N, int *counts) {
    if (!pos || !counts || np < 0 || N < 1) return ORC_EBADARG;
    const i64 nc = (i64)N * N * N;
    memset(counts, 0, sizeof(int) * (size_t)nc);
    for (i64 i = 0; i < np; ++i) {
        int c0 = cell_coord(pos[3 * i + 0], N);
        int c1 = cell_coord(pos[3 * i + 1], N);
        int c2 = cell_coord(pos[3 * i + 2], N);
        if (c0 < 0 || c1 < 0 || c2 < 0) continue;
        counts[c0 + (i64)N * (c1 + (i64)N * c2)]++;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Per-cell particle lists.  The reference keeps one std::forward_list per cell
 * filled with push_front (Particles.hpp:511), so a cell is walked from the most
 * recently added particle to the oldest; this mirrors that order with a
 * counting sort over the cell keys.
 * ------------------------------------------------------------------------- */
typedef struct {
    i64 *start; /* N^3 + 1 */
    i64 *order; /* particle indices grouped by cell */
} cell_lists;

static int build_lists(const double *pos, i64 np, int N, cell_lists *L) {
    const i64 nc = (i64)N * N * N;
    int *key = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    L->start = (i64 *)calloc((size_t)nc + 1, sizeof(i64));
    L->order = (i64 *)malloc(sizeof(i64) * (size_t)(np > 0 ? np : 1));
    if (!key || !L->start || !L->order) {
        free(key); free(L->start); free(L->order);
        return ORC_ENOMEM;
    }
    orc_bin(pos, np, N, key);
    for (i64 i = 0; i < np; ++i)
        if (key[i] >= 0) L->start[key[i] + 1]++;
    for (i64 c = 0; c < nc; ++c) L->start[c + 1] += L->start[c];
    i64 *fill = (i64 *)malloc(sizeof(i64) * (size_t)nc);
    if (!fill) { free(key); free(L->start); free(L->order); return ORC_ENOMEM; }
    memcpy(fill, L->start, sizeof(i64) * (size_t)nc);
    for (i64 i = np - 1; i >= 0; --i)
        if (key[i] >= 0) L->order[fill[key[i]]++] = i;
    free(fill);
    free(key);
    return 0;
}

static void free_lists(cell_lists *L) {
    free(L->start);
    free(L->order);
}

static inline double pmass(const double *mass, double mass_global, i64 i) {
    return mass ? mass[i] : mass_global;
}

/* scalarProjectionCIC_project (particles/projections.hpp:76-184).  Per cell, the
 * eight corner sums of m/res^6 * prod(length weights) are accumulated over the
 * cell's particles (the reference's localCube[8]) and then added to the corners
 * x+{0,1}, y+{0,1}, z+{0,1}; the +1 corners may be upper ghosts. */
int orc_scalar_project(const double *pos, const double *mass, double mass_global,
                       i64 np, double *rho, int N, int h) {
    if (!pos || !rho || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    cell_lists L;
    int rc = build_lists(pos, np, N, &L);
    if (rc) return rc;
    const double res = 1.0 / N;
    double cicVol = res * res * res;
    cicVol *= cicVol;
    /* cell planes z and z+1 write node planes {z,z+1} and {z+1,z+2}: two passes
     * over planes of equal parity keep the writes disjoint between threads */
    for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = parity; z < N; z += 2) {
            for (int y = 0; y < N; ++y) {
                for (int x = 0; x < N; ++x) {
                    const i64 cell = x + (i64)N * (y + (i64)N * z);
                    const i64 b = L.start[cell], e = L.start[cell + 1];
                    if (b == e) continue;
                    const double ref[3] = {x * res, y * res, z * res};
                    double cube[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (i64 q = b; q < e; ++q) {
                        const i64 p = L.order[q];
                        double up[3], dn[3];
                        for (int d = 0; d < 3; ++d) {
                            up[d] = pos[3 * p + d] - ref[d];
                            dn[d] = res - up[d];
                        }
                        const double m = pmass(mass, mass_global, p) / cicVol;
                        /* corner bits: 4 = +x, 2 = +y, 1 = +z */
                        for (int cb = 0; cb < 8; ++cb) {
                            const double wx = (cb & 4) ? up[0] : dn[0];
                            const double wy = (cb & 2) ? up[1] : dn[1];
                            const double wz = (cb & 1) ? up[2] : dn[2];
                            cube[cb] += wx * wy * wz * m;
                        }
                    }
                    for (int cb = 0; cb < 8; ++cb)
                        rho[sidx(N, h, x + ((cb >> 2) & 1), y + ((cb >> 1) & 1), z + (cb & 1))] += cube[cb];
                }
            }
        }
    }
    free_lists(&L);
    return 0;
}

/* One-axis weights shared by the vector and VecVec deposits
 * (particles/projections.hpp:1044-1071, :1426-1447):
 *   vertex grid : nodes c, c+1 with weights 1-u, u (u = (pos - c res)/res);
 *   link grid   : nodes at (n+1/2) res; t = (pos - (c+1/2) res)/res;
 *                 t < 0 -> nodes c-1, c with weights -t, 1+t;
 *                 else  -> nodes c, c+1 with weights 1-t, t. */
typedef struct {
    int lo;      /* node offset (relative to the cell) of the lower node */
    double w[2]; /* weights of nodes lo, lo+1 */
} axis_w;

static inline axis_w vertex_w(double p, int c, double res) {
    axis_w a;
    const double u = (p - c * res) / res;
    a.lo = 0;
    a.w[0] = 1.0 - u;
    a.w[1] = u;
    return a;
}

static inline axis_w link_w(double p, int c, double res) {
    axis_w a;
    const double t = (p - (c * res + res * 0.5)) / res;
    if (t < 0) {
        a.lo = -1;
        a.w[0] = -t;
        a.w[1] = 1.0 - a.w[0];
    } else {
        a.lo = 0;
        a.w[1] = t;
        a.w[0] = 1.0 - a.w[1];
    }
    return a;
}

static inline void add8(double *f, int C, int comp, int N, int h, int x, int y, int z,
                        const axis_w *ax, const axis_w *ay, const axis_w *az, double q) {
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const double w = ax->w[i] * ay->w[j] * az->w[k];
                f[sidx(N, h, x + ax->lo + i, y + ay->lo + j, z + az->lo + k) * C + comp] += q * w;
            }
}

/* vectorProjectionCIC_project (particles/projections.hpp:980-1189): component i of
 * m v / res^3 is deposited on the link grid along axis i (node n stands for
 * position (n + e_i/2) res) and on the vertex grid along the other axes.
 * The reference skips empty cells with a partList::size test that does not
 * compile; the list walk below visits exactly the non-empty cells. */
int orc_vector_project(const double *pos, const double *vel, const double *mass,
                       double mass_global, i64 np, double *B, int N, int h) {
    if (!pos || !vel || !B || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    cell_lists L;
    int rc = build_lists(pos, np, N, &L);
    if (rc) return rc;
    const double res = 1.0 / N;
    const double cicVol = res * res * res;
    /* cell plane z writes node planes z-1..z+1: three passes by z mod 3 */
    for (int phase = 0; phase < 3; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = phase; z < N; z += 3) {
            for (int y = 0; y < N; ++y) {
                for (int x = 0; x < N; ++x) {
                    const i64 cell = x + (i64)N * (y + (i64)N * z);
                    const int c3[3] = {x, y, z};
                    for (i64 q = L.start[cell]; q < L.start[cell + 1]; ++q) {
                        const i64 p = L.order[q];
                        axis_w v[3], l[3];
                        for (int d = 0; d < 3; ++d) {
                            v[d] = vertex_w(pos[3 * p + d], c3[d], res);
                            l[d] = link_w(pos[3 * p + d], c3[d], res);
                        }
                        const double m = pmass(mass, mass_global, p) / cicVol;
                        add8(B, 3, 0, N, h, x, y, z, &l[0], &v[1], &v[2], m * vel[3 * p + 0]);
                        add8(B, 3, 1, N, h, x, y, z, &v[0], &l[1], &v[2], m * vel[3 * p + 1]);
                        add8(B, 3, 2, N, h, x, y, z, &v[0], &v[1], &l[2], m * vel[3 * p + 2]);
                    }
                }
            }
        }
    }
    free_lists(&L);
    return 0;
}

/* VecVecProjectionCIC_project (particles/projections.hpp:1356-1622): T_ii += m v_i^2
 * on the vertex grid; T_ij (i<j) += m v_i v_j on the plaquette grid (link weights
 * along i and j, vertex weights along the third axis).  Symmetric storage
 * (0,0)=0 (0,1)=1 (0,2)=2 (1,1)=3 (1,2)=4 (2,2)=5 (LATfield2_Field_decl.hpp:773-788). */
int orc_vecvec_project(const double *pos, const double *vel, const double *mass,
                       double mass_global, i64 np, double *T, int N, int h) {
    if (!pos || !vel || !T || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    cell_lists L;
    int rc = build_lists(pos, np, N, &L);
    if (rc) return rc;
    const double res = 1.0 / N;
    const double cicVol = res * res * res;
    for (int phase = 0; phase < 3; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = phase; z < N; z += 3) {
            for (int y = 0; y < N; ++y) {
                for (int x = 0; x < N; ++x) {
                    const i64 cell = x + (i64)N * (y + (i64)N * z);
                    const int c3[3] = {x, y, z};
                    for (i64 q = L.start[cell]; q < L.start[cell + 1]; ++q) {
                        const i64 p = L.order[q];
                        axis_w v[3], l[3];
                        for (int d = 0; d < 3; ++d) {
                            v[d] = vertex_w(pos[3 * p + d], c3[d], res);
                            l[d] = link_w(pos[3 * p + d], c3[d], res);
                        }
                        const double m = pmass(mass, mass_global, p) / cicVol;
                        const double *u = vel + 3 * p;
                        add8(T, 6, 0, N, h, x, y, z, &v[0], &v[1], &v[2], m * u[0] * u[0]);
                        add8(T, 6, 3, N, h, x, y, z, &v[0], &v[1], &v[2], m * u[1] * u[1]);
                        add8(T, 6, 5, N, h, x, y, z, &v[0], &v[1], &v[2], m * u[2] * u[2]);
                        add8(T, 6, 1, N, h, x, y, z, &l[0], &l[1], &v[2], m * u[0] * u[1]);
                        add8(T, 6, 2, N, h, x, y, z, &l[0], &v[1], &l[2], m * u[0] * u[2]);
                        add8(T, 6, 4, N, h, x, y, z, &v[0], &l[1], &l[2], m * u[1] * u[2]);
                    }
                }
            }
        }
    }
    free_lists(&L);
    return 0;
}

/* vectorProjectionCICNGP_project (particles/projections.hpp:422-549): per cell, component
 * i of m v / res^3 goes to the cell's own node along axis i (the link x + e_i/2 that holds
 * the particle) and to the four vertex nodes {0,1}^2 of the other two axes with the CIC
 * weights; the reference's vi[12] per-cell sums are kept here too. */
int orc_vector_ngp_project(const double *pos, const double *vel, const double *mass,
                           double mass_global, i64 np, double *B, int N, int h) {
    if (!pos || !vel || !B || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    cell_lists L;
    int rc = build_lists(pos, np, N, &L);
    if (rc) return rc;
    const double res = 1.0 / N;
    const double cicVol = res * res * res;
    for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = parity; z < N; z += 2) {
            for (int y = 0; y < N; ++y) {
                for (int x = 0; x < N; ++x) {
                    const i64 cell = x + (i64)N * (y + (i64)N * z);
                    const i64 b = L.start[cell], e = L.start[cell + 1];
                    if (b == e) continue;
                    const double ref[3] = {x * res, y * res, z * res};
                    /* vi[4 i + 2 hi + lo]: component i, corner (lo, hi) on the two other
                     * axes taken in increasing order */
                    double vi[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for (i64 q = b; q < e; ++q) {
                        const i64 p = L.order[q];
                        double up[3], dn[3];
                        for (int d = 0; d < 3; ++d) {
                            up[d] = (pos[3 * p + d] - ref[d]) / res;
                            dn[d] = 1.0 - up[d];
                        }
                        const double m = pmass(mass, mass_global, p) / cicVol;
                        for (int i = 0; i < 3; ++i) {
                            const int a0 = i == 0 ? 1 : 0, a1 = i == 2 ? 1 : 2;
                            const double mv = m * vel[3 * p + i];
                            vi[4 * i + 0] += mv * dn[a0] * dn[a1];
                            vi[4 * i + 1] += mv * up[a0] * dn[a1];
                            vi[4 * i + 2] += mv * dn[a0] * up[a1];
                            vi[4 * i + 3] += mv * up[a0] * up[a1];
                        }
                    }
                    for (int i = 0; i < 3; ++i) {
                        const int a0 = i == 0 ? 1 : 0, a1 = i == 2 ? 1 : 2;
                        for (int k = 0; k < 4; ++k) {
                            int o[3] = {x, y, z};
                            o[a0] += k & 1;
                            o[a1] += (k >> 1) & 1;
                            B[sidx(N, h, o[0], o[1], o[2]) * 3 + i] += vi[4 * i + k];
                        }
                    }
                }
            }
        }
    }
    free_lists(&L);
    return 0;
}

/* symtensorProjectionCICNGP_project (particles/projections.hpp:709-847): T_ii += m v_i^2
 * with the eight CIC vertex weights; T_ij (i<j) += m v_i v_j at the cell's node along i and
 * j, split between the cell and its +1 neighbour along the third axis with the CIC weights.
 * Symmetric storage (0,0)=0 (0,1)=1 (0,2)=2 (1,1)=3 (1,2)=4 (2,2)=5. */
int orc_symtensor_ngp_project(const double *pos, const double *vel, const double *mass,
                              double mass_global, i64 np, double *T, int N, int h) {
    if (!pos || !vel || !T || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    cell_lists L;
    int rc = build_lists(pos, np, N, &L);
    if (rc) return rc;
    const double res = 1.0 / N;
    const double cicVol = res * res * res;
    static const int diag[3] = {0, 3, 5};
    static const int offd[3] = {1, 2, 4};     /* (0,1) (0,2) (1,2) */
    static const int third[3] = {2, 1, 0};    /* the CIC axis of each off-diagonal entry */
    static const int va[3] = {0, 0, 1}, vb[3] = {1, 2, 2};
    for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int z = parity; z < N; z += 2) {
            for (int y = 0; y < N; ++y) {
                for (int x = 0; x < N; ++x) {
                    const i64 cell = x + (i64)N * (y + (i64)N * z);
                    const i64 b = L.start[cell], e = L.start[cell + 1];
                    if (b == e) continue;
                    const double ref[3] = {x * res, y * res, z * res};
                    double tii[24], tij[6];
                    memset(tii, 0, sizeof tii);
                    memset(tij, 0, sizeof tij);
                    for (i64 q = b; q < e; ++q) {
                        const i64 p = L.order[q];
                        double up[3], dn[3];
                        for (int d = 0; d < 3; ++d) {
                            up[d] = (pos[3 * p + d] - ref[d]) / res;
                            dn[d] = 1.0 - up[d];
                        }
                        const double m = pmass(mass, mass_global, p) / cicVol;
                        const double *u = vel + 3 * p;
                        for (int i = 0; i < 3; ++i) {
                            const double mvv = m * u[i] * u[i];
                            /* corner bits: 4 = +x, 2 = +y, 1 = +z */
                            for (int cb = 0; cb < 8; ++cb) {
                                const double wx = (cb & 4) ? up[0] : dn[0];
                                const double wy = (cb & 2) ? up[1] : dn[1];
                                const double wz = (cb & 1) ? up[2] : dn[2];
                                tii[8 * i + cb] += mvv * wx * wy * wz;
                            }
                        }
                        for (int k = 0; k < 3; ++k) {
                            const double mvv = m * u[va[k]] * u[vb[k]];
                            tij[2 * k + 0] += mvv * dn[third[k]];
                            tij[2 * k + 1] += mvv * up[third[k]];
                        }
                    }
                    for (int i = 0; i < 3; ++i)
                        for (int cb = 0; cb < 8; ++cb)
                            T[sidx(N, h, x + ((cb >> 2) & 1), y + ((cb >> 1) & 1), z + (cb & 1)) * 6 + diag[i]] +=
                                tii[8 * i + cb];
                    for (int k = 0; k < 3; ++k)
                        for (int s = 0; s < 2; ++s) {
                            int o[3] = {x, y, z};
                            o[third[k]] += s;
                            T[sidx(N, h, o[0], o[1], o[2]) * 6 + offd[k]] += tij[2 * k + s];
                        }
                }
            }
        }
    }
    free_lists(&L);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Ghost folds (the *_comm functions) on one process.  Coordinates below are
 * interior coordinates; the reference loops use storage indices with
 * distHaloOne = h-1, so its ranges [h-1, N+h+1) are interior [-1, N].
 * On one rank sendUp/sendUpDown deliver the packed face to the same process.
 * ------------------------------------------------------------------------- */

/* scalarProjectionCIC_comm (particles/projections.hpp:194-300) */
int orc_scalar_comm(double *rho, int N, int h) {
    if (!rho || N < 1 || h < 1) return ORC_EBADARG;
    /* x: upper ghost plane onto x=0, for y,z in [-1, N] */
    for (int z = -1; z <= N; ++z)
        for (int y = -1; y <= N; ++y) rho[sidx(N, h, 0, y, z)] += rho[sidx(N, h, N, y, z)];
    /* y: upper ghost row onto y=0, for x in [0,N), z in [0, N] */
    for (int z = 0; z <= N; ++z)
        for (int x = 0; x < N; ++x) rho[sidx(N, h, x, 0, z)] += rho[sidx(N, h, x, N, z)];
    /* z: upper ghost plane onto z=0, for x,y in [0,N) */
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) rho[sidx(N, h, x, y, 0)] += rho[sidx(N, h, x, y, N)];
    return 0;
}

/* shared body of vectorProjectionCIC_comm (:1192-1351) and VecVecProjectionCIC_comm
 * (:1624-1814): all C comps fold from the upper ghosts onto index 0; the comps in
 * `down` fold from the lower ghosts onto index N-1.  x first (y,z in [-1,N]), then y
 * (x in [0,N), z in [-1,N]), then z (x,y in [0,N)). */
static int staggered_comm(double *f, int N, int h, int C, const int *down, int ndown) {
    if (!f || N < 1 || h < 1) return ORC_EBADARG;
    for (int z = -1; z <= N; ++z)
        for (int y = -1; y <= N; ++y) {
            for (int c = 0; c < C; ++c)
                f[sidx(N, h, 0, y, z) * C + c] += f[sidx(N, h, N, y, z) * C + c];
            for (int q = 0; q < ndown; ++q)
                f[sidx(N, h, N - 1, y, z) * C + down[q]] += f[sidx(N, h, -1, y, z) * C + down[q]];
        }
    for (int z = -1; z <= N; ++z)
        for (int x = 0; x < N; ++x) {
            for (int c = 0; c < C; ++c)
                f[sidx(N, h, x, 0, z) * C + c] += f[sidx(N, h, x, N, z) * C + c];
            for (int q = 0; q < ndown; ++q)
                f[sidx(N, h, x, N - 1, z) * C + down[q]] += f[sidx(N, h, x, -1, z) * C + down[q]];
        }
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            for (int c = 0; c < C; ++c)
                f[sidx(N, h, x, y, 0) * C + c] += f[sidx(N, h, x, y, N) * C + c];
            for (int q = 0; q < ndown; ++q)
                f[sidx(N, h, x, y, N - 1) * C + down[q]] += f[sidx(N, h, x, y, -1) * C + down[q]];
        }
    return 0;
}

int orc_vector_comm(double *B, int N, int h) {
    static const int down[3] = {0, 1, 2};
    return staggered_comm(B, N, h, 3, down, 3);
}

int orc_vecvec_comm(double *T, int N, int h) {
    static const int down[3] = {1, 2, 4}; /* (0,1) (0,2) (1,2) */
    return staggered_comm(T, N, h, 6, down, 3);
}

/* vectorProjectionCICNGP_comm (particles/projections.hpp:558-681) and
 * symtensorProjectionCICNGP_comm (:856-976): the ranges of scalarProjectionCIC_comm on
 * every component (x: y,z in [-1,N]; y: x in [0,N), z in [0,N]; z: x,y in [0,N)). */
static int upper_comm(double *f, int N, int h, int C) {
    if (!f || N < 1 || h < 1) return ORC_EBADARG;
    for (int z = -1; z <= N; ++z)
        for (int y = -1; y <= N; ++y)
            for (int c = 0; c < C; ++c) f[sidx(N, h, 0, y, z) * C + c] += f[sidx(N, h, N, y, z) * C + c];
    for (int z = 0; z <= N; ++z)
        for (int x = 0; x < N; ++x)
            for (int c = 0; c < C; ++c) f[sidx(N, h, x, 0, z) * C + c] += f[sidx(N, h, x, N, z) * C + c];
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x)
            for (int c = 0; c < C; ++c) f[sidx(N, h, x, y, 0) * C + c] += f[sidx(N, h, x, y, N) * C + c];
    return 0;
}

int orc_vector_ngp_comm(double *B, int N, int h) { return upper_comm(B, N, h, 3); }

int orc_symtensor_ngp_comm(double *T, int N, int h) { return upper_comm(T, N, h, 6); }

int orc_brute_deposit(const double *pos, const double *vel, const double *mass,
                      double mass_global, i64 np, int N, const double *stagger,
                      int va, int vb, int ngp, double *out) {
    if (!pos || !out || !stagger || np < 0 || N < 1) return ORC_EBADARG;
    if ((va >= 0 || vb >= 0) && !vel) return ORC_EBADARG;
    const double res = 1.0 / N;
    const double inv_vol = 1.0 / (res * res * res);
    memset(out, 0, sizeof(double) * (size_t)N * N * N);
    for (i64 p = 0; p < np; ++p) {
        double q = pmass(mass, mass_global, p) * inv_vol;
        if (va >= 0) q *= vel[3 * p + va];
        if (vb >= 0) q *= vel[3 * p + vb];
        int n0[3];
        double f[3];
        for (int d = 0; d < 3; ++d) {
            const double t = pos[3 * p + d] * N - stagger[d];
            const double fl = floor(t);
            n0[d] = (int)fl;
            f[d] = t - fl;
        }
        /* nearest-grid-point axes: all weight on the nearest node of the staggered grid */
        double wl[3][2];
        for (int d = 0; d < 3; ++d) {
            if ((ngp >> d) & 1) {
                if (f[d] >= 0.5) n0[d] += 1;
                wl[d][0] = 1.0;
                wl[d][1] = 0.0;
            } else {
                wl[d][0] = 1.0 - f[d];
                wl[d][1] = f[d];
            }
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const double w = wl[0][i] * wl[1][j] * wl[2][k];
                    const int x = ((n0[0] + i) % N + N) % N;
                    const int y = ((n0[1] + j) % N + N) % N;
                    const int z = ((n0[2] + k) % N + N) % N;
                    out[x + (i64)N * (y + (i64)N * z)] += q * w;
                }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * 1-D complex FFT (double).  Power-of-two lengths: iterative radix-2 with an
 * exact-angle twiddle table; other lengths: direct DFT.
 * ------------------------------------------------------------------------- */
typedef struct {
    int n;
    int pow2;
    double *tw; /* cos, sin of 2 pi m / n, m in [0, n) */
    int *rev;
} fft_plan1;

static int plan1_init(fft_plan1 *P, int n) {
    P->n = n;
    P->pow2 = (n & (n - 1)) == 0;
    P->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    P->rev = NULL;
    if (!P->tw) return ORC_ENOMEM;
    for (int m = 0; m < n; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)n;
        P->tw[2 * m] = cos(a);
        P->tw[2 * m + 1] = sin(a);
    }
    if (P->pow2) {
        P->rev = (int *)malloc(sizeof(int) * (size_t)n);
        if (!P->rev) return ORC_ENOMEM;
        int bits = 0;
        while ((1 << bits) < n) ++bits;
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < bits; ++b)
                if (i & (1 << b)) r |= 1 << (bits - 1 - b);
            P->rev[i] = r;
        }
    }
    return 0;
}

static void plan1_free(fft_plan1 *P) {
    free(P->tw);
    free(P->rev);
}

/* in-place transform of a[0..n) (complex interleaved); sign -1 forward, +1 backward.
 * tmp: scratch of 2n doubles for the direct DFT. */
static void fft1(const fft_plan1 *P, double *a, int sign, double *tmp) {
    const int n = P->n;
    if (n == 1) return;
    if (P->pow2) {
        for (int i = 0; i < n; ++i) {
            const int r = P->rev[i];
            if (r > i) {
                double t0 = a[2 * i], t1 = a[2 * i + 1];
                a[2 * i] = a[2 * r];
                a[2 * i + 1] = a[2 * r + 1];
                a[2 * r] = t0;
                a[2 * r + 1] = t1;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len >> 1, step = n / len;
            for (int s = 0; s < n; s += len) {
                for (int j = 0; j < half; ++j) {
                    const double wr = P->tw[2 * (j * step)];
                    const double wi = sign * P->tw[2 * (j * step) + 1];
                    double *u = a + 2 * (s + j), *v = a + 2 * (s + j + half);
                    const double tr = v[0] * wr - v[1] * wi;
                    const double ti = v[0] * wi + v[1] * wr;
                    v[0] = u[0] - tr;
                    v[1] = u[1] - ti;
                    u[0] += tr;
                    u[1] += ti;
                }
            }
        }
    } else {
        for (int k = 0; k < n; ++k) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                const int m = (int)(((long long)j * k) % n);
                const double wr = P->tw[2 * m], wi = sign * P->tw[2 * m + 1];
                sr += a[2 * j] * wr - a[2 * j + 1] * wi;
                si += a[2 * j] * wi + a[2 * j + 1] * wr;
            }
            tmp[2 * k] = sr;
            tmp[2 * k + 1] = si;
        }
        memcpy(a, tmp, sizeof(double) * 2 * (size_t)n);
    }
}

/* transform all lines of a strided 3-D complex array along one axis:
 * line l has elements base(l) + e*estride, e in [0,n). */
static void fft_lines(const fft_plan1 *P, double *buf, i64 nlines, i64 (*base)(i64, const i64 *),
                      const i64 *geo, i64 estride, int sign) {
    const int n = P->n;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 4 * (size_t)n);
        double *tmp = line + 2 * n;
#pragma omp for schedule(static)
        for (i64 l = 0; l < nlines; ++l) {
            const i64 b = base(l, geo);
            for (int e = 0; e < n; ++e) {
                line[2 * e] = buf[2 * (b + e * estride)];
                line[2 * e + 1] = buf[2 * (b + e * estride) + 1];
            }
            fft1(P, line, sign, tmp);
            for (int e = 0; e < n; ++e) {
                buf[2 * (b + e * estride)] = line[2 * e];
                buf[2 * (b + e * estride) + 1] = line[2 * e + 1];
            }
        }
        free(line);
    }
}

/* geo = {n0, n1}: array dims (n0 fastest, n1, n2).  Line bases for each axis. */
static i64 base_ax0(i64 l, const i64 *g) { return l * g[0]; }                       /* l over (i1,i2) */
static i64 base_ax1(i64 l, const i64 *g) { return (l % g[0]) + (l / g[0]) * g[0] * g[1]; } /* (i0,i2) */
static i64 base_ax2(i64 l, const i64 *g) { return l; }                              /* (i0,i1) */

/* PlanFFT::execute(FFT_FORWARD), R2C (LATfield2_PlanFFT_decl.hpp:855-950): per component,
 * the unnormalised DFT with e^{-2 pi i k.x/N}, written for kx in [0, N/2] at k-lattice
 * site (kx, kz, ky) (transpose_back_0_3 / implement_0, :1469-1510). */
int orc_fft_r2c(const double *r, double *k, int N, int C, int rh, int kh) {
    if (!r || !k || N < 1 || C < 1 || rh < 0 || kh < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    if (!buf) return ORC_ENOMEM;
    fft_plan1 P;
    if (plan1_init(&P, N)) { free(buf); return ORC_ENOMEM; }
    const int H = N / 2 + 1;
    const i64 R = H + 2 * kh, KG = (i64)N + 2 * kh;
    const i64 geo[2] = {N, N};
    for (int c = 0; c < C; ++c) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) {
                    const i64 b = x + (i64)N * (y + (i64)N * z);
                    buf[2 * b] = r[sidx(N, rh, x, y, z) * C + c];
                    buf[2 * b + 1] = 0.0;
                }
        fft_lines(&P, buf, (i64)N * N, base_ax0, geo, 1, -1);
        fft_lines(&P, buf, (i64)N * N, base_ax1, geo, N, -1);
        fft_lines(&P, buf, (i64)N * N, base_ax2, geo, (i64)N * N, -1);
#pragma omp parallel for schedule(static)
        for (int ky = 0; ky < N; ++ky)
            for (int kz = 0; kz < N; ++kz)
                for (int kx = 0; kx < H; ++kx) {
                    const i64 b = kx + (i64)N * (ky + (i64)N * kz);
                    const i64 o = ((kx + kh) + R * ((kz + kh) + KG * (ky + kh))) * C + c;
                    k[2 * o] = buf[2 * b];
                    k[2 * o + 1] = buf[2 * b + 1];
                }
    }
    plan1_free(&P);
    free(buf);
    return 0;
}

/* PlanFFT::execute(FFT_BACKWARD), C2R (LATfield2_PlanFFT_decl.hpp:951-1056): the inverse
 * complex transforms along ky and kz, then FFTW-style c2r rows along x (Hermitian
 * extension X[N-k] = conj X[k], real parts of kx = 0 and kx = N/2). */
int orc_fft_c2r(const double *k, double *r, int N, int C, int rh, int kh) {
    if (!r || !k || N < 1 || C < 1 || rh < 0 || kh < 0) return ORC_EBADARG;
    const int H = N / 2 + 1;
    const i64 nh = (i64)H * N * N;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nh);
    if (!buf) return ORC_ENOMEM;
    fft_plan1 P;
    if (plan1_init(&P, N)) { free(buf); return ORC_ENOMEM; }
    const i64 R = H + 2 * kh, KG = (i64)N + 2 * kh;
    const i64 geo[2] = {H, N};
    for (int c = 0; c < C; ++c) {
        /* buf[kx + H (ky + N kz)] */
#pragma omp parallel for schedule(static)
        for (int ky = 0; ky < N; ++ky)
            for (int kz = 0; kz < N; ++kz)
                for (int kx = 0; kx < H; ++kx) {
                    const i64 b = kx + (i64)H * (ky + (i64)N * kz);
                    const i64 o = ((kx + kh) + R * ((kz + kh) + KG * (ky + kh))) * C + c;
                    buf[2 * b] = k[2 * o];
                    buf[2 * b + 1] = k[2 * o + 1];
                }
        fft_lines(&P, buf, (i64)H * N, base_ax1, geo, H, +1);
        fft_lines(&P, buf, (i64)H * N, base_ax2, geo, (i64)H * N, +1);
#pragma omp parallel
        {
            double *line = (double *)malloc(sizeof(double) * 4 * (size_t)N);
            double *tmp = line + 2 * N;
#pragma omp for schedule(static)
            for (i64 row = 0; row < (i64)N * N; ++row) {
                const int y = (int)(row % N), z = (int)(row / N);
                const double *src = buf + 2 * ((i64)H * row);
                for (int kx = 0; kx < H; ++kx) {
                    line[2 * kx] = src[2 * kx];
                    line[2 * kx + 1] = src[2 * kx + 1];
                }
                line[1] = 0.0;                      /* DC: real part only */
                if (N % 2 == 0) line[2 * (N / 2) + 1] = 0.0; /* Nyquist: real part only */
                for (int kx = 1; kx < H; ++kx) {
                    const int m = N - kx;
                    if (m < H) continue;
                    line[2 * m] = src[2 * kx];
                    line[2 * m + 1] = -src[2 * kx + 1];
                }
                fft1(&P, line, +1, tmp);
                for (int x = 0; x < N; ++x) r[sidx(N, rh, x, y, z) * C + c] = line[2 * x];
            }
            free(line);
        }
    }
    plan1_free(&P);
    free(buf);
    return 0;
}

/* ---------------------------------------------------------------------------
 * C2C PlanFFT (two Field<Imag>), single rank.  Complex fields are interleaved
 * (re, im) doubles, C components per site.
 * ------------------------------------------------------------------------- */

/* one FFTW-style "many" 1-D transform: `howmany` lines of n points, element e of line
 * m at in[idist m + istride e], written to out[odist m + ostride e] (complex units) */
static void many_dft(const fft_plan1 *P, const double *in, i64 istride, i64 idist, double *out, i64 ostride,
                     i64 odist, i64 howmany, int sign) {
    const int n = P->n;
    double *line = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    double *tmp = line + 2 * n;
    for (i64 m = 0; m < howmany; ++m) {
        for (int e = 0; e < n; ++e) {
            line[2 * e] = in[2 * (idist * m + istride * e)];
            line[2 * e + 1] = in[2 * (idist * m + istride * e) + 1];
        }
        fft1(P, line, sign, tmp);
        for (int e = 0; e < n; ++e) {
            out[2 * (odist * m + ostride * e)] = line[2 * e];
            out[2 * (odist * m + ostride * e) + 1] = line[2 * e + 1];
        }
    }
    free(line);
}

/* transpose_0_2 / transpose_1_2 (LATfield2_PlanFFT_decl.hpp:1396-1466) */
static void transpose_0_2(const double *in, double *out, i64 di, i64 dj, i64 dk) {
    for (i64 i = 0; i < di; ++i)
        for (i64 j = 0; j < dj; ++j)
            for (i64 k = 0; k < dk; ++k) {
                out[2 * (k + dk * (j + i * dj))] = in[2 * (i + di * (j + k * dj))];
                out[2 * (k + dk * (j + i * dj)) + 1] = in[2 * (i + di * (j + k * dj)) + 1];
            }
}
static void transpose_1_2(const double *in, double *out, i64 di, i64 dj, i64 dk) {
    for (i64 i = 0; i < di; ++i)
        for (i64 j = 0; j < dj; ++j)
            for (i64 k = 0; k < dk; ++k) {
                out[2 * (i + di * (k + j * dk))] = in[2 * (i + di * (j + k * dj))];
                out[2 * (i + di * (k + j * dk)) + 1] = in[2 * (i + di * (j + k * dj)) + 1];
            }
}

/* PlanFFT::execute(FFT_FORWARD), C2C (LATfield2_PlanFFT_decl.hpp:1061-1133), the steps of
 * one rank: fPlan_i_ along x into temp [x][z][y], transpose_0_2, fPlan_j_ along the slowest
 * axis (y), transpose_1_2, fPlan_k_ along z into the k-field with kz on storage dim 0, kx
 * on dim 1, ky on dim 2 (plans at :678-680).  The reference steps the k-field with the
 * position lattice's jumps (rJump_); the k lattice's own jumps are used here, which is the
 * layout cKSite reads (LATfield2_Site.hpp:219-220) and the same when rh == kh. */
int orc_fft_c2c_forward(const double *f, double *k, int N, int C, int rh, int kh) {
    if (!f || !k || N < 1 || C < 1 || rh < 0 || kh < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N, N2 = (i64)N * N;
    double *temp = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    double *temp1 = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    fft_plan1 P;
    if (!temp || !temp1 || plan1_init(&P, N)) { free(temp); free(temp1); return ORC_ENOMEM; }
    const i64 G = (i64)N + 2 * rh, KG = (i64)N + 2 * kh;
    const double *cData = f + 2 * (i64)rh * (1 + G + G * G) * C;
    double *kData = k + 2 * (i64)kh * (1 + KG + KG * KG) * C;
    for (int comp = 0; comp < C; ++comp) {
        for (int l = 0; l < N; ++l)
            many_dft(&P, cData + 2 * (G * G * l * C + comp), C, G * C, temp + 2 * ((i64)l * N), N2, 1, N, -1);
        memcpy(temp1, temp, sizeof(double) * 2 * (size_t)n3);
        transpose_0_2(temp1, temp, N, N, N);
        many_dft(&P, temp, N2, 1, temp, N2, 1, N2, -1);
        memcpy(temp1, temp, sizeof(double) * 2 * (size_t)n3);
        transpose_1_2(temp1, temp, N, N, N);
        for (int l = 0; l < N; ++l)
            many_dft(&P, temp + 2 * ((i64)l * N), N2, 1, kData + 2 * (KG * KG * l * C + comp), C, KG * C, N, -1);
    }
    plan1_free(&P);
    free(temp);
    free(temp1);
    return 0;
}

/* The inverse of orc_fft_c2c_forward: the unnormalised e^{+2 pi i k.x/N} DFT from the
 * k-field (mode (kx, ky, kz) at storage (kz, kx, ky)) into the complex interior, so
 * backward(forward(f)) = N^3 f.  This is the published inverse DFT, not the reference's
 * backward (see orc_fft_c2c_backward_reference). */
int orc_fft_c2c_backward(const double *k, double *f, int N, int C, int rh, int kh) {
    if (!f || !k || N < 1 || C < 1 || rh < 0 || kh < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    fft_plan1 P;
    if (!buf || plan1_init(&P, N)) { free(buf); return ORC_ENOMEM; }
    const i64 geo[2] = {N, N};
    for (int c = 0; c < C; ++c) {
        /* buf[kx + N (ky + N kz)] */
        for (int ky = 0; ky < N; ++ky)
            for (int kx = 0; kx < N; ++kx)
                for (int kz = 0; kz < N; ++kz) {
                    const i64 b = kx + (i64)N * (ky + (i64)N * kz);
                    const i64 o = sidx(N, kh, kz, kx, ky) * C + c;
                    buf[2 * b] = k[2 * o];
                    buf[2 * b + 1] = k[2 * o + 1];
                }
        fft_lines(&P, buf, (i64)N * N, base_ax0, geo, 1, +1);
        fft_lines(&P, buf, (i64)N * N, base_ax1, geo, N, +1);
        fft_lines(&P, buf, (i64)N * N, base_ax2, geo, (i64)N * N, +1);
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) {
                    const i64 b = x + (i64)N * (y + (i64)N * z);
                    const i64 o = sidx(N, rh, x, y, z) * C + c;
                    f[2 * o] = buf[2 * b];
                    f[2 * o + 1] = buf[2 * b + 1];
                }
    }
    plan1_free(&P);
    free(buf);
    return 0;
}

/* PlanFFT::execute(FFT_BACKWARD), C2C, as the reference runs it on one rank
 * (LATfield2_PlanFFT_decl.hpp:1135-1196), with one ghost width h for both fields:
 * bPlan_k_ along kz into temp [z][ky][kx]; bPlan_j_ along the slowest axis, which is z
 * again because the transposition before it is commented out (:1165-1167); transpose_0_2;
 * bPlan_i_ along kx into position storage (x, l, m) <- temp line (l, m) = (z, ky). */
int orc_fft_c2c_backward_reference(const double *k, double *f, int N, int C, int h) {
    if (!f || !k || N < 1 || C < 1 || h < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N, N2 = (i64)N * N;
    double *temp = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    double *temp1 = (double *)malloc(sizeof(double) * 2 * (size_t)n3);
    fft_plan1 P;
    if (!temp || !temp1 || plan1_init(&P, N)) { free(temp); free(temp1); return ORC_ENOMEM; }
    const i64 G = (i64)N + 2 * h;
    const double *kData = k + 2 * (i64)h * (1 + G + G * G) * C;
    double *cData = f + 2 * (i64)h * (1 + G + G * G) * C;
    for (int comp = 0; comp < C; ++comp) {
        for (int l = 0; l < N; ++l)
            many_dft(&P, kData + 2 * (G * G * l * C + comp), C, G * C, temp + 2 * ((i64)l * N), N2, 1, N, +1);
        many_dft(&P, temp, N2, 1, temp, N2, 1, N2, +1);
        memcpy(temp1, temp, sizeof(double) * 2 * (size_t)n3);
        transpose_0_2(temp1, temp, N, N, N);
        for (int l = 0; l < N; ++l)
            many_dft(&P, temp + 2 * ((i64)l * N), N2, 1, cData + 2 * (G * G * l * C + comp), C, G * C, N, +1);
    }
    plan1_free(&P);
    free(temp);
    free(temp1);
    return 0;
}

/* examples/poissonSolver.cpp:99-111 (zero mode :100-104), k coordinates in the rKSite
 * sense: kx = lattice dim 0, ky = dim 2, kz = dim 1 (LATfield2_Site.hpp:319-320). */
int orc_poisson_green(double *k, int N, int C, int kh, double res, double scale) {
    if (!k || N < 1 || C < 1 || kh < 0 || !(res > 0)) return ORC_EBADARG;
    const int H = N / 2 + 1;
    const i64 R = H + 2 * kh, KG = (i64)N + 2 * kh;
    const double res2 = res * res;
#pragma omp parallel for schedule(static)
    for (int ky = 0; ky < N; ++ky)
        for (int kz = 0; kz < N; ++kz)
            for (int kx = 0; kx < H; ++kx) {
                const i64 o = ((kx + kh) + R * ((kz + kh) + KG * (ky + kh))) * C;
                if (kx == 0 && ky == 0 && kz == 0) {
                    for (int c = 0; c < C; ++c) k[2 * (o + c)] = k[2 * (o + c) + 1] = 0.0;
                    continue;
                }
                const double lam = 2.0 * (cos(2.0 * M_PI * kx / N) + cos(2.0 * M_PI * ky / N) +
                                          cos(2.0 * M_PI * kz / N) - 3.0) / res2;
                for (int c = 0; c < C; ++c) {
                    k[2 * (o + c)] = k[2 * (o + c)] / lam * scale;
                    k[2 * (o + c) + 1] = k[2 * (o + c) + 1] / lam * scale;
                }
            }
    return 0;
}

int orc_force_field(const double *phi, double *F, int N, int h) {
    if (!phi || !F || N < 1 || h < 1) return ORC_EBADARG;
    const double half_n = 0.5 * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const i64 s = sidx(N, h, x, y, z);
                F[3 * s + 0] = -(phi[sidx(N, h, x + 1, y, z)] - phi[sidx(N, h, x - 1, y, z)]) * half_n;
                F[3 * s + 1] = -(phi[sidx(N, h, x, y + 1, z)] - phi[sidx(N, h, x, y - 1, z)]) * half_n;
                F[3 * s + 2] = -(phi[sidx(N, h, x, y, z + 1)] - phi[sidx(N, h, x, y, z - 1)]) * half_n;
            }
    return orc_ghost_fill(F, N, h, 3);
}

/* Particles::updateVel driver (particles/LATfield2_Particles.hpp:753-893) with the CIC
 * gather + kick functor: frac = pos*N - cell (:825-828), corner weights as in the scalar
 * deposit, v += dt * F(pos).  Returns max v^2 (the driver returns its square root). */
int orc_gather_kick(const double *pos, double *vel, i64 np, const double *F, int N, int h,
                    double dt_kick, double *acc, double *vmax2) {
    if (!pos || !vel || !F || np < 0 || N < 1 || h < 1) return ORC_EBADARG;
    double vm = 0.0;
#pragma omp parallel for schedule(static) reduction(max : vm)
    for (i64 p = 0; p < np; ++p) {
        int c3[3];
        double fr[3];
        int bad = 0;
        for (int d = 0; d < 3; ++d) {
            c3[d] = cell_coord(pos[3 * p + d], N);
            if (c3[d] < 0) bad = 1;
            fr[d] = pos[3 * p + d] * N - c3[d];
        }
        double a[3] = {0, 0, 0};
        if (!bad) {
            for (int cb = 0; cb < 8; ++cb) {
                const int dx = (cb >> 2) & 1, dy = (cb >> 1) & 1, dz = cb & 1;
                const double w = (dx ? fr[0] : 1 - fr[0]) * (dy ? fr[1] : 1 - fr[1]) * (dz ? fr[2] : 1 - fr[2]);
                const i64 s = sidx(N, h, c3[0] + dx, c3[1] + dy, c3[2] + dz);
                for (int d = 0; d < 3; ++d) a[d] += w * F[3 * s + d];
            }
        }
        double v2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            vel[3 * p + d] += dt_kick * a[d];
            v2 += vel[3 * p + d] * vel[3 * p + d];
            if (acc) acc[3 * p + d] = a[d];
        }
        if (v2 > vm) vm = v2;
    }
    if (vmax2) *vmax2 = vm;
    return 0;
}

/* move_particles_simple drift (particles/move_function.hpp:30) and the sequential
 * periodic wrap of Particles::moveParticles (particles/LATfield2_Particles.hpp:2028-2032) */
int orc_drift(double *pos, const double *vel, i64 np, double dt) {
    if (!pos || !vel || np < 0) return ORC_EBADARG;
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < np; ++p)
        for (int d = 0; d < 3; ++d) {
            double x = pos[3 * p + d] + dt * vel[3 * p + d];
            if (x < 0) x += 1.0;
            if (x >= 1.0) x -= 1.0;
            pos[3 * p + d] = x;
        }
    return 0;
}

int orc_pm_step(double *pos, double *vel, const double *mass, double mass_global, i64 np,
                int N, double dt_kick, double dt_drift, double *rho, double *phi, double *acc,
                double *stage_ms) {
    if (!pos || !vel || !rho || !phi || np < 0 || N < 2) return ORC_EBADARG;
    const int h = 1;
    const i64 G = N + 2;
    const i64 gross = G * G * G;
    const int H = N / 2 + 1;
    double t[9];
    int rc;
    double *kf = (double *)malloc(sizeof(double) * 2 * (size_t)H * N * N);
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)gross);
    if (!kf || !F) { free(kf); free(F); return ORC_ENOMEM; }
    t[0] = now_ms();
    orc_zero(rho, gross);
    if ((rc = orc_scalar_project(pos, mass, mass_global, np, rho, N, h))) goto out;
    t[1] = now_ms();
    orc_scalar_comm(rho, N, h);
    t[2] = now_ms();
    {
        double mean = 0.0;
#pragma omp parallel for reduction(+ : mean) schedule(static)
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) mean += rho[sidx(N, h, x, y, z)];
        mean /= (double)N * N * N;
#pragma omp parallel for schedule(static)
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) rho[sidx(N, h, x, y, z)] -= mean;
    }
    t[3] = now_ms();
    if ((rc = orc_fft_r2c(rho, kf, N, 1, h, 0))) goto out;
    t[4] = now_ms();
    orc_poisson_green(kf, N, 1, 0, 1.0 / N, 1.0 / ((double)N * N * N));
    t[5] = now_ms();
    if ((rc = orc_fft_c2r(kf, phi, N, 1, h, 0))) goto out;
    t[6] = now_ms();
    orc_ghost_fill(phi, N, h, 1);
    orc_force_field(phi, F, N, h);
    t[7] = now_ms();
    orc_gather_kick(pos, vel, np, F, N, h, dt_kick, acc, NULL);
    orc_drift(pos, vel, np, dt_drift);
    t[8] = now_ms();
    if (stage_ms)
        for (int i = 0; i < 8; ++i) stage_ms[i] = t[i + 1] - t[i];
out:
    free(kf);
    free(F);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Differentiable PM: forward with checkpoints and the hand-derived adjoint.
 * Grids here are periodic N^3 arrays without ghosts, index x + N(y + N z).
 * ------------------------------------------------------------------------- */
static inline void cic_cell(double p, int N, int *c, double *u) {
    const double X = p * N;
    const double f = floor(X);
    *u = X - f;
    int ci = (int)f % N;
    if (ci < 0) ci += N;
    *c = ci;
}

static inline i64 gidx(int N, int x, int y, int z) {
    x = (x % N + N) % N;
    y = (y % N + N) % N;
    z = (z % N + N) % N;
    return x + (i64)N * (y + (i64)N * z);
}

/* value and gradient (per cell unit, d/dX with X = pos*N) of the CIC interpolant of g */
static double cic_interp_grad(const double *g, int N, const double *pos, double grad[3]) {
    int c[3];
    double u[3];
    for (int d = 0; d < 3; ++d) cic_cell(pos[d], N, &c[d], &u[d]);
    double val = 0.0;
    if (grad) grad[0] = grad[1] = grad[2] = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const double wx = i ? u[0] : 1 - u[0], wy = j ? u[1] : 1 - u[1], wz = k ? u[2] : 1 - u[2];
                const double gx = i ? 1.0 : -1.0, gy = j ? 1.0 : -1.0, gz = k ? 1.0 : -1.0;
                const double v = g[gidx(N, c[0] + i, c[1] + j, c[2] + k)];
                val += wx * wy * wz * v;
                if (grad) {
                    grad[0] += gx * wy * wz * v;
                    grad[1] += wx * gy * wz * v;
                    grad[2] += wx * wy * gz * v;
                }
            }
    return val;
}

static void cic_add(double *g, int N, const double *pos, double q) {
    int c[3];
    double u[3];
    for (int d = 0; d < 3; ++d) cic_cell(pos[d], N, &c[d], &u[d]);
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const double w = (i ? u[0] : 1 - u[0]) * (j ? u[1] : 1 - u[1]) * (k ? u[2] : 1 - u[2]);
                g[gidx(N, c[0] + i, c[1] + j, c[2] + k)] += q * w;
            }
}

int orc_cic_density(const double *pos, i64 np, int N, double mass, double *out) {
    if (!pos || !out || np < 0 || N < 2) return ORC_EBADARG;
    memset(out, 0, sizeof(double) * (size_t)N * N * N);
    const double q = mass * (double)N * N * N;
    for (i64 p = 0; p < np; ++p) cic_add(out, N, pos + 3 * p, q);
    return 0;
}

/* phi = G rho: mean removed, -1/K^2 per mode, zero mode 0, normalised (the PM Poisson solve).
 * G is symmetric, so it is also its own adjoint. */
static int poisson_periodic(const double *rho, double *phi, int N) {
    const i64 n3 = (i64)N * N * N;
    double *f = (double *)calloc((size_t)(N + 2) * (N + 2) * (N + 2), sizeof(double));
    if (!f) return ORC_ENOMEM;
    double mean = 0.0;
    for (i64 i = 0; i < n3; ++i) mean += rho[i];
    mean /= (double)n3;
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) f[sidx(N, 1, x, y, z)] = rho[gidx(N, x, y, z)] - mean;
    const int H = N / 2 + 1;
    double *k = (double *)malloc(sizeof(double) * 2 * (size_t)H * N * N);
    if (!k) { free(f); return ORC_ENOMEM; }
    int rc = orc_fft_r2c(f, k, N, 1, 1, 0);
    if (!rc) rc = orc_poisson_green(k, N, 1, 0, 1.0 / N, 1.0 / (double)n3);
    if (!rc) rc = orc_fft_c2r(k, f, N, 1, 1, 0);
    if (!rc)
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) phi[gidx(N, x, y, z)] = f[sidx(N, 1, x, y, z)];
    free(k);
    free(f);
    return rc;
}

/* F_d = -(phi(n+e_d) - phi(n-e_d)) N/2, three N^3 arrays */
static void force_periodic(const double *phi, double *F, int N) {
    const i64 n3 = (i64)N * N * N;
    const double hn = 0.5 * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const i64 i = gidx(N, x, y, z);
                F[i] = -(phi[gidx(N, x + 1, y, z)] - phi[gidx(N, x - 1, y, z)]) * hn;
                F[n3 + i] = -(phi[gidx(N, x, y + 1, z)] - phi[gidx(N, x, y - 1, z)]) * hn;
                F[2 * n3 + i] = -(phi[gidx(N, x, y, z + 1)] - phi[gidx(N, x, y, z - 1)]) * hn;
            }
}

static int step_forces(const double *pos, i64 np, int N, double mass, double *rho, double *phi, double *F) {
    orc_cic_density(pos, np, N, mass, rho);
    int rc = poisson_periodic(rho, phi, N);
    if (rc) return rc;
    force_periodic(phi, F, N);
    return 0;
}

static void wrap01(double *x) {
    if (*x < 0) *x += 1.0;
    if (*x >= 1.0) *x -= 1.0;
    if (*x < 0 || *x >= 1.0) *x -= floor(*x);
}

int orc_diffpm_forward(double *pos, double *vel, i64 np, int N, double mass, double dt, int nsteps, double *ckpt) {
    if (!pos || !vel || np < 0 || N < 2 || nsteps < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N;
    double *rho = (double *)malloc(sizeof(double) * (size_t)n3);
    double *phi = (double *)malloc(sizeof(double) * (size_t)n3);
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)n3);
    if (!rho || !phi || !F) { free(rho); free(phi); free(F); return ORC_ENOMEM; }
    int rc = 0;
    for (int s = 0; s < nsteps && !rc; ++s) {
        if (ckpt) {
            memcpy(ckpt + (i64)s * 6 * np, pos, sizeof(double) * 3 * (size_t)np);
            memcpy(ckpt + (i64)s * 6 * np + 3 * np, vel, sizeof(double) * 3 * (size_t)np);
        }
        rc = step_forces(pos, np, N, mass, rho, phi, F);
        if (rc) break;
#pragma omp parallel for schedule(static)
        for (i64 p = 0; p < np; ++p)
            for (int d = 0; d < 3; ++d) {
                const double a = cic_interp_grad(F + (i64)d * n3, N, pos + 3 * p, NULL);
                vel[3 * p + d] += dt * a;
            }
#pragma omp parallel for schedule(static)
        for (i64 p = 0; p < np; ++p)
            for (int d = 0; d < 3; ++d) {
                pos[3 * p + d] += dt * vel[3 * p + d];
                wrap01(&pos[3 * p + d]);
            }
    }
    free(rho);
    free(phi);
    free(F);
    return rc;
}

static double density_loss(const double *pos, i64 np, int N, double mass, const double *target, double *resid) {
    const i64 n3 = (i64)N * N * N;
    orc_cic_density(pos, np, N, mass, resid);
    double L = 0.0;
    for (i64 i = 0; i < n3; ++i) {
        resid[i] -= target[i];
        L += resid[i] * resid[i];
    }
    return L;
}

int orc_diffpm_loss(const double *pos0, const double *vel0, i64 np, int N, double mass, double dt, int nsteps,
                    const double *target, double *loss) {
    if (!pos0 || !vel0 || !target || !loss) return ORC_EBADARG;
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *vel = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *res = (double *)malloc(sizeof(double) * (size_t)N * N * N);
    if (!pos || !vel || !res) { free(pos); free(vel); free(res); return ORC_ENOMEM; }
    memcpy(pos, pos0, sizeof(double) * 3 * (size_t)np);
    memcpy(vel, vel0, sizeof(double) * 3 * (size_t)np);
    int rc = orc_diffpm_forward(pos, vel, np, N, mass, dt, nsteps, NULL);
    if (!rc) *loss = density_loss(pos, np, N, mass, target, res);
    free(pos);
    free(vel);
    free(res);
    return rc;
}

int orc_diffpm_loss_grad(const double *pos0, const double *vel0, i64 np, int N, double mass, double dt, int nsteps,
                         const double *target, double *loss, double *grad_pos, double *grad_vel) {
    if (!pos0 || !vel0 || !target || !loss || !grad_pos || np < 0 || N < 2 || nsteps < 0) return ORC_EBADARG;
    const i64 n3 = (i64)N * N * N;
    const double q = mass * (double)n3;
    double *ck = (double *)malloc(sizeof(double) * 6 * (size_t)np * (size_t)(nsteps > 0 ? nsteps : 1));
    double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *vel = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *lx = grad_pos;
    double *lv = (double *)calloc(3 * (size_t)np, sizeof(double));
    double *g = (double *)malloc(sizeof(double) * 5 * (size_t)n3);    /* resid/rho, phi, rhobar, phibar + spare */
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)n3);
    double *Fb = (double *)malloc(sizeof(double) * 3 * (size_t)n3);
    if (!ck || !pos || !vel || !lv || !g || !F || !Fb) {
        free(ck); free(pos); free(vel); free(lv); free(g); free(F); free(Fb);
        return ORC_ENOMEM;
    }
    double *rho = g, *phi = g + n3, *rhob = g + 2 * n3, *phib = g + 3 * n3;
    memcpy(pos, pos0, sizeof(double) * 3 * (size_t)np);
    memcpy(vel, vel0, sizeof(double) * 3 * (size_t)np);
    int rc = orc_diffpm_forward(pos, vel, np, N, mass, dt, nsteps, ck);
    if (rc) goto out;
    /* dL/dx_final: the position derivative of the deposit against 2 (rho - target) */
    *loss = density_loss(pos, np, N, mass, target, rho);
    for (i64 i = 0; i < n3; ++i) rho[i] *= 2.0;
#pragma omp parallel for schedule(static)
    for (i64 p = 0; p < np; ++p) {
        double gr[3];
        cic_interp_grad(rho, N, pos + 3 * p, gr);
        for (int d = 0; d < 3; ++d) lx[3 * p + d] = q * N * gr[d];
    }
    for (int s = nsteps - 1; s >= 0; --s) {
        const double *xs = ck + (i64)s * 6 * np;
        /* x_{s+1} = x_s + dt v_{s+1}: the velocity adjoint collects dt * lambda_x */
        for (i64 i = 0; i < 3 * np; ++i) lv[i] += dt * lx[i];
        /* mu = dt * lambda_v is the adjoint of the acceleration a_s(x_s) */
        if ((rc = step_forces(xs, np, N, mass, rho, phi, F))) goto out;
        memset(Fb, 0, sizeof(double) * 3 * (size_t)n3);
        for (i64 p = 0; p < np; ++p) {
            const double *x = xs + 3 * p;
            for (int d = 0; d < 3; ++d) {
                const double mu = dt * lv[3 * p + d];
                double gr[3];
                /* direct term: the interpolation weights move with the particle */
                cic_interp_grad(F + (i64)d * n3, N, x, gr);
                for (int e = 0; e < 3; ++e) lx[3 * p + e] += mu * N * gr[e];
                /* field term: a = I(x) F, so F-bar = I(x)^T mu (a CIC deposit of mu) */
                cic_add(Fb + (i64)d * n3, N, x, mu);
            }
        }
        /* F_d = -D_d phi with D_d antisymmetric: phi-bar = sum_d D_d F-bar_d */
        const double hn = 0.5 * N;
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x)
                    phib[gidx(N, x, y, z)] = (Fb[gidx(N, x + 1, y, z)] - Fb[gidx(N, x - 1, y, z)]) * hn +
                                             (Fb[n3 + gidx(N, x, y + 1, z)] - Fb[n3 + gidx(N, x, y - 1, z)]) * hn +
                                             (Fb[2 * n3 + gidx(N, x, y, z + 1)] - Fb[2 * n3 + gidx(N, x, y, z - 1)]) * hn;
        /* rho-bar = G^T phi-bar = G phi-bar; then the deposit's position derivative */
        if ((rc = poisson_periodic(phib, rhob, N))) goto out;
#pragma omp parallel for schedule(static)
        for (i64 p = 0; p < np; ++p) {
            double gr[3];
            cic_interp_grad(rhob, N, xs + 3 * p, gr);
            for (int d = 0; d < 3; ++d) lx[3 * p + d] += q * N * gr[d];
        }
    }
    if (grad_vel) memcpy(grad_vel, lv, sizeof(double) * 3 * (size_t)np);
out:
    free(ck); free(pos); free(vel); free(lv); free(g); free(F); free(Fb);
    return rc;
}
