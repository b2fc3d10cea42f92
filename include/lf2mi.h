/*
 * lf2mi.h -- C ABI of the MI355X-native LATfield2 particle-mesh path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (LF2MI_OK = 0, negative on error; lf2mi_last_error() gives the
 * message) and is stream-ordered: `stream` is a hipStream_t passed as void*
 * (NULL = the default stream).  Field and k-field pointers are DEVICE pointers.
 * Host pointers appear only in the explicit upload/download calls.
 *
 * Geometry (box length 1, cubic lattice of N sites, resolution 1/N):
 *   real field, C components, ghost width h, G = N + 2h  (Field<Real>):
 *     data[((x+h) + G*((y+h) + G*(z+h)))*C + c]
 *       -- LATfield2_Lattice.hpp:111-113 (strides), LATfield2_Field_decl.hpp:760-807
 *   k-field of the R2C transform, ghost width kh, R = N/2+1+2kh, KG = N+2kh
 *   (Field<Imag> on Lattice::initializeRealFFT, dims (kx, kz, ky)):
 *     k[((kx+kh) + R*((kz+kh) + KG*(ky+kh)))*C + c]   (complex = {re, im})
 *       -- LATfield2_PlanFFT_decl.hpp:1469-1510, LATfield2_Lattice.hpp:225-259
 *   Symmetric 3x3 fields use C = 6 in the order (0,0)(0,1)(0,2)(1,1)(1,2)(2,2)
 *       -- LATfield2_Field_decl.hpp:773-788
 * Precision is a run-time argument: LF2MI_F32 (float / float2) or LF2MI_F64.
 * The library is not thread-safe; use one host thread per device.
 */
#ifndef LF2MI_H
#define LF2MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LF2MI_OK 0
#define LF2MI_EINVAL (-1)      /* bad argument */
#define LF2MI_ENOMEM (-2)      /* device or host allocation failed */
#define LF2MI_EHIP (-3)        /* HIP runtime error */
#define LF2MI_ERANGE (-4)      /* size outside the supported range */
#define LF2MI_EUNSUPPORTED (-5)

/* precision */
#define LF2MI_F32 0
#define LF2MI_F64 1

/* FFT directions -- LATfield2_PlanFFT.hpp:14-15 */
#define LF2MI_FFT_FORWARD 1
#define LF2MI_FFT_BACKWARD (-1)

/* projection kinds */
#define LF2MI_PROJ_SCALAR 0 /* scalarProjectionCIC_*  (C = 1) */
#define LF2MI_PROJ_VECTOR 1 /* vectorProjectionCIC_*  (C = 3) */
#define LF2MI_PROJ_VECVEC 2 /* VecVecProjectionCIC_*  (C = 6) */
#define LF2MI_PROJ_VECTOR_NGP 3    /* vectorProjectionCICNGP_*    (C = 3) */
#define LF2MI_PROJ_SYMTENSOR_NGP 4 /* symtensorProjectionCICNGP_* (C = 6) */

/* deposit target */
#define LF2MI_DEP_UNFOLDED 0 /* like *_project: adds into interior and ghost nodes */
#define LF2MI_DEP_PERIODIC 1 /* like *_project then *_comm: adds into interior nodes */
/* projection_init, *_project and *_comm in one call: the interior receives the periodic
 * deposit, the ghosts zero (no counterpart in the reference, which zeroes the whole field
 * first: the FP32 deposits then store the interior instead of reading it back) */
#define LF2MI_DEP_PERIODIC_SET 2

/* deposit accumulation */
#define LF2MI_ACC_NATIVE 0  /* atomics in the container's precision */
#define LF2MI_ACC_FIXED64 1 /* int64 fixed point: run-to-run reproducible (see lf2mi_deposit) */

/* particle container flags */
#define LF2MI_PARTS_ID 1   /* keep 64-bit particle IDs (part_simple::ID) */
#define LF2MI_PARTS_MASS 2 /* per-particle mass (FROM_PART) instead of one global mass */

/* ------------------------------------------------------------------ library */
int lf2mi_version(void);
/* message of the last failing call on this host thread ("" if none) */
const char *lf2mi_last_error(void);
/* select and initialise a device for the calling thread */
int lf2mi_device_init(int device);
/* hipDeviceSynchronize with status */
int lf2mi_synchronize(void);

/* device memory for callers that do not link the HIP runtime themselves (the C++ Field
 * and Particles mirrors allocate through these, as Field::alloc does with new[] in
 * LATfield2_Field_decl.hpp:696-712).  lf2mi_copy synchronises for host transfers. */
#define LF2MI_COPY_H2D 0
#define LF2MI_COPY_D2H 1
#define LF2MI_COPY_D2D 2
int lf2mi_device_alloc(void **ptr, int64_t bytes);
int lf2mi_device_free(void *ptr);
int lf2mi_copy(void *dst, const void *src, int64_t bytes, int kind, void *stream);

/* Per-kernel timing (no counterpart in the reference, whose timer.hpp:64-177 and the
 * MPI_Wtime brackets of benchmarks_tests/benchmarks_particles.cpp:189-236 time stages
 * on the host).  When enabled, every kernel launch records a pair of HIP events on the
 * stream it is launched on; lf2mi_profile_get synchronises on pending events and
 * returns, for slot `index`, the kernel name, the summed duration and the launch count
 * (LF2MI_ERANGE past the last slot).  lf2mi_profile_reset clears all slots. */
int lf2mi_profile_enable(int on);
int lf2mi_profile_reset(void);
int lf2mi_profile_get(int index, char *name, int name_cap, double *total_ms, int64_t *count);

/* ------------------------------------------------------------------ fields */
/* projection_init: zero `count` values (particles/projections.hpp:31-34) */
int lf2mi_zero(void *ptr, int64_t count, int dtype, void *stream);

/* Field::updateHalo: periodic ghost fill of all C components
 * (LATfield2_Field_decl.hpp:889-935); bit-exact copies */
int lf2mi_ghost_fill(void *field, int N, int halo, int C, int dtype, void *stream);
/* the same on an nx x ny x nz lattice (Field::updateHalo takes any 3-D lattice; the
 * storage index is (x+h) + (nx+2h)((y+h) + (ny+2h)(z+h))); halo <= min(nx, ny, nz) */
int lf2mi_ghost_fill_box(void *field, int nx, int ny, int nz, int halo, int C, int dtype, void *stream);

/* the ghost folds after a deposit, in the reference's order (x, then y, then z):
 *   LF2MI_PROJ_SCALAR: scalarProjectionCIC_comm (particles/projections.hpp:194-300)
 *   LF2MI_PROJ_VECTOR: vectorProjectionCIC_comm (particles/projections.hpp:1192-1351)
 *   LF2MI_PROJ_VECVEC: VecVecProjectionCIC_comm (particles/projections.hpp:1624-1814)
 *   LF2MI_PROJ_VECTOR_NGP: vectorProjectionCICNGP_comm (particles/projections.hpp:558-681)
 *   LF2MI_PROJ_SYMTENSOR_NGP: symtensorProjectionCICNGP_comm (particles/projections.hpp:856-976)
 *     (the two CIC-NGP folds are the scalar fold on each component: upper ghosts only) */
int lf2mi_projection_comm(void *field, int N, int halo, int kind, int dtype, void *stream);

/* sum over the interior of component `comp` in FP64 (parallel.sum of the driver loop in
 * examples/poissonSolver.cpp:88-94); *out is written on the host, the call synchronises */
int lf2mi_field_sum(const void *field, int N, int halo, int C, int comp, int dtype,
                    double *out, void *stream);

/* field[interior, comp] = field * mul + add  (driver-side arithmetic of the Poisson
 * example: mean subtraction, 1/N^3 normalisation) */
int lf2mi_field_affine(void *field, int N, int halo, int C, int comp, double mul, double add,
                       int dtype, void *stream);

/* ------------------------------------------------------------------ FFT */
typedef struct lf2mi_fft_plan_s *lf2mi_fft_plan;

/* PlanFFT<Imag>(Field<Real>*, Field<Imag>*) for the R2C case
 * (LATfield2_PlanFFT_decl.hpp:731-846): N in [8, 1024], 2^k, 3 * 2^k (12 ... 768) or
 * 5 * 2^k (10 ... 640), C >= 1
 * components, real ghost width rhalo, k ghost width khalo.  Owns its twiddle tables
 * and workspace on the current device. */
int lf2mi_fft_plan_create(int N, int C, int rhalo, int khalo, int dtype, lf2mi_fft_plan *plan);
/* PlanFFT<Imag>(Field<Imag>*, Field<Imag>*) for the C2C case
 * (LATfield2_PlanFFT_decl.hpp:608-719): same sizes; the k-field lattice is N^3
 * (Lattice::initializeComplexFFT, LATfield2_Lattice.hpp:260-297) and holds mode
 * (kx, ky, kz) at storage (kz, kx, ky) as cKSite reads it (LATfield2_Site.hpp:219-220). */
int lf2mi_fft_plan_create_c2c(int N, int C, int rhalo, int khalo, int dtype, lf2mi_fft_plan *plan);
int lf2mi_fft_plan_destroy(lf2mi_fft_plan plan);
/* bytes of device memory the plan holds */
int lf2mi_fft_plan_bytes(lf2mi_fft_plan plan, int64_t *bytes);

/* PlanFFT::execute (LATfield2_PlanFFT_decl.hpp:851-1056).  FORWARD: unnormalised
 * e^{-2 pi i k.x/N} from the real interior into the k-field interior (input kept).
 * BACKWARD: unnormalised e^{+}, k-field kept, real interior written (ghosts untouched);
 * only the real parts of the kx = 0 and kx = N/2 modes are used, as in FFTW's c2r.
 * Any other direction is a no-op, as in the reference.
 * C2C plans (:1059-1198): FORWARD is the unnormalised e^{-} DFT of the complex interior
 * into the k-field interior (input kept); BACKWARD is the unnormalised e^{+} DFT back
 * (k-field kept), so BACKWARD(FORWARD(f)) = N^3 f.  The reference's single-rank backward
 * differs: it skips the transposition before its second pass (:1165-1167), so it applies
 * the z transform twice and no y transform. */
int lf2mi_fft_execute(lf2mi_fft_plan plan, int direction, void *rfield, void *kfield, void *stream);

/* Green's-function multiply of examples/poissonSolver.cpp:99-111 on a k-field:
 * k *= scale / lambda, lambda = 2 (cos 2pi kx/N + cos 2pi ky/N + cos 2pi kz/N - 3) / res^2,
 * zero mode set to 0; lambda and scale/lambda evaluated in FP64. */
int lf2mi_poisson_green(void *kfield, int N, int C, int khalo, double res, double scale,
                        int dtype, void *stream);

/* lf2mi_poisson_green without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The eigenvalue table lives in a static device array of
 * 1024 doubles (8 KiB per GPU, held until the process ends), so N <= 1024 (LF2MI_ERANGE
 * otherwise); no hipMalloc/hipFree, no host synchronisation.  Calls on one GPU share that
 * array: calls that overlap on one GPU (different streams) race.  Results equal
 * lf2mi_poisson_green's to FP64 rounding (the table is computed on the device). */
int lf2mi_poisson_green_noalloc(void *kfield, int N, int C, int khalo, double res, double scale,
                                int dtype, void *stream);

/* Fused Poisson solve on the plan's geometry (C must be 1):
 *   phi = C2R( scale * G(k) * R2C(rho - shift) ),  G = -1/K^2, G(0) = 0,
 *   K^2 = sum_i N^2 (2 - 2 cos(2 pi k_i / N))
 * i.e. the driver sequence mean-subtract -> execute(FORWARD) -> Green -> execute(BACKWARD)
 * with the 1/N^3 normalisation folded into `scale` by the caller.  rho is read only;
 * the plan workspace holds the k-space data.  No k-field is materialised. */
int lf2mi_poisson_solve(lf2mi_fft_plan plan, const void *rho, void *phi, double shift,
                        double scale, void *stream);

/* ------------------------------------------------------------------ particles */
/* Particles<part_simple, part_simple_info, ...> (particles/LATfield2_Particles.hpp:131-344)
 * on one device.  LF2MI_F32: each particle is kept as an int32 cell index
 * x + N(y + N z) plus an FP32 offset in cell units in [0,1) per axis and FP32
 * velocities, bucketed by 8^3-cell bricks (the analogue of the per-cell lists).
 * LF2MI_F64: absolute FP64 positions and velocities in upload order (plain path). */
typedef struct lf2mi_parts_s *lf2mi_parts;

int lf2mi_parts_create(int N, int64_t capacity, int dtype, int flags, lf2mi_parts *out);
int lf2mi_parts_destroy(lf2mi_parts p);
/* part_info mass (FROM_INFO), default 1 */
int lf2mi_parts_set_mass(lf2mi_parts p, double mass_global);
int lf2mi_parts_count(lf2mi_parts p, int64_t *np);
int lf2mi_parts_bytes(lf2mi_parts p, int64_t *bytes);

/* addParticle_global for np particles (particles/LATfield2_Particles.hpp:501-519):
 * host arrays pos[3*np] (box units), vel[3*np], optional mass[np] and id[np].  Cells
 * are (int)floor(pos*N) % N per axis computed on the device in FP64 (:464-467);
 * particles with a negative coordinate are dropped as the reference drops them.
 * Replaces the container's contents.  *kept (may be NULL) = particles stored.  Syncs. */
int lf2mi_parts_upload(lf2mi_parts p, int64_t np, const double *pos, const double *vel,
                       const double *mass, const int64_t *id, int64_t *kept, void *stream);

/* copy the particles to the host in container order: pos as (cell + offset)/N for F32
 * (exactly the positions the container represents), vel, mass, id, cell (any may be
 * NULL).  Synchronises. */
int lf2mi_parts_download(lf2mi_parts p, double *pos, double *vel, double *mass, int64_t *id,
                         int32_t *cell, void *stream);

/* The particles whose cell lies in the box [lo[d], hi[d]) on every axis (0 <= lo < hi <= N,
 * no wrap), in no particular order: positions (box units) and velocities np*3, masses np,
 * host arrays of room max_out.  *n_out receives the count; LF2MI_ERANGE when it exceeds
 * max_out.  No counterpart in the reference; it serves subvolume checks of meshes whose
 * particles do not fit the host at once (Particles::getParticles over a region). */
int lf2mi_parts_download_box(lf2mi_parts p, const int *lo, const int *hi, int64_t max_out,
                             double *pos, double *vel, double *mass, int64_t *n_out,
                             void *stream);

/* ---- user particle structs (Particles<part, part_info, part_dataType> with any part
 * type, particles/LATfield2_Particles.hpp:131-344, :358-428).  The reference keeps the
 * whole struct per particle and reads fields by byte offset (the mass of
 * projections.hpp:108-126, :144-152).  A container given a payload of `bytes` per
 * particle (a multiple of 8, at most 4096; before the first upload) keeps each particle's
 * struct on the device and moves it with the particle through every re-bucketing.
 * Particles made by the synthetic initial conditions get a zero payload. */
int lf2mi_parts_set_payload(lf2mi_parts p, int bytes);

/* addParticle_global from n structs of `stride` bytes at host address `base`: pos[3] and
 * vel[3] of type Real (real_dtype LF2MI_F32 or LF2MI_F64) at byte offsets off_pos and
 * off_vel, the ID (int64) at off_id and a per-particle mass (double) at off_mass, each
 * -1 when absent (IDs then count from 0; masses need a container with masses or a
 * payload).  With a payload, stride must equal it and the structs are kept whole; a
 * container created with LF2MI_PARTS_MASS then accepts off_mass -1 and starts at mass 0
 * until lf2mi_parts_mass_from gathers the masses (a float member, LF2MI_MASS_PART_F32).
 * Same filtering and replacement as lf2mi_parts_upload.  Syncs. */
int lf2mi_parts_upload_aos(lf2mi_parts p, int64_t n, const void *base, int64_t stride, int real_dtype,
                           int64_t off_pos, int64_t off_vel, int64_t off_id, int64_t off_mass,
                           int64_t *kept, void *stream);

/* The particles back into n_room structs of `stride` bytes (container order): the payload
 * first (when kept), then pos and vel (Real of real_dtype, positions as in
 * lf2mi_parts_download) and the ID at their offsets (-1: not written).  *n_out = count
 * (LF2MI_ERANGE when it exceeds n_room).  Syncs. */
int lf2mi_parts_download_aos(lf2mi_parts p, void *base, int64_t n_room, int64_t stride, int real_dtype,
                             int64_t off_pos, int64_t off_vel, int64_t off_id, int64_t *n_out,
                             void *stream);

/* The mass the deposits use (the projections' flag_where / oset, projections.hpp:110-126):
 *   LF2MI_MASS_INFO  the global mass of lf2mi_parts_set_mass (FROM_INFO);
 *   LF2MI_MASS_PART  per particle: the double at byte `offset` of its payload
 *                    (FROM_PART), gathered on the device into the per-particle masses;
 *                    offset -1 uses those masses as they stand (uploaded, or gathered
 *                    by an earlier call);
 *   LF2MI_MASS_PART_F32  the same from a float at `offset`: a mass member declared Real
 *                    under SINGLE, read with its own type (the reference reads every
 *                    per-particle mass as a double, projections.hpp:150).
 * The choice holds for later deposits and PM steps (whose mean density follows it).
 * A container created with LF2MI_PARTS_MASS starts on LF2MI_MASS_PART, others on
 * LF2MI_MASS_INFO.  Syncs. */
#define LF2MI_MASS_INFO 0
#define LF2MI_MASS_PART 1
#define LF2MI_MASS_PART_F32 2 /* a float mass member (Real mass under SINGLE) */
int lf2mi_parts_mass_from(lf2mi_parts p, int where, int64_t offset, void *stream);

/* The vector a projection's oset[1] names (the vector and tensor projections read v at that
 * byte offset of each part: projections.hpp:453-459, :741-747, :1019-1025, :1394-1400):
 * offset -1, or the vel offset given to lf2mi_parts_upload_aos, selects the
 * container's velocity (the default); any other offset selects the three Reals of type
 * real_dtype at that offset of the payload, gathered on the deposit's stream by every later
 * deposit of a vector or tensor kind.  The payload moves with its particles, so the choice
 * holds across moves. */
int lf2mi_parts_vector_from(lf2mi_parts p, int64_t offset, int real_dtype);

/* per-cell particle counts into a device int32 array of N^3 entries (bit-exact with
 * the histogram of Particles::getPartCoord keys) */
int lf2mi_parts_cell_counts(lf2mi_parts p, int32_t *counts, void *stream);

/* Particles::getPartCoord without a container (particles/LATfield2_Particles.hpp:464-467):
 * for np FP64 positions (device, x,y,z per particle, box length 1) the key
 * x + N (y + N z) with c = (int)floor(pos N) % N per axis, or -1 where a coordinate is
 * negative (the particles addParticle_global does not keep); counts (N^3 int32, zeroed
 * here) receives the histogram of the valid keys.  keys or counts may be NULL. */
int lf2mi_bin_positions(const double *pos, int64_t np, int N, int32_t *keys, int32_t *counts,
                        void *stream);

/* synthetic initial conditions generated on the device (counter-based RNG keyed by
 * seed and particle index; positions produced in FP64 and binned as in upload):
 *   lattice: ng^3 particles on i/ng plus a Gaussian displacement of sigma_cells/N per
 *            axis, Gaussian velocities of sigma_v;
 *   uniform: np uniform positions, Gaussian velocities of sigma_v;
 *   blobs  : (1-frac_uniform)*np particles in nblobs isotropic Gaussian blobs of
 *            sigma_cells/N around uniform centres with Pareto(alpha) occupancies
 *            (largest-remainder rounding), the rest uniform; velocities sigma_v.
 * IDs (if kept) are the generation index.  Synchronises. */
int lf2mi_parts_ic_lattice(lf2mi_parts p, int ng, double sigma_cells, double sigma_v,
                           uint64_t seed, void *stream);
int lf2mi_parts_ic_uniform(lf2mi_parts p, int64_t np, double sigma_v, uint64_t seed,
                           void *stream);
int lf2mi_parts_ic_blobs(lf2mi_parts p, int64_t np, int nblobs, double alpha,
                         double sigma_cells, double frac_uniform, double sigma_v,
                         uint64_t seed, void *stream);

/* CIC deposit of the container into `field` (device, the container's precision):
 *   kind  LF2MI_PROJ_SCALAR -> scalarProjectionCIC_project  (projections.hpp:76-184)
 *         LF2MI_PROJ_VECTOR -> vectorProjectionCIC_project  (projections.hpp:980-1189)
 *         LF2MI_PROJ_VECVEC -> VecVecProjectionCIC_project  (projections.hpp:1356-1622)
 *         LF2MI_PROJ_VECTOR_NGP -> vectorProjectionCICNGP_project (projections.hpp:422-549):
 *             component i nearest-grid-point along axis i, CIC along the other two
 *         LF2MI_PROJ_SYMTENSOR_NGP -> symtensorProjectionCICNGP_project (projections.hpp:709-847):
 *             T_ii CIC on the vertices, T_ij (i<j) NGP along i and j, CIC along the third
 *   target LF2MI_DEP_UNFOLDED adds exactly what the reference _project adds (ghost
 *          nodes included, follow with lf2mi_projection_comm); LF2MI_DEP_PERIODIC adds
 *          the folded result into the interior (== _project + _comm on the interior).
 *   acc   LF2MI_ACC_NATIVE, or LF2MI_ACC_FIXED64 (F32 containers): every particle's term is
 *         rounded to an int64 fixed point and summed in integers, so the result does not
 *         depend on the order of the additions (bitwise reproducible); scalar kind: 2^-32 of
 *         the mass per unit weight; vector and tensor kinds: per component 2^(e-62) with
 *         np max|q| max|v_a| max|v_b| < 2^e (finding the maxima synchronises the stream).
 * Values are added (+=): zero the field first, as projection_init does, or ask for
 * LF2MI_DEP_PERIODIC_SET, which replaces the field's contents. */
int lf2mi_deposit(lf2mi_parts p, int kind, void *field, int halo, int target, int acc,
                  void *stream);

/* Fused CIC gather + kick + drift (Particles::updateVel then moveParticles with the
 * gather functor, particles/LATfield2_Particles.hpp:753-893, :1811-2803):
 *   F = -grad phi by central differences (phi(n+e) - phi(n-e)) N/2 with periodic wrap
 *   (reads the interior of phi only); vel += dt_kick * sum_corners w F(corner) with the
 *   scalar-deposit weights; pos += dt_drift * vel, wrapped (:2028-2032) and re-keyed.
 * In F32 the new position is formed in FP64 from (cell, offset, vel) and re-anchored to
 * (cell, FP32 offset).  The container is re-bucketed afterwards.  vmax2 (host, may be
 * NULL) receives max |v|^2 after the kick (synchronises when non-NULL).  phi may be
 * NULL when dt_kick == 0: a drift only (move_particles_simple, move_function.hpp:30). */
int lf2mi_gather_kick_drift(lf2mi_parts p, const void *phi, int halo, double dt_kick,
                            double dt_drift, double *vmax2, void *stream);

/* The output[] / reduce_type[] statistics of Particles::updateVel and moveParticles
 * (particles/LATfield2_Particles.hpp:788-807, :846-885; reduction bits :81-86).  For each
 * i < n, out[i] receives quantity[i] of every particle reduced by reduce_type[i]:
 *   reduce_type & (SUM | SUM_LOCAL): the sum (starting from 0);
 *   else & (MIN | MIN_LOCAL): the minimum, starting from 9223372036854775807;
 *   else & (MAX | MAX_LOCAL): the maximum, starting from -9223372036854775807;
 *   else out[i] is left as it is.  On one device the _LOCAL forms equal the global ones.
 * Quantities, evaluated in FP64 from the container's state; m is the mass the deposits use
 * (the global mass or the per-particle one, lf2mi_parts_mass_from): */
#define LF2MI_PQ_V2 0   /* |v|^2 */
#define LF2MI_PQ_MVX 1  /* m v_x */
#define LF2MI_PQ_MVY 2  /* m v_y */
#define LF2MI_PQ_MVZ 3  /* m v_z */
#define LF2MI_PQ_MASS 4 /* m */
#define LF2MI_PQ_EKIN 5 /* m |v|^2 / 2 */
#define LF2MI_RED_SUM 1
#define LF2MI_RED_MIN 2
#define LF2MI_RED_MAX 4
#define LF2MI_RED_SUM_LOCAL 8
#define LF2MI_RED_MIN_LOCAL 16
#define LF2MI_RED_MAX_LOCAL 32
/* n <= 8; quantity, reduce_type and out are host arrays; the sums are formed in a fixed
 * order, so they are the same from run to run.  Synchronises. */
int lf2mi_parts_reduce(lf2mi_parts p, int n, const int *quantity, const int *reduce_type, double *out,
                       void *stream);

/* One PM step (the driver sequence of examples/poissonSolver.cpp plus the particle
 * update): zero rho, periodic CIC deposit, mean-subtracted Poisson solve with
 * G = -1/K^2 and 1/N^3, gather + kick + drift, re-bucket.  rho and phi are real fields
 * with ghost width `halo` in the container's precision; plan must match (N, C=1).
 * After a first call (which sizes the buffers) a step only launches kernels on `stream`, so
 * steps can be captured into a HIP graph; the container alternates between two particle
 * buffers, so a capture that is replayed holds an even number of steps. */
int lf2mi_pm_step(lf2mi_parts p, lf2mi_fft_plan plan, void *rho, void *phi, int halo,
                  double dt_kick, double dt_drift, int acc, void *stream);

/* ------------------------------------------------------------ differentiable PM */
/* No counterpart in the reference: the adjoint of nsteps of its PM cycle
 * (scalarProjectionCIC_project particles/projections.hpp:76-300, the Poisson solve of
 * examples/poissonSolver.cpp:99-111, the gather of updateVel Particles.hpp:825-828 and
 * the drift of moveParticles :2028-2032), for callers that fit initial conditions.
 * Forward, box length 1, periodic, particle order kept:
 *   rho = CIC deposit of q = mass N^3 per particle; phi = G rho, G = -1/K^2, G(0) = 0;
 *   F = -grad phi (central differences); v += dt * gather(F); x += dt * v (wrapped);
 *   loss = sum_cells (rho(x_final) - target)^2 over the N^3 interior.
 * The starting state of every step is checkpointed on the device; the backward pass
 * recomputes each step's fields and runs the hand-written adjoint: the gather's
 * transpose is a CIC deposit, the central difference is antisymmetric, G is symmetric
 * (the same FFT solve), and the CIC weights' position derivatives give the direct
 * terms.  dtype (F32/F64) is the precision of the whole path. */
typedef struct lf2mi_diffpm_s *lf2mi_diffpm;

int lf2mi_diffpm_create(int N, int64_t np, int nsteps, double mass, double dt, int dtype,
                        lf2mi_diffpm *out);
int lf2mi_diffpm_destroy(lf2mi_diffpm m);
/* device bytes held by the handle (checkpoints, fields, adjoints, FFT workspace) */
int lf2mi_diffpm_bytes(lf2mi_diffpm m, int64_t *bytes);
/* pos0, vel0: np*3 (x,y,z per particle) in box units, device, io_dtype F32 or F64;
 * vel0 may be NULL (zero).  target: N^3 device values, x fastest, io_dtype, may be
 * NULL (zero).  rho_out (may be NULL) receives the final density, N^3, x fastest.
 * loss (host, may be NULL) receives L and synchronises the stream. */
int lf2mi_diffpm_forward(lf2mi_diffpm m, const void *pos0, const void *vel0,
                         const void *target, int io_dtype, void *rho_out, double *loss,
                         void *stream);
/* dL/dpos0 and dL/dvel0 (np*3 each, io_dtype of the last forward; grad_vel may be NULL) */
int lf2mi_diffpm_backward(lf2mi_diffpm m, void *grad_pos, void *grad_vel, void *stream);
/* positions (box units) and velocities at the start of step `step` (0..nsteps; nsteps
 * is the final state) of the last forward; either output may be NULL */
int lf2mi_diffpm_state(lf2mi_diffpm m, int step, void *pos, void *vel, int io_dtype,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif
