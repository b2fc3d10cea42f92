/*
 * lf2_oracle.h -- FP64 CPU restatement of the LATfield2 particle-mesh (PM) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels in
 * latfield2_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  No product code path calls into it.
 *
 * Conventions (all restated from the reference, citations relative to the
 * LATfield2 source tree):
 *   - cubic lattice of N sites per axis, box length 1, resolution res = 1/N;
 *   - real fields are AoS with C components and a ghost layer of width h:
 *       data[idx*C + c], idx = (x+h) + G*((y+h) + G*(z+h)), G = N + 2h
 *     (LATfield2_Lattice.hpp:111-113 strides, LATfield2_Field_decl.hpp:760-807);
 *   - the R2C k-field has lattice dims (kx, kz, ky) = (N/2+1, N, N) and its own
 *     ghost width kh:  k[((kx+kh) + R*((kz+kh) + (N+2kh)*(ky+kh)))*C + c],
 *     R = N/2+1+2kh  (LATfield2_PlanFFT_decl.hpp:1469-1510);
 *   - particle positions are absolute, in box units, AoS pos[3*i+d];
 *     the cell key is (int)floor(pos*N) % N
 *     (particles/LATfield2_Particles.hpp:464-467).
 * Every function returns 0 on success and a negative code on bad arguments.
 */
#ifndef LF2_ORACLE_H
#define LF2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef long long orc_i64;

/* threads used by the OpenMP loops (0 = OpenMP default) */
int orc_set_threads(int n);
int orc_get_threads(void);

/* projection_init: zero every value including ghosts (particles/projections.hpp:31-34) */
int orc_zero(double *f, orc_i64 count);

/* Field::updateHalo periodic ghost fill (LATfield2_Field_decl.hpp:889-935) */
int orc_ghost_fill(double *f, int N, int h, int C);

/* cell keys: key[i] = x + N(y + N z) with x = (int)floor(pos*N) % N per axis;
 * -1 when any axis gives a negative coordinate (the reference's setCoord fails
 * and addParticle_global drops the particle, Particles.hpp:501-519). */
int orc_bin(const double *pos, orc_i64 np, int N, int *key);

/* per-cell particle counts (histogram of orc_bin keys), counts has N^3 entries */
int orc_cell_counts(const double *pos, orc_i64 np, int N, int *counts);

/* CIC deposits.  mass == NULL: every particle has mass mass_global.  The field
 * is accumulated into (+=), exactly like the reference: call orc_zero first.
 *   scalar : scalarProjectionCIC_project   (particles/projections.hpp:76-184)
 *   vector : vectorProjectionCIC_project   (particles/projections.hpp:980-1189)
 *   vecvec : VecVecProjectionCIC_project   (particles/projections.hpp:1356-1622)
 * vector writes C=3 comps, vecvec writes C=6 comps in the symmetric order
 * (0,0)(0,1)(0,2)(1,1)(1,2)(2,2) (LATfield2_Field_decl.hpp:773-788). */
int orc_scalar_project(const double *pos, const double *mass, double mass_global,
                       orc_i64 np, double *rho, int N, int h);
int orc_vector_project(const double *pos, const double *vel, const double *mass,
                       double mass_global, orc_i64 np, double *B, int N, int h);
int orc_vecvec_project(const double *pos, const double *vel, const double *mass,
                       double mass_global, orc_i64 np, double *T, int N, int h);
/* the CIC-NGP hybrids:
 *   vector_ngp    : vectorProjectionCICNGP_project    (particles/projections.hpp:422-549)
 *   symtensor_ngp : symtensorProjectionCICNGP_project (particles/projections.hpp:709-847) */
int orc_vector_ngp_project(const double *pos, const double *vel, const double *mass,
                           double mass_global, orc_i64 np, double *B, int N, int h);
int orc_symtensor_ngp_project(const double *pos, const double *vel, const double *mass,
                              double mass_global, orc_i64 np, double *T, int N, int h);

/* the periodic ghost folds that follow each deposit (the *_comm functions):
 *   scalar : particles/projections.hpp:194-300
 *   vector : particles/projections.hpp:1192-1351
 *   vecvec : particles/projections.hpp:1624-1814 */
int orc_scalar_comm(double *rho, int N, int h);
int orc_vector_comm(double *B, int N, int h);
int orc_vecvec_comm(double *T, int N, int h);
/*   vector_ngp    : particles/projections.hpp:558-681
 *   symtensor_ngp : particles/projections.hpp:856-976 */
int orc_vector_ngp_comm(double *B, int N, int h);
int orc_symtensor_ngp_comm(double *T, int N, int h);

/* Brute-force periodic CIC deposit straight from the hat-function definition:
 * value(node n) = sum_p q_p * prod_d hat(p_d*N - n_d - s_d), over all periodic
 * images, for a single component with stagger s (0 or 0.5 per axis).
 * q = (mass/res^3) * vfac where vfac = 1, v_a, or v_a*v_b (a,b = -1 for none).
 * Axes in the bit mask `ngp` use the nearest-grid-point kernel instead of the hat:
 * all of q on the node of the staggered grid nearest to the particle.
 * Writes the N^3 interior of a C=1 field without ghosts (out[x + N(y + N z)]).
 * Independent of the per-cell loop structure used by the *_project functions. */
int orc_brute_deposit(const double *pos, const double *vel, const double *mass,
                      double mass_global, orc_i64 np, int N, const double *stagger,
                      int va, int vb, int ngp, double *out);

/* 3-D R2C / C2R FFT with the PlanFFT layout (LATfield2_PlanFFT_decl.hpp:851-1056).
 * forward  : k = sum_x f(x) exp(-2 pi i k.x/N), unnormalised, input preserved;
 * backward : f = sum_k F(k) exp(+2 pi i k.x/N), unnormalised (caller divides by
 *            N^3), k-field preserved, real ghosts untouched; like FFTW's c2r only
 *            the real parts of the kx = 0 and kx = N/2 modes are used.
 * k is complex interleaved (re, im). */
int orc_fft_r2c(const double *r, double *k, int N, int C, int rh, int kh);
int orc_fft_c2r(const double *k, double *r, int N, int C, int rh, int kh);

/* C2C PlanFFT of two Field<Imag> (LATfield2_PlanFFT_decl.hpp:608-719, :1059-1198).
 * Complex fields: interleaved (re, im) doubles, C components per site.
 * forward : unnormalised e^{-} DFT of the complex interior; mode (kx, ky, kz) at k-lattice
 *           storage (kz, kx, ky) (cKSite, LATfield2_Site.hpp:219-220); input kept;
 * backward: the unnormalised e^{+} inverse, backward(forward(f)) = N^3 f;
 * backward_reference: what the reference's backward computes on one rank (its z pass
 *           runs twice and its y pass not at all), with one ghost width h for both. */
int orc_fft_c2c_forward(const double *f, double *k, int N, int C, int rh, int kh);
int orc_fft_c2c_backward(const double *k, double *f, int N, int C, int rh, int kh);
int orc_fft_c2c_backward_reference(const double *k, double *f, int N, int C, int h);

/* Green's function multiply of examples/poissonSolver.cpp:99-111:
 *   phi_k = rho_k * scale / lambda,  lambda = 2 (sum_i cos(2 pi k_i/N) - 3) / res^2,
 *   zero mode set to 0.  In place on a k-field in the layout above. */
int orc_poisson_green(double *k, int N, int C, int kh, double res, double scale);

/* Force field F_d = -(phi(n+e_d) - phi(n-e_d)) * N/2 on the interior (central
 * difference, res = 1/N), phi must have ghosts filled; F (C=3, same halo h) gets
 * its ghosts filled too. */
int orc_force_field(const double *phi, double *F, int N, int h);

/* CIC gather + kick: acc_d(p) = sum_corners w * F_d(corner), with the corner
 * weights of the scalar deposit and frac = pos*N - cell (Particles.hpp:825-828);
 * vel += dt_kick * acc.  acc (np*3) may be NULL.  Returns max |v|^2 via vmax2. */
int orc_gather_kick(const double *pos, double *vel, orc_i64 np, const double *F,
                    int N, int h, double dt_kick, double *acc, double *vmax2);

/* drift pos += dt*vel, then the sequential periodic wrap of Particles.hpp:2028-2032 */
int orc_drift(double *pos, const double *vel, orc_i64 np, double dt);

/* One PM step with box length 1 (the driver sequence of examples/poissonSolver.cpp
 * and benchmarks_tests/benchmarks_particles.cpp):
 *   rho = CIC(particles) folded; rho -= mean; rho_k = FFT(rho);
 *   phi_k = -rho_k / K^2 (K^2 = sum_i N^2 (2 - 2 cos(2 pi k_i/N)), zero mode 0);
 *   phi = C2R(phi_k) / N^3; ghosts; F = -grad phi; vel += dt_kick * F(p);
 *   pos += dt_drift * vel (wrapped).
 * rho, phi: (N+2)^3 fields (h = 1); acc: np*3 or NULL.
 * stage_ms (8 doubles or NULL) receives wall-clock per stage:
 *   deposit, fold, mean, fft_fwd, green, fft_bwd, halo+force, gather+kick+drift. */
int orc_pm_step(double *pos, double *vel, const double *mass, double mass_global,
                orc_i64 np, int N, double dt_kick, double dt_drift,
                double *rho, double *phi, double *acc, double *stage_ms);

/* ---------------------------------------------------------------------------
 * Differentiable PM (no counterpart in the reference; the reference gives the
 * forward pieces).  Periodic CIC deposit of mass m (density m N^3 per particle),
 * Poisson phi = -K^-2 (rho - mean) with the zero mode removed, F = -grad phi by
 * central differences, CIC gather.  One leapfrog step in kick-drift form:
 *   v += dt a(x);  x += dt v;  x wrapped into [0,1).
 * Loss after nsteps: L = sum_cells (rho(x_final) - target)^2 on the N^3 interior.
 * ------------------------------------------------------------------------- */
/* periodic CIC deposit of q_p = m N^3 onto out[N^3] (x fastest), cleared first */
int orc_cic_density(const double *pos, orc_i64 np, int N, double mass, double *out);

/* nsteps of the leapfrog in place; ckpt (may be NULL) receives pos and vel at the start
 * of every step: ckpt[s*6*np + ...] = pos (np*3) then vel (np*3) */
int orc_diffpm_forward(double *pos, double *vel, orc_i64 np, int N, double mass, double dt,
                       int nsteps, double *ckpt);

/* L and dL/dpos0, dL/dvel0 (box units) by the adjoint of the forward above:
 * backward through the checkpointed steps with the transposed gather (deposit),
 * the transposed central difference, the symmetric Green's operator and the position
 * derivatives of the CIC weights.  grad_vel may be NULL. */
int orc_diffpm_loss_grad(const double *pos0, const double *vel0, orc_i64 np, int N, double mass,
                         double dt, int nsteps, const double *target, double *loss,
                         double *grad_pos, double *grad_vel);

/* the loss alone (forward only), for finite differences */
int orc_diffpm_loss(const double *pos0, const double *vel0, orc_i64 np, int N, double mass,
                    double dt, int nsteps, const double *target, double *loss);

#ifdef __cplusplus
}
#endif
#endif
