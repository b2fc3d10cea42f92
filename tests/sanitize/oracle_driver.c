/* Runs every oracle entry point once on small inputs, for a build of oracle/lf2_oracle.c with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitizers.py).  The
 * sizes cover ghost widths 0-2, several components, non-power-of-two FFT lengths and
 * particles on cell faces.  Exit code: the number of calls that returned an error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lf2_oracle.h"

static double frand(unsigned *s) {
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0;
}

static double *field(int N, int h, int C) {
    const long G = N + 2 * h;
    return (double *)calloc((size_t)(G * G * G * C), sizeof(double));
}

int main(void) {
    int bad = 0;
    unsigned seed = 7;
    const orc_i64 np = 200;
    double *pos = (double *)malloc(sizeof(double) * 3 * np), *vel = (double *)malloc(sizeof(double) * 3 * np);
    double *mass = (double *)malloc(sizeof(double) * np);
    for (orc_i64 i = 0; i < np; ++i) {
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = frand(&seed);
            vel[3 * i + d] = frand(&seed) - 0.5;
        }
        mass[i] = 0.5 + frand(&seed);
    }
    pos[0] = 0.0;  /* on a node */
    pos[4] = 0.5;  /* on a cell face */
    orc_set_threads(2);
    for (int h = 1; h <= 2; ++h) {
        const int N = 9;
        double *rho = field(N, h, 1), *B = field(N, h, 3), *T = field(N, h, 6);
        int *key = (int *)malloc(sizeof(int) * np), *cnt = (int *)malloc(sizeof(int) * N * N * N);
        bad += orc_bin(pos, np, N, key) != 0;
        bad += orc_cell_counts(pos, np, N, cnt) != 0;
        bad += orc_scalar_project(pos, mass, 1.0, np, rho, N, h) != 0;
        bad += orc_vector_project(pos, vel, mass, 1.0, np, B, N, h) != 0;
        bad += orc_vecvec_project(pos, vel, NULL, 2.0, np, T, N, h) != 0;
        bad += orc_scalar_comm(rho, N, h) != 0;
        bad += orc_vector_comm(B, N, h) != 0;
        bad += orc_vecvec_comm(T, N, h) != 0;
        memset(B, 0, sizeof(double) * (size_t)((N + 2 * h) * (N + 2 * h) * (N + 2 * h) * 3));
        memset(T, 0, sizeof(double) * (size_t)((N + 2 * h) * (N + 2 * h) * (N + 2 * h) * 6));
        bad += orc_vector_ngp_project(pos, vel, mass, 1.0, np, B, N, h) != 0;
        bad += orc_symtensor_ngp_project(pos, vel, NULL, 1.0, np, T, N, h) != 0;
        bad += orc_vector_ngp_comm(B, N, h) != 0;
        bad += orc_symtensor_ngp_comm(T, N, h) != 0;
        bad += orc_ghost_fill(T, N, h, 6) != 0;
        double *out = (double *)malloc(sizeof(double) * N * N * N);
        const double st[3] = {0.5, 0.0, 0.5};
        bad += orc_brute_deposit(pos, vel, mass, 1.0, np, N, st, 0, 2, 5, out) != 0;
        free(out);
        free(rho); free(B); free(T); free(key); free(cnt);
    }
    /* FFTs: power of two and not, ghosts on both sides, several components */
    const int sizes[3][4] = {{8, 1, 0, 2}, {12, 2, 1, 1}, {10, 0, 2, 3}};
    for (int t = 0; t < 3; ++t) {
        const int N = sizes[t][0], rh = sizes[t][1], kh = sizes[t][2], C = sizes[t][3];
        double *r = field(N, rh, C), *r2 = field(N, rh, C);
        const long KG = N + 2 * kh, R = N / 2 + 1 + 2 * kh;
        double *k = (double *)calloc((size_t)(2 * KG * KG * R * C), sizeof(double));
        double *kc = (double *)calloc((size_t)(2 * KG * KG * KG * C), sizeof(double));
        double *c = (double *)calloc((size_t)(2 * (N + 2 * rh) * (N + 2 * rh) * (N + 2 * rh) * C), sizeof(double));
        double *c2 = (double *)calloc((size_t)(2 * (N + 2 * rh) * (N + 2 * rh) * (N + 2 * rh) * C), sizeof(double));
        const long nr = (long)(N + 2 * rh) * (N + 2 * rh) * (N + 2 * rh) * C;
        for (long i = 0; i < nr; ++i) {
            r[i] = frand(&seed);
            c[2 * i] = frand(&seed);
            c[2 * i + 1] = frand(&seed);
        }
        bad += orc_fft_r2c(r, k, N, C, rh, kh) != 0;
        bad += orc_poisson_green(k, N, C, kh, 1.0 / N, 1.0) != 0;
        bad += orc_fft_c2r(k, r2, N, C, rh, kh) != 0;
        bad += orc_fft_c2c_forward(c, kc, N, C, rh, kh) != 0;
        bad += orc_fft_c2c_backward(kc, c2, N, C, rh, kh) != 0;
        if (rh == kh) bad += orc_fft_c2c_backward_reference(kc, c2, N, C, rh) != 0;
        free(r); free(r2); free(k); free(kc); free(c); free(c2);
    }
    /* PM step, gather, drift and the differentiable PM */
    {
        const int N = 8;
        double *rho = field(N, 1, 1), *phi = field(N, 1, 1), *F = field(N, 1, 3);
        double *acc = (double *)malloc(sizeof(double) * 3 * np), stage[8], vmax2;
        bad += orc_pm_step(pos, vel, mass, 1.0, np, N, 0.01, 0.01, rho, phi, acc, stage) != 0;
        bad += orc_force_field(phi, F, N, 1) != 0;
        bad += orc_gather_kick(pos, vel, np, F, N, 1, 0.01, acc, &vmax2) != 0;
        bad += orc_drift(pos, vel, np, 0.01) != 0;
        double *dens = (double *)malloc(sizeof(double) * N * N * N), *target = (double *)calloc(N * N * N, sizeof(double));
        double *gp = (double *)malloc(sizeof(double) * 3 * np), *gv = (double *)malloc(sizeof(double) * 3 * np), loss;
        bad += orc_cic_density(pos, np, N, 1.0 / np, dens) != 0;
        bad += orc_diffpm_loss_grad(pos, vel, np, N, 1.0 / np, 0.02, 2, target, &loss, gp, gv) != 0;
        bad += orc_diffpm_loss(pos, vel, np, N, 1.0 / np, 0.02, 2, target, &loss) != 0;
        free(rho); free(phi); free(F); free(acc); free(dens); free(target); free(gp); free(gv);
    }
    free(pos); free(vel); free(mass);
    printf("oracle driver: %d failing calls\n", bad);
    return bad;
}
