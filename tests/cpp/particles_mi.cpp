// Particles + projections + updateVel/moveParticles through the C++ mirror.  Particles on
// every node with unit velocity (the initial condition of benchmarks_tests/
// benchmarks_particles.cpp:131-149, with box length 1 and mass 0.1):
//   - scalar, vector and VecVec CIC projections followed by their _comm are uniform
//     (rho = m N^3, B_i = m v_i N^3, T_ij = m v_i v_j N^3) in the interior;
//   - the gathered force of a uniform potential is zero, so updateVel leaves v = 1 and
//     returns |v| = sqrt(3);
//   - moveParticles(gpu_drift, dt) moves every particle by dt in each direction.
// Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

int main(int argc, char **argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 32;
    const double tol = sizeof(Real) == 8 ? 1e-12 : 1e-5;
    int fails = 0;
    Lattice lat_part(3, N, 0), lat(3, N, 1);
    Field<Real> rho(lat, 1), B(lat, 3), T(lat, 3, 3, symmetric), phi(lat, 1);
    part_simple_info info;
    info.mass = 0.1;
    part_simple_dataType dtype;
    Real box[3] = {1, 1, 1};
    Particles<part_simple, part_simple_info, part_simple_dataType> parts;
    parts.initialize(info, dtype, &lat_part, box);
    Site xp(lat_part);
    long id = 0;
    for (xp.first(); xp.test(); xp.next()) {
        part_simple p;
        p.ID = id++;
        for (int d = 0; d < 3; ++d) {
            p.pos[d] = (Real)xp.coord(d) / N;
            p.vel[d] = 1.0;
        }
        parts.addParticle_global(p);
    }
    if (parts.numParticles() != (long)N * N * N) ++fails;
    projection_init(&rho);
    projection_init(&B);
    projection_init(&T);
    scalarProjectionCIC_project(&parts, &rho);
    scalarProjectionCIC_comm(&rho);
    vectorProjectionCIC_project(&parts, &B);
    vectorProjectionCIC_comm(&B);
    VecVecProjectionCIC_project(&parts, &T);
    VecVecProjectionCIC_comm(&T);
    rho.toHost();
    B.toHost();
    T.toHost();
    const double want = 0.1 * N * N * N;
    double e_rho = 0, e_b = 0, e_t = 0;
    Site x(lat);
    for (x.first(); x.test(); x.next()) {
        e_rho = std::fmax(e_rho, std::fabs(rho(x) / want - 1));
        for (int c = 0; c < 3; ++c) e_b = std::fmax(e_b, std::fabs(B(x, c) / want - 1));
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) e_t = std::fmax(e_t, std::fabs(T(x, i, j) / want - 1));
    }
    std::printf("projections: max rel error rho %.2e B %.2e T %.2e\n", e_rho, e_b, e_t);
    if (e_rho > tol || e_b > tol || e_t > tol) ++fails;
    // uniform potential: no force
    for (x.first(); x.test(); x.next()) phi(x) = 3.0;
    phi.toDevice();
    Field<Real> *fields[1] = {&phi};
    const Real vmax = parts.updateVel(gpu_gather_kick(), 0.01, fields, 1);
    std::printf("updateVel: max |v| = %.9f (expected sqrt 3)\n", (double)vmax);
    if (std::fabs(vmax - std::sqrt(3.0)) > 10 * tol) ++fails;
    const double dt = 0.25 / N;  // a quarter cell
    parts.moveParticles(gpu_drift(), dt);
    std::vector<part_simple> out;
    parts.getParticles(out);
    double e_pos = 0.0;
    for (const part_simple &p : out) {
        const long i = p.ID;
        const int c[3] = {(int)(i % N), (int)((i / N) % N), (int)(i / ((long)N * N))};
        for (int d = 0; d < 3; ++d) {
            double w = (double)c[d] / N + dt;
            if (w >= 1.0) w -= 1.0;
            e_pos = std::fmax(e_pos, std::fabs(p.pos[d] - w));
        }
    }
    std::printf("moveParticles: max position error %.3e\n", e_pos);
    if (e_pos > 1e-7) ++fails;
    std::printf("particles_mi N=%d %s: %d failed checks\n", N, sizeof(Real) == 8 ? "double" : "single", fails);
    return fails;
}
