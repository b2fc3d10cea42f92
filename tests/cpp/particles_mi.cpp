// Particles + projections + updateVel/moveParticles through the C++ mirror.  Particles on
// every node with unit velocity (the initial condition of benchmarks_tests/
// benchmarks_particles.cpp:131-149, with box length 1 and mass 0.1):
//   - scalar, vector and VecVec CIC projections followed by their _comm are uniform
//     (rho = m N^3, B_i = m v_i N^3, T_ij = m v_i v_j N^3) in the interior, and so are the
//     CIC-NGP vector and symmetric tensor projections (projections.hpp:405-976);
//   - the gathered force of a uniform potential is zero, so updateVel leaves v = 1 and
//     returns |v| = sqrt(3);
//   - moveParticles(gpu_drift, dt) moves every particle by dt in each direction;
//   - a part type with its own `mass` and another double member: the projection takes the
//     individual mass by default, any double member with FROM_PART and an offset, and
//     part_info's mass with FROM_INFO (projections.hpp:110-126); the whole structs come
//     back from getParticles after a move (projections.hpp:144-152, Particles.hpp:398-427).
// Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <vector>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

// a user part type: an individual mass and a weight beside the reference's members
struct part_weighted {
    long ID;
    Real pos[3];
    Real vel[3];
    double mass;
    double weight;
    Real mom[3];
};

// a part type whose mass is a Real (a float under SINGLE), read with its declared type
struct part_real_mass {
    long ID;
    Real pos[3];
    Real vel[3];
    Real mass;
};

// sum of the interior of a scalar field (host copy)
static double interior_sum(Field<Real> &f, Lattice &lat) {
    f.toHost();
    double s = 0.0;
    Site x(lat);
    for (x.first(); x.test(); x.next()) s += f(x);
    return s;
}

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
    // the CIC-NGP hybrids of the same lattice are uniform too
    projection_init(&B);
    projection_init(&T);
    vectorProjectionCICNGP_project(&parts, &B);
    vectorProjectionCICNGP_comm(&B);
    symtensorProjectionCICNGP_project(&parts, &T);
    symtensorProjectionCICNGP_comm(&T);
    B.toHost();
    T.toHost();
    e_b = 0;
    e_t = 0;
    for (x.first(); x.test(); x.next()) {
        for (int c = 0; c < 3; ++c) e_b = std::fmax(e_b, std::fabs(B(x, c) / want - 1));
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) e_t = std::fmax(e_t, std::fabs(T(x, i, j) / want - 1));
    }
    std::printf("CIC-NGP projections: max rel error B %.2e T %.2e\n", e_b, e_t);
    if (e_b > tol || e_t > tol) ++fails;
    // uniform potential: no force
    for (x.first(); x.test(); x.next()) phi(x) = 3.0;
    phi.toDevice();
    Field<Real> *fields[1] = {&phi};
    // the reference's updateVel signature with outputs (particles/LATfield2_Particles.hpp:217-224):
    // slots |v|^2, m v_x, m v_y, m v_z, m, m |v|^2 / 2 of the device functors
    double stats[6] = {-1, -1, -1, -1, -1, 42.0};
    int rtype[6] = {SUM, MAX_LOCAL, SUM, MIN, SUM_LOCAL, 0};
    const Real vmax = parts.updateVel(updateVel_cicKick, 0.01, fields, 1, NULL, stats, rtype, 6);
    const double np3 = (double)N * N * N;
    std::printf("updateVel: max |v| = %.9f (expected sqrt 3); sum |v|^2 %.9g, max m v_x %.9g, sum m v_y %.9g, "
                "min m v_z %.9g, sum m %.9g, untouched %.3g\n",
                (double)vmax, stats[0], stats[1], stats[2], stats[3], stats[4], stats[5]);
    if (std::fabs(vmax - std::sqrt(3.0)) > 10 * tol) ++fails;
    if (std::fabs(stats[0] / (3.0 * np3) - 1) > 10 * tol || std::fabs(stats[1] / 0.1 - 1) > 10 * tol ||
        std::fabs(stats[2] / (0.1 * np3) - 1) > 10 * tol || std::fabs(stats[3] / 0.1 - 1) > 10 * tol ||
        std::fabs(stats[4] / (0.1 * np3) - 1) > 10 * tol || stats[5] != 42.0)
        ++fails;
    // the identity functor: velocities unchanged, the same outputs
    double out2[2] = {0, 0};
    int rtype2[2] = {MIN, SUM};
    const Real vmax2 = parts.updateVel(updateVel_simple, 0.01, NULL, 0, NULL, out2, rtype2, 2);
    std::printf("updateVel_simple: max |v| = %.9f, min |v|^2 %.9g, sum m v_x %.9g\n", (double)vmax2, out2[0], out2[1]);
    if (std::fabs(vmax2 - std::sqrt(3.0)) > 10 * tol || std::fabs(out2[0] - 3.0) > 10 * tol ||
        std::fabs(out2[1] / (0.1 * np3) - 1) > 10 * tol)
        ++fails;
    const double dt = 0.25 / N;  // a quarter cell
    double mout[1] = {0};
    int mtype[1] = {SUM};
    parts.moveParticles(move_particles_simple, dt, NULL, 0, NULL, mout, mtype, 1);
    if (std::fabs(mout[0] / (3.0 * np3) - 1) > 10 * tol) ++fails;
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
    // user structs: masses 0.2 + 0.01 (ID mod 5), weights 0.3 + 0.02 (ID mod 3)
    {
        Particles<part_weighted, part_simple_info, part_simple_dataType> wp;
        wp.initialize(info, dtype, &lat_part, box);
        double msum = 0.0, wsum = 0.0, momsum[3] = {0.0, 0.0, 0.0};
        long k = 0;
        for (xp.first(); xp.test(); xp.next()) {
            part_weighted p;
            p.ID = k;
            for (int d = 0; d < 3; ++d) {
                p.pos[d] = ((Real)xp.coord(d) + (Real)0.37) / N;
                p.vel[d] = 0.5;
            }
            p.mass = 0.2 + 0.01 * (k % 5);
            p.weight = 0.3 + 0.02 * (k % 3);
            for (int d = 0; d < 3; ++d) {
                p.mom[d] = (Real)(1.0 + 0.125 * (k % 7) + d);
                momsum[d] += p.mom[d];
            }
            msum += p.mass;
            wsum += p.weight;
            wp.addParticle_global(p);
            ++k;
        }
        const double nv = (double)N * N * N;
        size_t ow = offsetof(part_weighted, weight), om = offsetof(part_simple_info, mass);
        projection_init(&rho);
        scalarProjectionCIC_project(&wp, &rho);  // the individual mass
        scalarProjectionCIC_comm(&rho);
        const double s_mass = interior_sum(rho, lat) / nv;
        projection_init(&rho);
        scalarProjectionCIC_project(&wp, &rho, &ow, FROM_PART);
        scalarProjectionCIC_comm(&rho);
        const double s_weight = interior_sum(rho, lat) / nv;
        projection_init(&rho);
        scalarProjectionCIC_project(&wp, &rho, &om, FROM_INFO);
        scalarProjectionCIC_comm(&rho);
        const double s_info = interior_sum(rho, lat) / nv;
        const double stol = sizeof(Real) == 8 ? 1e-10 : 2e-4;
        std::printf("struct masses: individual %.9g (%.9g), FROM_PART weight %.9g (%.9g), FROM_INFO %.9g (%.9g)\n",
                    s_mass, msum, s_weight, wsum, s_info, 0.1 * k);
        if (std::fabs(s_mass / msum - 1) > stol || std::fabs(s_weight / wsum - 1) > stol ||
            std::fabs(s_info / (0.1 * k) - 1) > stol)
            ++fails;
        // vector projections: the velocity with the individual mass by default; any Real[3]
        // member with oset[1] (projections.hpp:453-459, :1019-1025)
        auto component_sum = [&](int c) {
            B.toHost();
            Site y(lat);
            double t = 0.0;
            for (y.first(); y.test(); y.next()) t += B(y, c);
            return t / nv;
        };
        projection_init(&B);
        vectorProjectionCIC_project(&wp, &B);
        vectorProjectionCIC_comm(&B);
        const double b_vel = component_sum(1);
        size_t os[2] = {om, offsetof(part_weighted, mom)};
        projection_init(&B);
        vectorProjectionCICNGP_project(&wp, &B, os, FROM_INFO);
        vectorProjectionCICNGP_comm(&B);
        double b_err = std::fabs(b_vel / (0.5 * msum) - 1);
        for (int c = 0; c < 3; ++c) b_err = std::fmax(b_err, std::fabs(component_sum(c) / (0.1 * momsum[c]) - 1));
        std::printf("struct vectors: default (vel, individual mass) and oset = {info mass, mom}: max rel error %.2e\n",
                    b_err);
        if (b_err > stol) ++fails;
        wp.moveParticles(move_particles_simple, 3.0 / N);  // 1.5 cells: many cross a brick face
        std::vector<part_weighted> back;
        wp.getParticles(back);
        long bad = (long)back.size() != k ? 1 : 0;
        for (const part_weighted &p : back) {
            if (p.mass != 0.2 + 0.01 * (p.ID % 5) || p.weight != 0.3 + 0.02 * (p.ID % 3)) ++bad;
            const long c[3] = {p.ID % N, (p.ID / N) % N, p.ID / ((long)N * N)};
            for (int d = 0; d < 3; ++d) {
                double w = ((double)c[d] + 0.37) / N + 1.5 / N;
                if (w >= 1.0) w -= 1.0;
                if (std::fabs(p.pos[d] - w) > (sizeof(Real) == 8 ? 1e-12 : 1e-6)) ++bad;
            }
        }
        std::printf("struct round trip after a move: %ld mismatches\n", bad);
        if (bad) ++fails;
    }
    // a Real mass member (projections.hpp:150 reads it as a double; here with its own type)
    {
        Particles<part_real_mass, part_simple_info, part_simple_dataType> rp;
        rp.initialize(info, dtype, &lat_part, box);
        double msum = 0.0;
        long k = 0;
        for (xp.first(); xp.test(); xp.next()) {
            part_real_mass p;
            p.ID = k;
            for (int d = 0; d < 3; ++d) {
                p.pos[d] = ((Real)xp.coord(d) + (Real)0.25) / N;
                p.vel[d] = 0.0;
            }
            p.mass = (Real)(0.5 + 0.125 * (k % 3));
            msum += p.mass;
            rp.addParticle_global(p);
            ++k;
        }
        projection_init(&rho);
        scalarProjectionCIC_project(&rp, &rho);
        scalarProjectionCIC_comm(&rho);
        const double s_real = interior_sum(rho, lat) / ((double)N * N * N);
        const double stol = sizeof(Real) == 8 ? 1e-10 : 2e-4;
        std::printf("Real mass member (%zu bytes): deposit %.9g (%.9g)\n", sizeof(Real), s_real, msum);
        if (std::fabs(s_real / msum - 1) > stol) ++fails;
    }
    std::printf("particles_mi N=%d %s: %d failed checks\n", N, sizeof(Real) == 8 ? "double" : "single", fails);
    return fails;
}
