// Site displacements and ghost indexing through the C++ mirror, restating
// examples/boundary_test.cpp:41-123 (128^3, halo 3, a 3-component field holding each site's
// coordinates).  Two checks:
//   - after updateHalo (on the device), x -/+ 1..3 steps along direction k from every interior
//     site reads the periodic image: (coord_k -/+ s) mod N;
//   - after every ghost site is overwritten with its own (unwrapped) coordinates, as the
//     reference's forallboundary loop does, the same steps read coord_k -/+ s exactly.
// Exit code = number of failed reads (capped at 255).
#include <cstdio>
#include <cstdlib>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

int main(int argc, char **argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 128;
    const int halo = 3;
    Lattice lat(3, N, halo);
    Field<Real> rho(lat, 3);
    Site x(lat);
    for (x.first(); x.test(); x.next())
        for (int j = 0; j < 3; ++j) rho(x, j) = x.coord(j);
    rho.toDevice();
    rho.updateHalo();
    rho.toHost();
    long err_periodic = 0, err_ghost = 0;
    auto check = [&](bool wrapped) {
        long err = 0;
        for (x.first(); x.test(); x.next())
            for (int k = 0; k < 3; ++k) {
                Site up = x, dn = x;
                for (int s = 1; s <= 3; ++s) {
                    up = up + k;
                    dn = dn - k;
                    const int c = x.coord(k);
                    const int wu = wrapped ? (c + s) % N : c + s;
                    const int wd = wrapped ? (c - s + N) % N : c - s;
                    if (rho(up, k) != (Real)wu) ++err;
                    if (rho(dn, k) != (Real)wd) ++err;
                }
            }
        return err;
    };
    err_periodic = check(true);
    // every ghost site takes its own coordinates (storage index -> coordinates)
    const long G = N + 2 * halo;
    for (long i = 0; i < G * G * G; ++i) {
        const int c[3] = {(int)(i % G) - halo, (int)((i / G) % G) - halo, (int)(i / (G * G)) - halo};
        const bool ghost = c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= N || c[1] >= N || c[2] >= N;
        if (ghost)
            for (int j = 0; j < 3; ++j) rho(i, j) = c[j];
    }
    err_ghost = check(false);
    std::printf("boundary_test_mi N=%d halo=%d %s: %ld periodic-image errors, %ld ghost-coordinate errors\n", N, halo,
                sizeof(Real) == 8 ? "double" : "single", err_periodic, err_ghost);
    const long err = err_periodic + err_ghost;
    return err > 255 ? 255 : (int)err;
}
