// Host-only checks of the C++ mirror's geometry (no device calls): Site walks the interior
// in storage order with index = sum (coord + h) jump; rKSite maps asked directions
// (kx, ky, kz) to storage dims (0, 2, 1) in coord, operator+ and setCoord
// (LATfield2_Site.hpp:319-416); symmetric 3x3 fields use the component order
// (0,0)(0,1)(0,2)(1,1)(1,2)(2,2), unsymmetric ones (i, j) at j * rows + i.
// Exit code = number of failures.
#include <cstdio>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

int main() {
    int fails = 0;
    const int N = 6, h = 2;
    Lattice lat(3, N, h);
    Site x(lat);
    long n = 0;
    int ex = 0, ey = 0, ez = 0;
    for (x.first(); x.test(); x.next(), ++n) {
        if (x.coord(0) != ex || x.coord(1) != ey || x.coord(2) != ez) ++fails;
        const long G = N + 2 * h;
        if (x.index() != (ex + h) + G * ((ey + h) + G * (ez + h))) ++fails;
        if ((x + 2).index() - x.index() != G * G || (x - 1).index() - x.index() != -G) ++fails;
        if (++ex == N) { ex = 0; if (++ey == N) { ey = 0; ++ez; } }
    }
    if (n != (long)N * N * N) ++fails;
    Lattice latK;
    latK.initializeRealFFT(lat, 1);
    if (latK.size(0) != N / 2 + 1 || latK.size(1) != N || latK.size(2) != N) ++fails;
    rKSite k(latK);
    k.setCoord(1, 2, 3);  // asked coordinates (kx, ky, kz), stored at (kx, kz, ky)
    if (k.coord(0) != 1 || k.coord(1) != 2 || k.coord(2) != 3) ++fails;
    if (k.latCoord(0) != 1 || k.latCoord(1) != 3 || k.latCoord(2) != 2) ++fails;
    if ((k + 1).coord(1) != 3 || (k + 2).coord(2) != 4) ++fails;
    {
        const long KG = N + 2, R = N / 2 + 1 + 2;
        if (k.index() != (1 + 1) + R * ((3 + 1) + KG * (2 + 1))) ++fails;
    }
    Field<Real> T;
    T.initialize(lat, 3, 3, symmetric);
    const int want[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (T.component(i, j) != want[i][j]) ++fails;
    if (T.components() != 6) ++fails;
    Field<Real> U;
    U.initialize(lat, 3, 3, unsymmetric);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (U.component(i, j) != j * 3 + i) ++fails;
    Field<Real> V;
    V.initialize(lat, 2, 3, 3, symmetric);
    if (V.components() != 12 || V.component(1, 2, 1) != 6 + 4) ++fails;
    // halo walk: every ghost site once
    long ghosts = 0;
    Site g(lat);
    for (g.haloFirst(); g.haloTest(); g.haloNext()) ++ghosts;
    if (ghosts != (long)(N + 2 * h) * (N + 2 * h) * (N + 2 * h) - (long)N * N * N) ++fails;
    std::printf("shim_host_check: %d failures\n", fails);
    return fails;
}
