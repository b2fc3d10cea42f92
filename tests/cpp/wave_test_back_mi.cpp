// Known-answer test of PlanFFT::execute(FFT_BACKWARD) through the C++ mirror: restates
// benchmarks_tests/wave_test_back.cpp:76-138 (phiK = cos(2 pi i kx/N) cos(2 pi j ky/N) cos(2 pi l kz/N)
// in rKSite coordinates; expected phi/N^3 = 1, 1/2, 1/4, 1/8 at x = +-i, y = +-j, z = +-l).
#include <cmath>
#include <cstdio>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

int main(int argc, char **argv) {
    int N = argc > 1 ? atoi(argv[1]) : 64;
    const double tol = sizeof(Real) == 8 ? 1.0e-10 : 2.0e-6;
    Lattice lat(3, N, 1);
    Lattice latK;
    latK.initializeRealFFT(lat, 0);
    Field<Real> phi(lat, 1);
    Field<Imag> phiK(latK, 1);
    PlanFFT<Imag> plan(&phi, &phiK);
    Site x(lat);
    rKSite k(latK);
    long count = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int l = 0; l < 4; ++l) {
                for (k.first(); k.test(); k.next())
                    phiK(k) = Imag((Real)(std::cos(2.0 * M_PI * i * k.coord(0) / N) * std::cos(2.0 * M_PI * j * k.coord(1) / N) *
                                          std::cos(2.0 * M_PI * l * k.coord(2) / N)),
                                   0);
                phiK.toDevice();
                plan.execute(FFT_BACKWARD);
                phi.toHost();
                const int nz = (i != 0) + (j != 0) + (l != 0);
                const double amp = 1.0 / (1 << nz);
                for (x.first(); x.test(); x.next()) {
                    double want = 0.0;
                    if ((x.coord(0) == i || x.coord(0) == N - i) && (x.coord(1) == j || x.coord(1) == N - j) &&
                        (x.coord(2) == l || x.coord(2) == N - l))
                        want = amp;
                    if (std::fabs(phi(x) / lat.sites() - want) > tol) {
                        if (count < 10) std::printf(" %d-%d-%d  %g exceeding tolerance\n", x.coord(0), x.coord(1), x.coord(2), (double)phi(x) / lat.sites());
                        ++count;
                    }
                }
            }
    std::printf("wave_test_back_mi N=%d %s: %ld sites outside %g\n", N, sizeof(Real) == 8 ? "double" : "single", count, tol);
    return count > 255 ? 255 : (int)count;
}
