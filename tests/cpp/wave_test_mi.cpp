// Known-answer test of PlanFFT::execute(FFT_FORWARD) through the C++ mirror: restates
// benchmarks_tests/wave_test.cpp:72-152 (inputs cos(2 pi i x/N) cos(2 pi j y/N) cos(2 pi l z/N),
// expected amplitudes 1, 1/2, 1/4, 1/8 at (kx=i, ky=+-j, kz=+-l) read through rKSite).
// Exit code = number of k-sites outside the tolerance.
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
                for (x.first(); x.test(); x.next())
                    phi(x) = (Real)(std::cos(2.0 * M_PI * i * x.coord(0) / N) * std::cos(2.0 * M_PI * j * x.coord(1) / N) *
                                    std::cos(2.0 * M_PI * l * x.coord(2) / N));
                phi.toDevice();
                plan.execute(FFT_FORWARD);
                phiK.toHost();
                const int nz = (i != 0) + (j != 0) + (l != 0);
                const double amp = 1.0 / (1 << nz);
                for (k.first(); k.test(); k.next()) {
                    double want = 0.0;
                    if (k.coord(0) == i && (k.coord(1) == j || k.coord(1) == N - j) &&
                        (k.coord(2) == l || k.coord(2) == N - l))
                        want = amp;
                    const Imag v = phiK(k);
                    if (std::fabs(v.real() / lat.sites() - want) > tol || std::fabs(v.imag() / lat.sites()) > tol) {
                        if (count < 10)
                            std::printf(" %d-%d-%d  %g+%gi exceeding tolerance (mode %d %d %d)\n", k.coord(0), k.coord(1),
                                        k.coord(2), (double)v.real() / lat.sites(), (double)v.imag() / lat.sites(), i, j, l);
                        ++count;
                    }
                }
            }
    std::printf("wave_test_mi N=%d %s: %ld sites outside %g\n", N, sizeof(Real) == 8 ? "double" : "single", count, tol);
    return count > 255 ? 255 : (int)count;
}
