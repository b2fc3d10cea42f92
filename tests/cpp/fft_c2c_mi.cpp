// PlanFFT<Imag> of two complex fields through the C++ mirror.
//   - examples/fft_c2c_r2c_sametime.cpp:44-130 restated (BoxSize 16, khalo 1): the R2C pair
//     forward + backward of phi = 1 averages to 1 after dividing by N^3 twice; the C2C pair
//     on rho = 1 gives rho = N^3 at every site (forward then backward is N^3 times the
//     identity here; the reference's own backward is not an inverse and its example divides
//     by N^3 N^2 instead);
//   - a plane wave exp(2 pi i (a x + b y + c z)/N) transforms to N^3 at the single cKSite
//     (kx, ky, kz) = (a, b, c) (LATfield2_Site.hpp:219-263), for several (a, b, c);
//   - a random field comes back from forward + backward / N^3.
// Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "LATfield2_mi.hpp"

using namespace LATfield2;

int main(int argc, char **argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 16;
    const double tol = sizeof(Real) == 8 ? 1e-10 : 2e-5;
    int fails = 0;
    Lattice lat(3, N, 0);
    Lattice latKreal, latKcomplex;
    latKreal.initializeRealFFT(lat, 1);
    latKcomplex.initializeComplexFFT(lat, 1);
    Field<Real> phi(lat, 1);
    Field<Imag> phiK(latKreal, 1);
    PlanFFT<Imag> planReal(&phi, &phiK);
    Field<Imag> rho(lat, 1);
    Field<Imag> rhoK(latKcomplex, 1);
    PlanFFT<Imag> planComplex(&rho, &rhoK);
    const double l3 = (double)N * N * N;
    Site x(lat);
    for (x.first(); x.test(); x.next()) phi(x) = 1.0;
    phi.toDevice();
    planReal.execute(FFT_FORWARD);
    planReal.execute(FFT_BACKWARD);
    phi.toHost();
    double avreal = 0.0;
    for (x.first(); x.test(); x.next()) avreal += phi(x);
    avreal *= 1.0 / l3 / l3;
    for (x.first(); x.test(); x.next()) {
        rho(x).real() = 1.0;
        rho(x).imag() = 0.0;
    }
    rho.toDevice();
    planComplex.execute(FFT_FORWARD);
    planComplex.execute(FFT_BACKWARD);
    rho.toHost();
    double avcom = 0.0, emax = 0.0;
    for (x.first(); x.test(); x.next()) {
        avcom += rho(x).real();
        emax = std::fmax(emax, std::fabs(rho(x).real() / l3 - 1.0) + std::fabs(rho(x).imag() / l3));
    }
    avcom *= 1.0 / l3 / l3;
    std::printf("Average real = %.12g Average complex = %.12g (max site error %.2e)\n", avreal, avcom, emax);
    if (std::fabs(avreal - 1) > tol || std::fabs(avcom - 1) > tol || emax > tol) ++fails;
    // plane waves land on one cKSite
    cKSite k(latKcomplex);
    const int modes[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {3, N - 5, 2}};
    for (const auto &m : modes) {
        for (x.first(); x.test(); x.next()) {
            const double a = 2.0 * M_PI * (m[0] * x.coord(0) + m[1] * x.coord(1) + m[2] * x.coord(2)) / N;
            rho(x) = Imag((Real)std::cos(a), (Real)std::sin(a));
        }
        rho.toDevice();
        planComplex.execute(FFT_FORWARD);
        rhoK.toHost();
        double err = 0.0;
        for (k.first(); k.test(); k.next()) {
            const bool hit = k.coord(0) == m[0] && k.coord(1) == m[1] && k.coord(2) == m[2];
            const Imag v = rhoK(k);
            err = std::fmax(err, std::fabs(v.real() / l3 - (hit ? 1.0 : 0.0)) + std::fabs(v.imag() / l3));
        }
        std::printf("plane wave (%d %d %d): max error %.2e\n", m[0], m[1], m[2], err);
        if (err > tol) ++fails;
    }
    // random field round trip
    srand(7);
    for (x.first(); x.test(); x.next())
        rho(x) = Imag((Real)(rand() / (double)RAND_MAX - 0.5), (Real)(rand() / (double)RAND_MAX - 0.5));
    rho.toDevice();
    Field<Imag> keep(lat, 1);
    for (x.first(); x.test(); x.next()) keep(x) = rho(x);
    planComplex.execute(FFT_FORWARD);
    planComplex.execute(FFT_BACKWARD);
    rho.toHost();
    double rerr = 0.0;
    for (x.first(); x.test(); x.next())
        rerr = std::fmax(rerr, std::fabs(rho(x).real() / l3 - keep(x).real()) + std::fabs(rho(x).imag() / l3 - keep(x).imag()));
    std::printf("random round trip: max error %.2e\n", rerr);
    if (rerr > tol) ++fails;
    std::printf("fft_c2c_mi N=%d %s: %d failed checks\n", N, sizeof(Real) == 8 ? "double" : "single", fails);
    return fails;
}
