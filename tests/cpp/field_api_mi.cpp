// Field, Site and PlanFFT semantics of the drop-in header that reference drivers rely on,
// each checked against the reference's own formula:
//   - matrix components (LATfield2_Field_decl.hpp:773-806): unsymmetric (i, j) at j*rows + i,
//     symmetric at the upper-triangle order, matrix k of nMatrix at + k * matrixSize;
//   - operator()(Site) is data_[index] and operator()(Site, c) is data_[index*C + c]
//     (:760-770, :808-818), data() is the host array (:1984);
//   - host writes reach the device kernels and device results reach host reads with no
//     explicit transfer; updateHalo fills every ghost with its periodic image (:889-935),
//     also on a non-cubic lattice, visited with haloFirst/haloNext and forallboundary
//     (LATfield2_Site.hpp:99-116, looping_macro.hpp:10-21);
//   - FFT_IN_PLACE shares one buffer between the real and the k field (PlanFFT_decl.hpp:
//     804-824) and still transforms correctly;
//   - Field::write / read round trip (LATfield2_Field_decl.hpp:1179-1340, one process);
//   - the one-rank parallel object (LATfield2_parallel2d.hpp:42-314).
// Exit code = number of failed checks.
#include <cstdio>

#include "LATfield2.hpp"

using namespace LATfield2;

static int fails = 0;
static void check(bool ok, const char *what) {
    std::printf("%-72s %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) ++fails;
}

static double value_at(int x, int y, int z, int c) { return 1.0 + x + 100.0 * y + 10000.0 * z + 0.25 * c; }

int main(int argc, char **argv) {
    int n = 1, m = 1;
    parallel.initialize(n, m);
    const int N = argc > 1 ? atoi(argv[1]) : 12;
    const double tol = sizeof(Real) == 8 ? 1e-9 : 1e-3;

    // ---- the one-rank process grid
    {
        double s = 2.5;
        int a[3] = {1, 2, 3};
        parallel.sum(s);
        parallel.max(a, 3);
        parallel.broadcast(s, 0);
        long msg = 77, got = 0;
        parallel.send(msg, 0);
        parallel.receive(got, 0);
        check(parallel.size() == 1 && parallel.rank() == 0 && parallel.isRoot() && parallel.grid_size()[0] == 1 &&
                  parallel.grid_size()[1] == 1 && parallel.grid_rank()[0] == 0 && parallel.last_proc()[1] && s == 2.5 &&
                  a[2] == 3 && got == 77,
              "parallel: one rank, collectives are identities, self messages");
    }

    Lattice lat(3, N, 2);
    Site x(lat);

    // ---- matrix components
    {
        Field<Real> A(lat, 3, 3, unsymmetric);
        Field<Real> R(lat, 2, 3, unsymmetric);  // 2 rows, 3 columns
        Field<Real> S(lat, 3, 3, symmetric);
        Field<Real> V(lat, 2, 3, 3, symmetric);  // two symmetric 3x3 matrices
        bool ok = A.components() == 9 && R.components() == 6 && S.components() == 6 && V.components() == 12 &&
                  V.nMatrix() == 2 && V.rows() == 3 && V.cols() == 3 && V.symmetry() == symmetric;
        for (x.first(); x.test(); x.next()) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) A(x, i, j) = 10 * i + j + x.index();
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 3; ++j) R(x, i, j) = 10 * i + j;
            for (int i = 0; i < 3; ++i)
                for (int j = i; j < 3; ++j) S(x, i, j) = 10 * i + j;
            for (int k = 0; k < 2; ++k)
                for (int i = 0; i < 3; ++i)
                    for (int j = i; j < 3; ++j) V(x, k, i, j) = 100 * k + 10 * i + j;
        }
        const int sym_order[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        Real *a = A.data(), *r = R.data(), *s = S.data(), *v = V.data();
        for (x.first(); x.test(); x.next()) {
            const long i0 = x.index();
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ok = ok && a[i0 * 9 + j * 3 + i] == (Real)(10 * i + j + i0);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 3; ++j) ok = ok && r[i0 * 6 + j * 2 + i] == (Real)(10 * i + j);
            for (int c = 0; c < 6; ++c) {
                const int i = sym_order[c][0], j = sym_order[c][1];
                ok = ok && s[i0 * 6 + c] == (Real)(10 * i + j) && S(x, j, i) == S(x, i, j);
                for (int k = 0; k < 2; ++k) ok = ok && v[i0 * 12 + 6 * k + c] == (Real)(100 * k + 10 * i + j);
            }
        }
        check(ok, "matrix fields: unsymmetric (i,j) at j*rows+i, symmetric order, nMatrix");
    }

    // ---- host writes -> device ghost fill -> host reads, no explicit transfers
    {
        Field<Real> f(lat, 2);
        for (x.first(); x.test(); x.next())
            for (int c = 0; c < 2; ++c) f(x, c) = value_at(x.coord(0), x.coord(1), x.coord(2), c);
        f.updateHalo();
        long ghosts = 0, bad = 0;
        forallboundary_start(lat, g)
            ++ghosts;
            const int px = (g.coord(0) + N) % N, py = (g.coord(1) + N) % N, pz = (g.coord(2) + N) % N;
            for (int c = 0; c < 2; ++c)
                if (f(g, c) != (Real)value_at(px, py, pz, c)) ++bad;
        forallboundary_stop
        const long G = N + 4;
        check(ghosts == G * G * G - (long)N * N * N && bad == 0,
              "updateHalo: every ghost of a halo-2 field is its periodic image (forallboundary)");
        // operator()(Site) is data_[index], as in the reference
        x.first();
        x.setCoord(3, 4, 5);
        check(&f(x) == f.data() + x.index() && &f(x, 1) == f.data() + 2 * x.index() + 1,
              "operator()(Site) is data_[index], operator()(Site, c) is data_[index*C + c]");
        // the device copy follows host writes made through data()
        Real *h = f.data();
        h[2 * x.index()] = -7;
        f.updateHalo();
        Site xg(lat);
        xg.setCoord(3, 4, 5);
        check(f(xg, 0) == -7, "a host write through data() is uploaded before the next device operation");
        // Site coordinates and neighbours
        Site s1(lat);
        s1.setCoord(N - 1, 0, 2);
        check(s1.coord(0) == N - 1 && s1.coord(1) == 0 && s1.coord(2) == 2 && (s1 + 0).coord(0) == N &&
                  (s1 - 1).coord(1) == -1 && s1.move(2, 3).coord(2) == 5,
              "Site: setCoord, coord, operator+/-, move");
    }

    // ---- halo iteration order: every ghost site once, in storage order
    {
        Lattice l1(3, 5, 1);
        Site h(l1);
        long count = 0, prev = -1;
        bool ordered = true, ghost = true;
        for (h.haloFirst(); h.haloTest(); h.haloNext()) {
            ordered = ordered && h.index() > prev;
            prev = h.index();
            bool in = true;
            for (int d = 0; d < 3; ++d) in = in && h.coord(d) >= 0 && h.coord(d) < 5;
            ghost = ghost && !in;
            ++count;
        }
        check(count == 7 * 7 * 7 - 5 * 5 * 5 && ordered && ghost, "haloFirst/haloTest/haloNext: all ghost sites, once, in order");
    }

    // ---- non-cubic lattice (the reference's gettingStarted uses 25 x 57 x 32)
    {
        int sz[3] = {5, 7, 6};
        Lattice ln(3, sz, 1);
        Field<Real> f(ln, 1);
        Site y(ln);
        for (y.first(); y.test(); y.next()) f(y) = value_at(y.coord(0), y.coord(1), y.coord(2), 0);
        f.updateHalo();
        long bad = 0, ghosts = 0;
        forallboundary_start(ln, g)
            ++ghosts;
            if (f(g) != (Real)value_at((g.coord(0) + 5) % 5, (g.coord(1) + 7) % 7, (g.coord(2) + 6) % 6, 0)) ++bad;
        forallboundary_stop
        check(ghosts == 7 * 9 * 8 - 5 * 7 * 6 && bad == 0, "updateHalo on a 5 x 7 x 6 lattice");
    }

    // ---- FFT_IN_PLACE
    {
        Lattice l(3, N, 1), lk;
        lk.initializeRealFFT(l, 0);
        Field<Real> r;
        r.initialize(l, 1);
        Field<Imag> kf;
        kf.initialize(lk, 1);
        PlanFFT<Imag> plan(&r, &kf, FFT_IN_PLACE);
        check((void *)r.data() == (void *)kf.data(), "FFT_IN_PLACE: the real and the k field share one buffer");
        Site y(l);
        const int a = 1, b = 2, c = 3;
        for (y.first(); y.test(); y.next())
            r(y) = cos(2 * M_PI * (a * y.coord(0) + b * y.coord(1) + c * y.coord(2)) / N);
        plan.execute(FFT_FORWARD);
        rKSite k(lk);
        double err = 0.0;
        const double n3 = (double)N * N * N;
        for (k.first(); k.test(); k.next()) {
            // cos(2 pi (a x + b y + c z)/N): N^3/2 at (a, b, c); the half spectrum keeps kx = a only
            const bool hit = k.coord(0) == a && k.coord(1) == b && k.coord(2) == c;
            err = max(err, fabs(kf(k).real() / n3 - (hit ? 0.5 : 0.0)) + fabs(kf(k).imag() / n3));
        }
        check(err < tol, "FFT_IN_PLACE forward: one mode at (1, 2, 3) with N^3/2");
        plan.execute(FFT_BACKWARD);
        double err2 = 0.0;
        for (y.first(); y.test(); y.next())
            err2 = max(err2, fabs(r(y) / n3 - cos(2 * M_PI * (a * y.coord(0) + b * y.coord(1) + c * y.coord(2)) / N)));
        check(err2 < tol, "FFT_IN_PLACE backward: N^3 times the input");
    }

    // ---- write / read
    {
        Field<Real> f(lat, 3), g(lat, 3);
        for (x.first(); x.test(); x.next())
            for (int c = 0; c < 3; ++c) f(x, c) = value_at(x.coord(0), x.coord(1), x.coord(2), c);
        const std::string path = "field_api_mi.bin";
        f.write(path);
        g.read(path);
        bool ok = true;
        for (x.first(); x.test(); x.next())
            for (int c = 0; c < 3; ++c) ok = ok && g(x, c) == f(x, c);
        std::ifstream in(path.c_str(), std::ios::binary | std::ios::ate);
        ok = ok && (long)in.tellg() == (long)N * N * N * 3 * (long)sizeof(Real);
        std::remove(path.c_str());
        check(ok, "Field::write / read: interior sites in storage order, round trip");
    }

    std::printf("field_api_mi N=%d %s: %d failed checks\n", N, sizeof(Real) == 8 ? "double" : "single", fails);
    return fails;
}
