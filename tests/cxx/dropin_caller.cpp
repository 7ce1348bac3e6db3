// A C++ caller of the drop-in interface, written the way the reference's own tests call it
// (vlovero/xTGEVC3 tests/test_lapack_d.cpp:278-286, tests/test_lapack_z.cpp:199-205): it
// includes include/tgevc3.h, links libxtgevc3.so, and calls the C++-linkage dtgevc3 /
// ztgevc3 with the two-call lwork idiom (query with lwork = -1, allocate work[0], call).
//
//   dropin_caller early             argument / early-return cases, no GPU work; one line each:
//                                   "<case> m=<m> info=<info> work0=<work[0]>"
//   dropin_caller solve d|z N FILE  a seeded pencil in the reference tests' construction,
//                                   side = howmny = 'B' with VL = VR = I, written to FILE as
//                                   int32 n, m, info, then S, P, eigenvalues, VL, VR (column-major)
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "tgevc3.h"

typedef std::complex<double> zc;

static void real_pencil(int n, std::mt19937_64 &g, std::vector<double> &S, std::vector<double> &P,
                        std::vector<double> &ar, std::vector<double> &ai, std::vector<double> &be)
{
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::uniform_int_distribution<int> pct(0, 99);
    S.assign((size_t)n * n, 0.0);
    P.assign((size_t)n * n, 0.0);
    ar.assign(n, 0.0);
    ai.assign(n, 0.0);
    be.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i <= j; ++i) {
            S[i + (size_t)j * n] = u(g);
            P[i + (size_t)j * n] = u(g);
        }
        double &d = P[j + (size_t)j * n];
        d += d >= 0.0 ? 1.0 : -1.0;
    }
    for (int k = 0; k < n;) {
        if (k + 1 < n && pct(g) < 40) {
            P[k + (size_t)k * n] = P[k + 1 + (size_t)(k + 1) * n] = 1.0;
            P[k + (size_t)(k + 1) * n] = 0.0;
            const double a = S[k + (size_t)k * n];
            double b = S[k + (size_t)(k + 1) * n];
            if (b == 0.0) {
                b = 1.0;
                S[k + (size_t)(k + 1) * n] = b;
            }
            S[k + 1 + (size_t)(k + 1) * n] = a;
            S[k + 1 + (size_t)k * n] = -b;
            ar[k] = ar[k + 1] = a;
            ai[k] = std::abs(b);
            ai[k + 1] = -std::abs(b);
            be[k] = be[k + 1] = 1.0;
            k += 2;
        }
        else {
            ar[k] = S[k + (size_t)k * n];
            be[k] = P[k + (size_t)k * n];
            ++k;
        }
    }
}

static void complex_pencil(int n, std::mt19937_64 &g, std::vector<zc> &S, std::vector<zc> &P, std::vector<zc> &al,
                           std::vector<zc> &be)
{
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    S.assign((size_t)n * n, 0.0);
    P.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i <= j; ++i) {
            S[i + (size_t)j * n] = zc(u(g), u(g));
            P[i + (size_t)j * n] = zc(u(g), u(g));
        }
        zc &d = P[j + (size_t)j * n];
        d += d.real() >= 0.0 ? 1.0 : -1.0;
    }
    al.resize(n);
    be.resize(n);
    for (int k = 0; k < n; ++k) {
        al[k] = S[k + (size_t)k * n];
        be[k] = P[k + (size_t)k * n];
    }
}

static int early()
{
    const int n = 8;
    std::mt19937_64 g(7);
    std::vector<double> S, P, ar, ai, be;
    real_pencil(n, g, S, P, ar, ai, be);
    std::vector<double> VL((size_t)n * n), VR((size_t)n * n);
    double q = -1.0;
    int m = -7, info = -99;
    dtgevc3('B', 'B', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
            VR.data(), n, n, &m, &q, -1, &info);
    std::printf("d_query m=%d info=%d work0=%.17g\n", m, info, q);
    const int lwork = (int)q;
    std::vector<double> work(lwork);
    m = -7;
    dtgevc3('B', 'X', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
            VR.data(), n, n, &m, work.data(), lwork, &info);
    std::printf("d_bad_howmny m=%d info=%d work0=0\n", m, info);
    m = -7;
    dtgevc3('R', 'A', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
            VR.data(), n, n - 1, &m, work.data(), lwork, &info);
    std::printf("d_mm_short m=%d info=%d work0=0\n", m, info);
    m = -7;
    dtgevc3('R', 'A', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
            VR.data(), n, n, &m, work.data(), 3, &info);
    std::printf("d_lwork_short m=%d info=%d work0=0\n", m, info);
    m = -7;
    dtgevc3('R', 'A', nullptr, 0, S.data(), 1, P.data(), 1, ar.data(), ai.data(), be.data(), VL.data(), 1,
            VR.data(), 1, 0, &m, work.data(), lwork, &info);
    std::printf("d_n0 m=%d info=%d work0=0\n", m, info);
    m = -7;
    dtgevc3('R', 'S', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
            VR.data(), n, n, &m, work.data(), lwork, &info);
    std::printf("d_select_null m=%d info=%d work0=0\n", m, info);

    std::vector<zc> Sz, Pz, al, bz;
    complex_pencil(n, g, Sz, Pz, al, bz);
    std::vector<zc> VLz((size_t)n * n), VRz((size_t)n * n);
    zc qz(-1.0, 0.0);
    m = -7;
    ztgevc3('B', 'B', nullptr, n, Sz.data(), n, Pz.data(), n, al.data(), bz.data(), VLz.data(), n, VRz.data(), n, n,
            &m, &qz, -1, &info);
    std::printf("z_query m=%d info=%d work0=%.17g\n", m, info, qz.real());
    const int lz = (int)qz.real();
    std::vector<zc> wz(lz);
    m = -7;
    ztgevc3('B', 'Q', nullptr, n, Sz.data(), n, Pz.data(), n, al.data(), bz.data(), VLz.data(), n, VRz.data(), n, n,
            &m, wz.data(), lz, &info);
    std::printf("z_bad_howmny m=%d info=%d work0=0\n", m, info);
    m = -7;
    ztgevc3('L', 'A', nullptr, n, Sz.data(), n, Pz.data(), n, al.data(), bz.data(), VLz.data(), n, VRz.data(), n, 2,
            &m, wz.data(), lz, &info);
    std::printf("z_mm_short m=%d info=%d work0=0\n", m, info);
    m = -7;
    ztgevc3('L', 'A', nullptr, n, Sz.data(), n, Pz.data(), n, al.data(), bz.data(), VLz.data(), n, VRz.data(), n, n,
            &m, wz.data(), 5, &info);
    std::printf("z_lwork_short m=%d info=%d work0=0\n", m, info);
    return 0;
}

template <class T> static void put(FILE *f, const std::vector<T> &v) { std::fwrite(v.data(), sizeof(T), v.size(), f); }

static int solve(char prefix, int n, const char *path)
{
    std::mt19937_64 g(1000 + n);
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        std::perror(path);
        return 2;
    }
    int m = -7, info = -99;
    if (prefix == 'd') {
        std::vector<double> S, P, ar, ai, be;
        real_pencil(n, g, S, P, ar, ai, be);
        std::vector<double> VL((size_t)n * n, 0.0), VR((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) {
            VL[i + (size_t)i * n] = VR[i + (size_t)i * n] = 1.0;
        }
        double q = 0.0;
        dtgevc3('B', 'B', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
                VR.data(), n, n, &m, &q, -1, &info);
        std::vector<double> work((size_t)q);
        dtgevc3('B', 'B', nullptr, n, S.data(), n, P.data(), n, ar.data(), ai.data(), be.data(), VL.data(), n,
                VR.data(), n, n, &m, work.data(), (int)q, &info);
        const int hdr[3] = {n, m, info};
        std::fwrite(hdr, sizeof(int), 3, f);
        put(f, S), put(f, P), put(f, ar), put(f, ai), put(f, be), put(f, VL), put(f, VR);
    }
    else {
        std::vector<zc> S, P, al, be;
        complex_pencil(n, g, S, P, al, be);
        std::vector<zc> VL((size_t)n * n, 0.0), VR((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) {
            VL[i + (size_t)i * n] = VR[i + (size_t)i * n] = 1.0;
        }
        zc q = 0.0;
        ztgevc3('B', 'B', nullptr, n, S.data(), n, P.data(), n, al.data(), be.data(), VL.data(), n, VR.data(), n, n,
                &m, &q, -1, &info);
        std::vector<zc> work((size_t)q.real());
        ztgevc3('B', 'B', nullptr, n, S.data(), n, P.data(), n, al.data(), be.data(), VL.data(), n, VR.data(), n, n,
                &m, work.data(), (int)q.real(), &info);
        const int hdr[3] = {n, m, info};
        std::fwrite(hdr, sizeof(int), 3, f);
        put(f, S), put(f, P), put(f, al), put(f, be), put(f, VL), put(f, VR);
    }
    std::fclose(f);
    std::printf("solve %c n=%d m=%d info=%d\n", prefix, n, m, info);
    return info == 0 ? 0 : 1;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "early") == 0) {
        return early();
    }
    if (argc >= 5 && std::strcmp(argv[1], "solve") == 0) {
        return solve(argv[2][0], std::atoi(argv[3]), argv[4]);
    }
    std::fprintf(stderr, "usage: %s early | solve d|z N FILE\n", argv[0]);
    return 2;
}
