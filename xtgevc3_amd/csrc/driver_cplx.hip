// Host driver of the complex128 xTGEVC3 path (ztgevc3; ctgevc3 through it).
//
// Reference behaviour restated (src/ztgevc3.cpp):
//   selection: one column per selected eigenvalue            :476-482
//   global scales from |z| column sums over rows <= col      :528-545
//   complex shift coefficients acoeff, bcoeff                :145-153, :619-627
//   right: (a S - b P) x = 0, x_k = 1                       :134-210
//   left : y^H (a S - b P) = 0, y_k = 1, via conj(a), conj(b) :281-357, :775-780
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <limits>
#include <mutex>
#include <vector>

#include "../../include/xtgevc3.h"
#include "common.h"
#include "driver.h"

namespace xtg {

__global__ void k_round_f32(const double *, long, long, float *, long);
template <int MAXB>
__global__ void kz_solve_units(const double2 *, long, const double2 *, long, int, int, double2 *, long, double2 *, long,
                               const ZColDesc *, int, int, double *, int *, int *, int *, int, int, double *, int,
                               GroupInfo, SplitOut);
__global__ void k_slice_digits(const double *, long, int, int, const double *, long, int, int, int, int, const int *,
                               signed char *, long, long, int, signed char *);
__global__ void kz_rowexp(const double2 *, long, const double2 *, long, int, int, int *, const int *);
__global__ void k_fill_int(int *, long, int);
__global__ void k_xexp_final(const int *, const int *, long, const int *, const int *, int, int, int *);
int launch_gemm_i8(hipStream_t st, int M, int N, int K1, int K2, const signed char *L1, const signed char *L2,
                   long lslice, long kbl, const signed char *R, long rslice, long kbr, const int *EL, const int *ER,
                   double *C, long ldc, int cst, const int *outcol, int mode, const double *sig, const int *kmax);
template <int TAG, int TA>
__global__ void kz_gemm(int, int, int, int, const double2 *, long, const double2 *, long, const double2 *, long,
                        double2 *, long, const int *, int, const double *, const int *, const double2 *);
__global__ void kz_fixup_scale(double2 *, long, int, int, const int *, const int *, long, const int *, const int *);
__global__ void kz_copy_out(const double2 *, long, int, int, double2 *, long, const int *, int);
__global__ void kz_normalize(double2 *, long, int);
__global__ void kz_reverse_conjtranspose(const double2 *, long, int, double2 *, long);
__global__ void kz_colnorms(const double2 *, long, const double2 *, long, int, double *);
__global__ void kz_to_rowmajor(const double2 *, long, int, int, double2 *, long);
__global__ void kz_tgsna(const double2 *, long, const double2 *, long, const double2 *, const double2 *, long, int,
                         double *);
__global__ void k_zero(double *, long);
__global__ void k_zero_int(int *, long);

constexpr int NB_Z = 32;
constexpr int ZGROUP = 8;        // row blocks per far update (K = 2 * 256)
constexpr int ZCOLS_PER_BLOCK = 16; // solve kernel: 16 lanes per column, 256 threads
constexpr int ZTA_H = 128, ZTB_H = 64;
static_assert(ZGROUP <= GROUP_MAX_BLOCKS, "group size");
// split-int8 form: complex entries expand to (re, im) pairs along K; K step 32 of the
// expanded index; K extent of the group operand (two segments of up to 2 * 256 expanded rows
// shifted to the K step boundary: 544 each)
constexpr int ZSPLIT_KSTEP = 32;
constexpr long ZWG_K = 2 * 544;
static_assert(2 * ZGROUP * NB_Z + ZSPLIT_KSTEP - 2 <= 544, "split operand extent");
static inline int zround_up(int v, int q) { return (v + q - 1) / q * q; }

int complex_units(char howmny, const int *select, int n, std::vector<Unit> &units)
{
    bool do_all = howmny == 'A' || howmny == 'B';
    bool do_sel = howmny == 'S';
    units.clear();
    int m = 0;
    for (int k = 0; k < n; ++k) {
        if (do_all || (do_sel && select != nullptr && select[k])) {
            units.push_back(Unit{k, 1, m});
            ++m;
        }
    }
    return m;
}

namespace {

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) {
            return hipSuccess;
        }
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) {
                return e;
            }
            p = nullptr;
            cap = 0;
        }
        size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) {
            cap = b;
        }
        return e;
    }
    template <class T> T *as() { return static_cast<T *>(p); }
};

struct ZWork {
    DBuf X, W, sig, sig_grp, colexp, blkexp, flag, desc, outcol, kmax, rowstep, Sw, Pw, Zt, norms;
    // split-int8 form: pencil digits expanded as (re, -im) and (im, re), group operand, X and Z
    // digits, exponents
    DBuf Sx, Sy, Px, Py, rowexp, Wg, wgexp, wbexp, xbexp, sstart, xexp, Xx, Xy, Zq, zexp, kmax2, cfar;
    DBuf Vt, YA, YB; // eigenvalue condition numbers
    // far-update lookahead: the rows above the next group run on `fs`
    hipStream_t fs = nullptr;
    hipEvent_t ev_wready[2] = {nullptr, nullptr}, ev_fdone[2] = {nullptr, nullptr};
    hipEvent_t ev_last = nullptr; // end of the previous call's work on this workspace
};

// Orders a call after the previous call's work on the device's workspace, whatever stream
// that ran on, and records the end of this call when it goes out of scope.
struct ZWsOrder {
    hipEvent_t *ev;
    hipStream_t st;
    bool on = false; // off while st is being captured into a graph (the graph orders itself)
    hipError_t begin()
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            return hipSuccess;
        }
        on = true;
        if (*ev == nullptr) {
            return hipEventCreateWithFlags(ev, hipEventDisableTiming);
        }
        return hipStreamWaitEvent(st, *ev, 0);
    }
    ~ZWsOrder()
    {
        if (on && *ev != nullptr) {
            (void)hipEventRecord(*ev, st);
        }
    }
};

std::mutex g_zmutex;
ZWork g_zws[64];

typedef std::complex<double> zt;

// complex shift coefficients, in the reference's operation order (ztgevc3.cpp:145-153)
// safemin: smallest normal number of the interface precision (FLT_MIN for ctgevc3)
void shift_cplx(zt al, zt bv, double ascale, double bscale, zt &a, zt &b, double safemin)
{
    double t = 1.0 / std::max(std::max(std::abs(al) * ascale, std::abs(bv) * bscale), safemin);
    a = zt(t * ascale, 0.0) * bv * zt(bscale, 0.0);
    b = zt(t * bscale, 0.0) * al * zt(ascale, 0.0);
    if (std::abs(bv) <= safemin && std::abs(al) > safemin) {
        a = zt(0.0, 0.0);
        b = zt(1.0, 0.0);
    }
}

#define ZENSURE(buf, bytes) XTG_CHECK((buf).ensure(bytes))

// Creates the far-update lookahead stream and its events once per device workspace.
static hipError_t zfar_stream_init(ZWork &ws)
{
    if (ws.fs == nullptr) {
        hipError_t e = hipStreamCreateWithFlags(&ws.fs, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ws.fs = nullptr;
            return e;
        }
        for (int b = 0; b < 2; ++b) {
            for (hipEvent_t *ev : {&ws.ev_wready[b], &ws.ev_fdone[b]}) {
                e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
                if (e != hipSuccess) {
                    return e;
                }
            }
        }
    }
    return hipSuccess;
}

int run_cplx_side(ZWork &ws, int n, const double2 *Sw, long ldsw, const double2 *Pw, long ldpw, const zt *al,
                  const zt *be, double ascale, double bscale, bool left, bool back, const double2 *Z, long ldz,
                  double2 *V, long ldv, const std::vector<Unit> &units, int m, hipStream_t st, int prec,
                  const ZFeed *zf, double safemin)
{
    const bool split = prec == XTGEVC3_PREC_SPLIT_I8;
    std::vector<ZColDesc> desc(m);
    std::vector<int> outcol(m), kmaxv(m);
    for (int p = 0; p < m; ++p) {
        const Unit &un = left ? units[m - 1 - p] : units[p];
        zt a, b;
        shift_cplx(al[un.k], be[un.k], ascale, bscale, a, b, safemin);
        if (left) {
            a = std::conj(a);
            b = std::conj(b);
        }
        ZColDesc c{};
        c.aR = a.real();
        c.aI = a.imag();
        c.bR = b.real();
        c.bI = b.imag();
        c.d = left ? n - 1 - un.k : un.k;
        c.lo = 0;
        c.flags = INIT_UNIT << 2;
        desc[p] = c;
        outcol[p] = un.outc;
        kmaxv[p] = c.d + 1;
    }
    std::vector<int> i0s, nbs;
    for (int cur = n; cur > 0;) {
        int i = std::max(0, cur - NB_Z);
        i0s.push_back(i);
        nbs.push_back(cur - i);
        cur = i;
    }
    const int nsteps = (int)i0s.size();
    std::vector<int> rowstep(n);
    for (int s = 0; s < nsteps; ++s) {
        for (int r = i0s[s]; r < i0s[s] + nbs[s]; ++r) {
            rowstep[r] = s;
        }
    }
    const long ldx = ((long)m + 7) / 8 * 8;
    ZENSURE(ws.X, sizeof(double2) * (size_t)n * ldx);
    const long ldw = ldx;
    // group operands, group scales (and their digits / exponents) alternate between two copies
    // by group parity: the lookahead far update of one group reads its copy while the next
    // group's steps write the other
    const size_t wsz = (size_t)(2 * ZGROUP * NB_Z) * ldw;
    ZENSURE(ws.W, sizeof(double2) * 2 * wsz);
    ZENSURE(ws.sig, sizeof(double) * (size_t)m);
    ZENSURE(ws.sig_grp, sizeof(double) * 2 * (size_t)m);
    ZENSURE(ws.colexp, sizeof(int) * (size_t)m);
    ZENSURE(ws.blkexp, sizeof(int) * (size_t)nsteps * ldx);
    ZENSURE(ws.flag, sizeof(int) * 4);
    ZENSURE(ws.desc, sizeof(ZColDesc) * (size_t)m);
    ZENSURE(ws.outcol, sizeof(int) * (size_t)m);
    ZENSURE(ws.kmax, sizeof(int) * (size_t)m);
    ZENSURE(ws.rowstep, sizeof(int) * (size_t)n);
    XTG_CHECK(hipMemcpyAsync(ws.desc.p, desc.data(), sizeof(ZColDesc) * m, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.outcol.p, outcol.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.kmax.p, kmaxv.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.rowstep.p, rowstep.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    double2 *X = ws.X.as<double2>();
    double2 *W = ws.W.as<double2>();
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, (double *)X, 2L * n * ldx); XTG_LAUNCH_CHECK("k_zero");
    hipLaunchKernelGGL(k_zero_int, dim3(256), dim3(256), 0, st, ws.blkexp.as<int>(), (long)nsteps * ldx); XTG_LAUNCH_CHECK("k_zero_int");
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(256), 0, st, ws.colexp.as<int>(), (long)m); XTG_LAUNCH_CHECK("k_zero_int");
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, ws.flag.as<int>(), 4L); XTG_LAUNCH_CHECK("k_zero_int");
    long qld = 0, qsl = 0, kbs = 0;
    if (split) {
        // digits of the pencil rows, complex entries expanded along K (exponent per row over
        // |Re|, |Im| of S and P)
        qld = zround_up(2 * n, ZSPLIT_KSTEP) + ZSPLIT_KSTEP;
        qsl = (long)n * qld;
        kbs = (long)n * ZSPLIT_KSTEP;
        for (DBuf *b : {&ws.Sx, &ws.Sy, &ws.Px, &ws.Py}) {
            ZENSURE(*b, (size_t)SPLIT_DIGITS * qsl);
        }
        ZENSURE(ws.rowexp, sizeof(int) * (size_t)n);
        ZENSURE(ws.Wg, 2 * (size_t)SPLIT_DIGITS * m * ZWG_K);
        ZENSURE(ws.wgexp, sizeof(int) * 2 * (size_t)m);
        ZENSURE(ws.wbexp, sizeof(int) * (size_t)nsteps * ldx);
        ZENSURE(ws.xbexp, sizeof(int) * (size_t)nsteps * ldx);
        ZENSURE(ws.sstart, sizeof(int) * (size_t)m);
        ZENSURE(ws.xexp, sizeof(int) * (size_t)m);
        std::vector<int> sstart(m);
        for (int p = 0; p < m; ++p) {
            sstart[p] = rowstep[desc[p].d];
        }
        XTG_CHECK(hipMemcpyAsync(ws.sstart.p, sstart.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        // row exponents over each row's far-update region only: columns from the bottom edge of
        // the row's group on (the far updates of a row read no other columns; digits left of
        // it may saturate and only meet zero digits of the group operand)
        std::vector<int> cfar(n, n);
        for (int s = 0; s < nsteps; ++s) {
            const int gs = s / ZGROUP * ZGROUP;
            for (int r = i0s[s]; r < i0s[s] + nbs[s]; ++r) {
                cfar[r] = i0s[gs] + nbs[gs];
            }
        }
        ZENSURE(ws.cfar, sizeof(int) * (size_t)n);
        XTG_CHECK(hipMemcpyAsync(ws.cfar.p, cfar.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        int pr = prof_begin(st);
        hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.rowexp.as<int>(), (long)n, -1000); XTG_LAUNCH_CHECK("k_fill_int");
        hipLaunchKernelGGL(kz_rowexp, dim3((n + 255) / 256, (n + 63) / 64), dim3(256), 0, st, Sw, ldsw, Pw, ldpw, n, n,
                           ws.rowexp.as<int>(), (const int *)ws.cfar.as<int>()); XTG_LAUNCH_CHECK("kz_rowexp");
        const dim3 grid((n + 63) / 64, (unsigned)((qld + 15) / 16));
        hipLaunchKernelGGL(k_slice_digits, grid, dim3(256), 0, st, (const double *)Sw, ldsw, 0, n,
                           (const double *)nullptr, 0L, 0, 0, n, (int)qld, (const int *)ws.rowexp.as<int>(),
                           ws.Sx.as<signed char>(), qsl, kbs, 1, ws.Sy.as<signed char>()); XTG_LAUNCH_CHECK("k_slice_digits");
        hipLaunchKernelGGL(k_slice_digits, grid, dim3(256), 0, st, (const double *)Pw, ldpw, 0, n,
                           (const double *)nullptr, 0L, 0, 0, n, (int)qld, (const int *)ws.rowexp.as<int>(),
                           ws.Px.as<signed char>(), qsl, kbs, 1, ws.Py.as<signed char>()); XTG_LAUNCH_CHECK("k_slice_digits");
        prof_end(pr, st, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 16.0 * 3.0 * n * (double)n + 4.0 * SPLIT_DIGITS * qsl);
    }

    // Steps go bottom-up in groups of ZGROUP row blocks [g0, g1).  Within a group each step
    // updates only the group's rows above it ("near" update, K = 2 nb); after the group one
    // "far" update applies all of its blocks to the rows above g0 (K = 2 (g1 - g0)).  The
    // group's operand W holds rows [0, g1-g0) (S part: -a x) and [g1-g0, 2(g1-g0)) (P part: b x).
    int c_begin = m;
    static const bool lookahead = std::getenv("XTGEVC3_NO_LOOKAHEAD") == nullptr;
    if (lookahead) {
        XTG_CHECK(zfar_stream_init(ws));
    }
    bool f_pending = false; // a far update on ws.fs has not been waited for yet
    int last_par = 0;       // its group parity
    for (int gs = 0; gs < nsteps; gs += ZGROUP) {
        const int ge = std::min(nsteps, gs + ZGROUP);
        const int g1 = i0s[gs] + nbs[gs];
        const int g0 = i0s[ge - 1];
        const int wr = g1 - g0;
        // this group's copies of the operand, group scales and their digits / exponents
        const int par = (gs / ZGROUP) & 1;
        if (f_pending && last_par == par) {
            XTG_CHECK(hipStreamWaitEvent(st, ws.ev_fdone[last_par], 0));
            f_pending = false;
        }
        double2 *Wb = W + (size_t)par * wsz;
        double *sgb = ws.sig_grp.as<double>() + (size_t)par * m;
        int *wgexpb = split ? ws.wgexp.as<int>() + (size_t)par * m : nullptr;
        signed char *Wgb = split ? ws.Wg.as<signed char>() + (size_t)par * SPLIT_DIGITS * m * ZWG_K : nullptr;
        GroupInfo grp{};
        grp.wr = wr;
        {
            int off = 0;
            for (int jb = 0; jb < GROUP_MAX_BLOCKS; ++jb) {
                const int sj = ge - 1 - jb;
                grp.off[jb] = off;
                grp.step[jb] = sj >= gs ? sj : ge - 1;
                if (sj >= gs) {
                    off += nbs[sj];
                }
            }
            grp.off[GROUP_MAX_BLOCKS] = off;
        }
        bool first = true;
        for (int s = gs; s < ge; ++s) {
            const int i0 = i0s[s], nb = nbs[s];
            while (c_begin > 0 && desc[c_begin - 1].d >= i0) {
                --c_begin;
            }
            if (c_begin >= m) {
                continue;
            }
            const int N = m - c_begin;
            const int woff = i0 - g0;
            grp.last = s == ge - 1 ? 1 : 0;
            int pr = prof_begin(st);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(kz_solve_units<NB_Z>),
                               dim3((N + ZCOLS_PER_BLOCK - 1) / ZCOLS_PER_BLOCK), dim3(256), 0, st, Sw, ldsw, Pw,
                               ldpw, i0, nb, X, ldx, Wb, ldw, ws.desc.as<ZColDesc>(), c_begin, m, ws.sig.as<double>(),
                               ws.colexp.as<int>(), ws.blkexp.as<int>() + (long)s * ldx, ws.flag.as<int>(), woff,
                               wr + woff, sgb, s == gs ? 1 : 0, grp,
                               split ? SplitOut{ws.wbexp.as<int>() + (long)s * ldx, wgexpb,
                                                ws.xbexp.as<int>() + (long)s * ldx}
                                     : SplitOut{nullptr, nullptr, nullptr}); XTG_LAUNCH_CHECK("kz_solve_units");
            prof_end(pr, st, 0, 8.0 * nb * nb * (double)N, 16.0 * (4.0 * nb * N + 2.0 * nb * nb));
            first = false;
            // near update: the group's rows above this block
            const int M = i0 - g0;
            if (M > 0) {
                // few rows: 32-row tiles keep more CUs busy than 128-row tiles
                const bool small = ((N + ZTB_H - 1) / ZTB_H) * ((M + ZTA_H - 1) / ZTA_H) < 256;
                const int ta = small ? 32 : ZTA_H;
                dim3 grid(((N + ZTB_H - 1) / ZTB_H) * ((M + ta - 1) / ta));
                pr = prof_begin(st);
                auto kern = small ? kz_gemm<0, 32> : kz_gemm<0, ZTA_H>;
                hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, M, N, nb, nb, Sw + g0 + (long)i0 * ldsw, ldsw,
                                   Pw + g0 + (long)i0 * ldpw, ldpw, (const double2 *)(Wb + (long)woff * ldw + c_begin),
                                   ldw, X + (long)g0 * ldx + c_begin, ldx, (const int *)nullptr, 1,
                                   (const double *)(ws.sig.as<double>() + c_begin), (const int *)nullptr,
                                   (const double2 *)(Wb + (long)(wr + woff) * ldw + c_begin)); XTG_LAUNCH_CHECK("kz_gemm<0>");
                prof_end(pr, st, 1, 8.0 * M * N * (2.0 * nb), 16.0 * (2.0 * M * N + 2.0 * nb * M + 2.0 * nb * N));
            }
        }
        if (f_pending && !(first || g0 == 0)) {
            // the previous group's rows above this group are written again below
            XTG_CHECK(hipStreamWaitEvent(st, ws.ev_fdone[last_par], 0));
            f_pending = false;
        }
        if (first || g0 == 0) {
            continue;
        }
        // far update: rows above the group, all of its blocks at once.  With the lookahead the
        // rows of the next group run on st and the rows above it on ws.fs, overlapping the
        // next group's steps; the next far update waits for them.
        const int N = m - c_begin;
        const int gen = std::min(nsteps, ge + ZGROUP);
        const int rsplit = lookahead && ge < nsteps ? i0s[gen - 1] : 0;
        int klo = 0, kp1 = 0;
        signed char *wg = nullptr;
        const long wslice = (long)m * ZWG_K, wkb = (long)m * ZSPLIT_KSTEP;
        if (split) {
            // digits of the group operand, (re, im) expanded and shifted by off so the expanded K
            // range starts at the K step boundary klo <= 2 g0; Re and Im as two real products
            klo = (2 * g0) / ZSPLIT_KSTEP * ZSPLIT_KSTEP;
            const int off = 2 * g0 - klo;
            kp1 = zround_up(off + 2 * wr, ZSPLIT_KSTEP);
            wg = Wgb + (long)c_begin * ZSPLIT_KSTEP;
            int pr = prof_begin(st);
            hipLaunchKernelGGL(k_slice_digits, dim3((N + 63) / 64, (2 * kp1 + 15) / 16), dim3(256), 0, st,
                               (const double *)(Wb + c_begin), ldw, off / 2, wr,
                               (const double *)(Wb + (long)wr * ldw + c_begin), ldw, (kp1 + off) / 2, wr, N, 2 * kp1,
                               (const int *)(wgexpb + c_begin), wg, wslice, wkb, 3,
                               (signed char *)nullptr); XTG_LAUNCH_CHECK("k_slice_digits");
            prof_end(pr, st, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 16.0 * (2.0 * wr * N) + SPLIT_DIGITS * 2.0 * kp1 * N);
        }
        // the product for rows [ra, rb) on stream fst
        auto far = [&](int ra, int rb, hipStream_t fst) -> int {
            const int Mp = rb - ra;
            if (Mp <= 0) {
                return 0;
            }
            int pr = prof_begin(fst);
            if (split) {
                const long ko = (long)(klo / ZSPLIT_KSTEP) * kbs + (long)ra * ZSPLIT_KSTEP;
                for (int part = 0; part < 2; ++part) {
                    const signed char *ls = (part == 0 ? ws.Sx : ws.Sy).as<signed char>() + ko;
                    const signed char *lp = (part == 0 ? ws.Px : ws.Py).as<signed char>() + ko;
                    XTG_CHECK((hipError_t)launch_gemm_i8(fst, Mp, N, kp1, kp1, ls, lp, qsl, kbs, wg, wslice, wkb,
                                                         ws.rowexp.as<int>() + ra, wgexpb + c_begin,
                                                         (double *)(X + (long)ra * ldx + c_begin) + part, 2 * ldx, 2,
                                                         nullptr, 1, sgb + c_begin, nullptr));
                }
                prof_end(pr, fst, XTGEVC3_PROF_UPDATE_SPLIT, 8.0 * Mp * N * (2.0 * wr),
                         16.0 * (2.0 * Mp * N) + 2.0 * SPLIT_DIGITS * (2.0 * kp1 * Mp + 2.0 * kp1 * N));
            }
            else {
                dim3 grid(((N + ZTB_H - 1) / ZTB_H) * ((Mp + ZTA_H - 1) / ZTA_H));
                hipLaunchKernelGGL(HIP_KERNEL_NAME(kz_gemm<0, ZTA_H>), grid, dim3(256), 0, fst, Mp, N, wr, wr,
                                   Sw + ra + (long)g0 * ldsw, ldsw, Pw + ra + (long)g0 * ldpw, ldpw,
                                   (const double2 *)(Wb + c_begin), ldw, X + (long)ra * ldx + c_begin, ldx,
                                   (const int *)nullptr, 1, (const double *)(sgb + c_begin), (const int *)nullptr,
                                   (const double2 *)(Wb + (long)wr * ldw + c_begin)); XTG_LAUNCH_CHECK("kz_gemm<0>");
                prof_end(pr, fst, 1, 8.0 * Mp * N * (2.0 * wr), 16.0 * (2.0 * Mp * N + 2.0 * wr * Mp + 2.0 * wr * N));
            }
            return 0;
        };
        if (const int frc = far(rsplit, g0, st)) {
            return frc;
        }
        if (rsplit > 0) {
            XTG_CHECK(hipEventRecord(ws.ev_wready[par], st));
            XTG_CHECK(hipStreamWaitEvent(ws.fs, ws.ev_wready[par], 0));
            if (const int frc = far(0, rsplit, ws.fs)) {
                return frc;
            }
            XTG_CHECK(hipEventRecord(ws.ev_fdone[par], ws.fs));
            f_pending = true;
            last_par = par;
        }
    }
    if (f_pending) {
        XTG_CHECK(hipStreamWaitEvent(st, ws.ev_fdone[last_par], 0));
    }
    XTG_CHECK(hipGetLastError());
    {
        dim3 grid((m + 255) / 256, n);
        hipLaunchKernelGGL(kz_fixup_scale, grid, dim3(256), 0, st, X, ldx, n, m, ws.rowstep.as<int>(),
                           ws.blkexp.as<int>(), ldx, ws.colexp.as<int>(), ws.flag.as<int>()); XTG_LAUNCH_CHECK("kz_fixup_scale");
    }
    if (split) {
        hipLaunchKernelGGL(k_xexp_final, dim3((m + 255) / 256), dim3(256), 0, st, (const int *)ws.xbexp.as<int>(),
                           (const int *)ws.blkexp.as<int>(), ldx, (const int *)ws.colexp.as<int>(),
                           (const int *)ws.sstart.as<int>(), nsteps, m, ws.xexp.as<int>()); XTG_LAUNCH_CHECK("k_xexp_final");
    }
    if (back && zf != nullptr) {
        const hipEvent_t zready = zf->ready(left);
        if (zready == nullptr) {
            return -1000;
        }
        XTG_CHECK(hipStreamWaitEvent(st, zready, 0));
    }
    if (back && split) {
        // digits of X's columns as (re, -im) and (im, re), of Z's rows as (re, im); V = Re + i Im
        const double2 *Rp = left ? Z + (long)(n - 1) * ldz : Z;
        const long ldr = left ? -ldz : ldz;
        const int kb = zround_up(2 * n, ZSPLIT_KSTEP);
        ZENSURE(ws.Xx, (size_t)SPLIT_DIGITS * m * kb);
        ZENSURE(ws.Xy, (size_t)SPLIT_DIGITS * m * kb);
        ZENSURE(ws.Zq, (size_t)SPLIT_DIGITS * n * kb);
        ZENSURE(ws.zexp, sizeof(int) * (size_t)n);
        ZENSURE(ws.kmax2, sizeof(int) * (size_t)m);
        std::vector<int> kmax2(m);
        double ksum = 0.0;
        int kmx = 0;
        for (int q = 0; q < m; ++q) {
            kmax2[q] = 2 * kmaxv[q];
            ksum += kmaxv[q];
            kmx = std::max(kmx, kmaxv[q]);
        }
        XTG_CHECK(hipMemcpyAsync(ws.kmax2.p, kmax2.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        int pr = prof_begin(st);
        hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.zexp.as<int>(), (long)n, -1000); XTG_LAUNCH_CHECK("k_fill_int");
        hipLaunchKernelGGL(kz_rowexp, dim3((n + 255) / 256, (n + 63) / 64), dim3(256), 0, st, Rp, ldr,
                           (const double2 *)nullptr, 0L, n, n, ws.zexp.as<int>(), (const int *)nullptr); XTG_LAUNCH_CHECK("kz_rowexp");
        const dim3 gx((m + 63) / 64, (kb + 15) / 16), gz((n + 63) / 64, (kb + 15) / 16);
        hipLaunchKernelGGL(k_slice_digits, gx, dim3(256), 0, st, (const double *)X, ldx, 0, n, (const double *)nullptr,
                           0L, 0, 0, m, kb, (const int *)ws.xexp.as<int>(), ws.Xx.as<signed char>(), (long)m * kb,
                           (long)m * ZSPLIT_KSTEP, 1, ws.Xy.as<signed char>()); XTG_LAUNCH_CHECK("k_slice_digits");
        hipLaunchKernelGGL(k_slice_digits, gz, dim3(256), 0, st, (const double *)Rp, ldr, 0, n, (const double *)nullptr,
                           0L, 0, 0, n, kb, (const int *)ws.zexp.as<int>(), ws.Zq.as<signed char>(), (long)n * kb,
                           (long)n * ZSPLIT_KSTEP, 3, (signed char *)nullptr); XTG_LAUNCH_CHECK("k_slice_digits");
        prof_end(pr, st, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 16.0 * 2.0 * ((double)n * m + (double)n * n));
        pr = prof_begin(st);
        for (int part = 0; part < 2; ++part) {
            XTG_CHECK((hipError_t)launch_gemm_i8(st, m, n, kb, 0, (part == 0 ? ws.Xx : ws.Xy).as<signed char>(),
                                                 (part == 0 ? ws.Xx : ws.Xy).as<signed char>(), (long)m * kb,
                                                 (long)m * ZSPLIT_KSTEP, ws.Zq.as<signed char>(), (long)n * kb,
                                                 (long)n * ZSPLIT_KSTEP, ws.xexp.as<int>(), ws.zexp.as<int>(),
                                                 (double *)V + part, 2 * ldv, 2, ws.outcol.as<int>(), 0, nullptr,
                                                 ws.kmax2.as<int>()));
        }
        prof_end(pr, st, XTGEVC3_PROF_BACK_SPLIT, 8.0 * n * ksum,
                 16.0 * (double)n * m + 2.0 * SPLIT_DIGITS * (2.0 * (double)n * kmx + 2.0 * ksum));
    }
    else if (back && zf != nullptr && zf->done && zf->chunks > 1) {
        // pieces of consecutive working columns, each normalised and handed to the caller as
        // soon as it is queued; a piece's output columns are contiguous (output order, reversed
        // for the left side)
        const double2 *Rp = left ? Z + (long)(n - 1) * ldz : Z;
        const long ldr = left ? -ldz : ldz;
        const int C = std::min(zf->chunks, m);
        for (int i = 0; i < C; ++i) {
            const int p0 = (int)((long)m * i / C), p1 = (int)((long)m * (i + 1) / C);
            int c0 = n, c1 = 0;
            double ks = 0.0;
            int kx = 0;
            for (int q = p0; q < p1; ++q) {
                c0 = std::min(c0, outcol[q]);
                c1 = std::max(c1, outcol[q] + 1);
                ks += kmaxv[q];
                kx = std::max(kx, kmaxv[q]);
            }
            const dim3 grid(((n + ZTB_H - 1) / ZTB_H) * ((p1 - p0 + ZTA_H - 1) / ZTA_H));
            int pr = prof_begin(st);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(kz_gemm<1, ZTA_H>), grid, dim3(256), 0, st, p1 - p0, n, n, 0,
                               (const double2 *)X + p0, ldx, (const double2 *)nullptr, 0L, Rp, ldr, V, ldv,
                               ws.outcol.as<int>() + p0, 0, (const double *)nullptr, ws.kmax.as<int>() + p0,
                               (const double2 *)nullptr); XTG_LAUNCH_CHECK("kz_gemm<1>");
            prof_end(pr, st, 2, 8.0 * n * ks, 16.0 * ((double)n * kx + ks + (double)n * (p1 - p0)));
            hipLaunchKernelGGL(kz_normalize, dim3(c1 - c0), dim3(256), 0, st, V + (long)c0 * ldv, ldv, n); XTG_LAUNCH_CHECK("kz_normalize");
            if (float *o = static_cast<float *>(zf->out32[left ? 1 : 0])) {
                // (re, im) pairs as 2n scalars per column
                hipLaunchKernelGGL(k_round_f32, dim3(c1 - c0), dim3(256), 0, st, (const double *)(V + (long)c0 * ldv),
                                   2 * ldv, 2L * n, o + 2L * c0 * n, 2L * n); XTG_LAUNCH_CHECK("k_round_f32");
            }
            if (const int drc = zf->done(left, c0, c1, st)) {
                return drc;
            }
        }
        XTG_CHECK(hipGetLastError());
        return 0;
    }
    else if (back) {
        const double2 *Rp = left ? Z + (long)(n - 1) * ldz : Z;
        long ldr = left ? -ldz : ldz;
        dim3 grid(((n + ZTB_H - 1) / ZTB_H) * ((m + ZTA_H - 1) / ZTA_H));
        double ksum = 0.0;
        int kmx = 0;
        for (int q = 0; q < m; ++q) {
            ksum += kmaxv[q];
            kmx = std::max(kmx, kmaxv[q]);
        }
        int pr = prof_begin(st);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kz_gemm<1, ZTA_H>), grid, dim3(256), 0, st, m, n, n, 0, (const double2 *)X, ldx,
                           (const double2 *)nullptr, 0L, Rp, ldr, V, ldv, ws.outcol.as<int>(), 0,
                           (const double *)nullptr, ws.kmax.as<int>(), (const double2 *)nullptr); XTG_LAUNCH_CHECK("kz_gemm<1>");
        prof_end(pr, st, 2, 8.0 * n * ksum, 16.0 * ((double)n * kmx + ksum + (double)n * m));
    }
    else {
        dim3 grid((m + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(kz_copy_out, grid, dim3(32, 8), 0, st, (const double2 *)X, ldx, n, m, V, ldv,
                           ws.outcol.as<int>(), left ? 1 : 0); XTG_LAUNCH_CHECK("kz_copy_out");
    }
    hipLaunchKernelGGL(kz_normalize, dim3(m), dim3(256), 0, st, V, ldv, n); XTG_LAUNCH_CHECK("kz_normalize");
    XTG_CHECK(hipGetLastError());
    return 0;
}

} // namespace

int complex_device(char side, char howmny, const int *select, int n, const std::complex<double> *S, long lds,
                   const std::complex<double> *P, long ldp, const std::complex<double> *al,
                   const std::complex<double> *be, std::complex<double> *VL, long ldvl, std::complex<double> *VR,
                   long ldvr, hipStream_t st, int prec, const ZFeed *zf, double safemin)
{
    if (zf != nullptr && prec == XTGEVC3_PREC_SPLIT_I8) {
        set_error("Z delivered beside the sweep is not supported in the split form");
        return -1000;
    }
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_zmutex);
    ZWork &ws = g_zws[dev & 63];
    ZWsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    bool right = side == 'R' || side == 'B';
    bool left = side == 'L' || side == 'B';
    bool back = howmny == 'B';
    std::vector<Unit> units;
    int m = complex_units(howmny, select, n, units);
    if (m == 0 || (!(right && VR) && !(left && VL))) {
        return 0;
    }
    const double2 *dS = reinterpret_cast<const double2 *>(S);
    const double2 *dP = reinterpret_cast<const double2 *>(P);
    ZENSURE(ws.norms, sizeof(double) * 2 * (size_t)n);
    hipLaunchKernelGGL(kz_colnorms, dim3(n), dim3(256), 0, st, dS, lds, dP, ldp, n, ws.norms.as<double>()); XTG_LAUNCH_CHECK("kz_colnorms");
    std::vector<double> nrm(2 * (size_t)n);
    XTG_CHECK(hipMemcpyAsync(nrm.data(), ws.norms.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
    XTG_CHECK(hipStreamSynchronize(st));
    double an = 0.0, bn = 0.0;
    for (int c = 0; c < n; ++c) {
        an = std::max(an, nrm[c]);
        bn = std::max(bn, nrm[n + c]);
    }
    double ascale = 1.0 / std::max(an, safemin);
    double bscale = 1.0 / std::max(bn, safemin);
    if (right && VR) {
        const double2 *Zp = nullptr;
        long zld = n;
        if (back && zf != nullptr) {
            Zp = static_cast<const double2 *>(zf->zr);
        }
        else if (back && prec == XTGEVC3_PREC_SPLIT_I8) {
            // the split form reads Z only through its digits, made before the output is written
            Zp = reinterpret_cast<const double2 *>(VR);
            zld = ldvr;
        }
        else if (back) {
            ZENSURE(ws.Zt, sizeof(double2) * (size_t)n * n);
            XTG_CHECK(hipMemcpy2DAsync(ws.Zt.p, sizeof(double2) * n, VR, sizeof(double2) * ldvr, sizeof(double2) * n,
                                       n, hipMemcpyDeviceToDevice, st));
            Zp = ws.Zt.as<double2>();
        }
        int rc = run_cplx_side(ws, n, dS, lds, dP, ldp, al, be, ascale, bscale, false, back, Zp, zld,
                               reinterpret_cast<double2 *>(VR), ldvr, units, m, st, prec, zf, safemin);
        if (rc) {
            return rc;
        }
    }
    if (left && VL) {
        ZENSURE(ws.Sw, sizeof(double2) * (size_t)n * n);
        ZENSURE(ws.Pw, sizeof(double2) * (size_t)n * n);
        dim3 grid((n + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(kz_reverse_conjtranspose, grid, dim3(32, 8), 0, st, dS, lds, n, ws.Sw.as<double2>(), (long)n); XTG_LAUNCH_CHECK("kz_reverse_conjtranspose");
        hipLaunchKernelGGL(kz_reverse_conjtranspose, grid, dim3(32, 8), 0, st, dP, ldp, n, ws.Pw.as<double2>(), (long)n); XTG_LAUNCH_CHECK("kz_reverse_conjtranspose");
        const double2 *Zp = nullptr;
        long zld = n;
        if (back && zf != nullptr) {
            Zp = static_cast<const double2 *>(zf->zl);
        }
        else if (back && prec == XTGEVC3_PREC_SPLIT_I8) {
            // the split form reads Z only through its digits, made before the output is written
            Zp = reinterpret_cast<const double2 *>(VL);
            zld = ldvl;
        }
        else if (back) {
            ZENSURE(ws.Zt, sizeof(double2) * (size_t)n * n);
            XTG_CHECK(hipMemcpy2DAsync(ws.Zt.p, sizeof(double2) * n, VL, sizeof(double2) * ldvl, sizeof(double2) * n,
                                       n, hipMemcpyDeviceToDevice, st));
            Zp = ws.Zt.as<double2>();
        }
        int rc = run_cplx_side(ws, n, ws.Sw.as<double2>(), n, ws.Pw.as<double2>(), n, al, be, ascale, bscale, true,
                               back, Zp, zld, reinterpret_cast<double2 *>(VL), ldvl, units, m, st, prec, zf, safemin);
        if (rc) {
            return rc;
        }
    }
    return 0;
}

// Eigenvalue condition numbers of the complex pencil (LAPACK ztgsna, JOB = 'E'): for the
// columns howmny / select describes (dtgevc3's rule: one column per eigenvalue), with left and
// right vectors VL, VR (device, the basis of S and P): s[c] = |(u^H S v, u^H P v)| / (|u| |v|).
// YA = S V and YB = P V on the complex back-transform GEMM, then one wave per column.
int complex_tgsna_device(char howmny, const int *select, int n, const std::complex<double> *S, long lds,
                         const std::complex<double> *P, long ldp, const std::complex<double> *VL, long ldvl,
                         const std::complex<double> *VR, long ldvr, int m, double *s, hipStream_t st)
{
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_zmutex);
    ZWork &ws = g_zws[dev & 63];
    ZWsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    std::vector<Unit> units;
    const int mc = complex_units(howmny, select, n, units);
    if (mc != m) {
        set_error("m = " + std::to_string(m) + " does not match the " + std::to_string(mc) +
                  " columns the selection describes");
        return -1;
    }
    if (m == 0) {
        return 0;
    }
    const long ldt = ((long)m + 7) / 8 * 8;
    ZENSURE(ws.Vt, sizeof(double2) * (size_t)n * ldt);
    ZENSURE(ws.YA, sizeof(double2) * (size_t)n * m);
    ZENSURE(ws.YB, sizeof(double2) * (size_t)n * m);
    hipLaunchKernelGGL(kz_to_rowmajor, dim3((m + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, st, (const double2 *)VR,
                       ldvr, n, m, ws.Vt.as<double2>(), ldt); XTG_LAUNCH_CHECK("kz_to_rowmajor");
    // D[a][b] = sum_k V(k, a) M(b, k): column a of Y = M v_a (TAG 1 form, full K, no column map)
    const dim3 grid(((n + ZTB_H - 1) / ZTB_H) * ((m + ZTA_H - 1) / ZTA_H));
    for (int part = 0; part < 2; ++part) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kz_gemm<1, ZTA_H>), grid, dim3(256), 0, st, m, n, n, 0,
                           (const double2 *)ws.Vt.as<double2>(), ldt, (const double2 *)nullptr, 0L,
                           (const double2 *)(part == 0 ? S : P), part == 0 ? lds : ldp,
                           part == 0 ? ws.YA.as<double2>() : ws.YB.as<double2>(), (long)n, (const int *)nullptr, 0,
                           (const double *)nullptr, (const int *)nullptr, (const double2 *)nullptr); XTG_LAUNCH_CHECK("kz_gemm<1>");
    }
    hipLaunchKernelGGL(kz_tgsna, dim3(m), dim3(64), 0, st, (const double2 *)VL, ldvl, (const double2 *)VR, ldvr,
                       (const double2 *)ws.YA.as<double2>(), (const double2 *)ws.YB.as<double2>(), (long)n, n, s); XTG_LAUNCH_CHECK("kz_tgsna");
    XTG_CHECK(hipGetLastError());
    return 0;
}

} // namespace xtg

