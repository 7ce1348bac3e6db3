// Host driver of the real xTGEVC3 path: argument semantics of the reference entry point,
// the column scheduler, and the launch sequence of the gfx950 kernels.
//
// Reference behaviour restated here (src/dtgevc3.cpp):
//   side/howmny decoding and info = -2            :1173-1182
//   n == 0 early exit                              :1184-1190
//   selection count (pairs count twice), *m        :1192-1211
//   mm < m -> info = -16                           :1213-1216
//   lwork query / shrink / info = -19              :1219-1242
//   global scales ascale, bscale                   :1252-1269
//   shift coefficients acoeff, bcoeff              :647-655, :716-725, :1372-1381
//   pair detection alphai[k] != 0 && k+1 < n       :1198, :1204-1206
//   output columns packed in eigenvalue order      :1282, :1326, :1477, :1658
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xtgevc3.h"
#include "common.h"
#include "driver.h"

namespace xtg {

template <int MAXB>
__global__ void k_solve_real(const double2 *, const double2 *, int, int, int, int, unsigned long long, double *, long,
                             double *, long, const UnitDesc *, int, int, int, int, double *, int *, int *,
                             int *, int, int, double *, int, GroupInfo, MixedExp, SplitOut);
__global__ void k_pack_steps(const double *, long, const double *, long, const StepPack *, double2 *);
__global__ void k_rowexp(const double *, long, const double *, long, int, int, int *, const int *);
__global__ void k_tgsna_part(const double *, const double *, const double *, const double *, long, int, int,
                             const int *, const int *, int, double *);
__global__ void k_tgsna_fin(const double *, int, int, const int *, const int *, double *);
__global__ void k_slice_digits(const double *, long, int, int, const double *, long, int, int, int, int, const int *,
                               signed char *, long, long, int, signed char *);
__global__ void k_xexp_final(const int *, const int *, long, const int *, const int *, int, int, int *);
__global__ void k_slice_xfix(const double *, long, int, int, int, const int *, const int *, long, const int *,
                             const int *, signed char *, long, long);
int launch_gemm_i8(hipStream_t st, int M, int N, int K1, int K2, const signed char *L1, const signed char *L2,
                   long lslice, long kbl, const signed char *R, long rslice, long kbr, const int *EL, const int *ER,
                   double *C, long ldc, int cst, const int *outcol, int mode, const double *sig, const int *kmax);
__global__ void k_xfin(const int *, const int *, int, int *);
__global__ void k_fill_int(int *, long, int);
constexpr int SOLVE_UNITS = 4;   // eigenvalue units per wave of the solve kernel (16 lanes per unit)
constexpr int SOLVE_WAVES_H = 4; // waves per solve workgroup (they share the staged block)
int launch_gemm_f64(int tag, hipStream_t st, int M, int N, int K1, int K2, const double *L, long ldl, const double *L2,
                     long ldl2, const double *R, long ldr, double *C, long ldc, const int *outcol, int mode,
                     const double *sig, const int *kmax, const double *R2, int prec, const int *sa,
                     const int *sb, int batch = 1, long bsr = 0, long bsc = 0, const int *afl = nullptr,
                     int aoff = 0);
constexpr int GROUP_STEPS = 4; // row blocks per far update (K = 2 * ~256)
__global__ void k_fixup_scale(double *, long, int, int, const int *, const int *, long, const int *, const int *);
__global__ void k_fixup_steps(double *, long, int, const int *, const int *, long, const int *, const int *);
__global__ void k_copy_out(const double *, long, int, int, double *, long, const int *, int);
__global__ void k_normalize(double *, long, int, const int *, const int *);
__global__ void k_round_f32(const double *, long, long, float *, long);
__global__ void k_reverse_transpose(const double *, long, int, double *, long);
__global__ void k_colnorms(const double *, long, const double *, long, int, int, double *);
template <int TQ>
__global__ void k_bwd_rhs(const double *, long, const double *, long, int, int, int, const int *, double *, long, int);
__global__ void k_bwd_permute(const double *, long, int, int, int, int, int, double *, long);
__global__ void k_bwd_gcorr(const double *, long, const double *, long, int, int, double *, long);
__global__ void k_bwd_trunc(double *, long, int, int, const int *);
__global__ void k_bwd_lambda_t(const double *, long, int, int, int, const int *, const int *, const int *, double *,
                               long, int *, int *);
__global__ void k_to_rowmajor(const double *, long, int, int, double *, long);
__global__ void k_bwd_diag(const double *, long, const double *, long, int, int, int, const int *, const double *,
                           const int *, const int *, double *);
__global__ void k_bwd_diag_add(const double *, int, int, int, const int *, double *, long, double *, long);
__global__ void k_subdiag(const double *, long, int, double *);
__global__ void k_bwd_outer(int, int, const double *, long, const double *, long, const double *, const double *, double *,
                            long, double *, long, const int *, const int *, const int *, long, long, long, long);
__global__ void k_bwd_zero_tiles(int, double *, long, double *, long, const int *, long, long);
__global__ void k_zero(double *, long);
__global__ void k_zero_int(int *, long);

// profiling kind of an update GEMM launch by its tile configuration
static int update_kind(int cfg)
{
    switch (cfg) {
    case GEMM_128x128:
        return XTGEVC3_PROF_UPDATE_128x128;
    case GEMM_128x64:
        return XTGEVC3_PROF_UPDATE_128x64;
    case GEMM_64x128:
        return XTGEVC3_PROF_UPDATE_64x128;
    default:
        return XTGEVC3_PROF_UPDATE;
    }
}

constexpr int MAXB_REAL = 65;  // nominal block 64, +1 when a boundary would split a 2x2 block
constexpr int NB_REAL = 64;
// split-int8 form: K step of the digit products (digits are k-blocked by it); K extent of the
// far update's group operand (two segments, each the group's rows shifted to the K step
// boundary below the group and rounded up: at most 320)
constexpr int SPLIT_KSTEP = 32;
constexpr long WG_K = 2 * 320;
static_assert(GROUP_STEPS * MAXB_REAL + SPLIT_KSTEP - 1 <= 320, "split operand extent");
static inline int round_up(int v, int q) { return (v + q - 1) / q * q; }
// row pitch of the pencil digits: the far update reads up to SPLIT_KSTEP - 1 columns past n
static inline long split_qld(int n) { return round_up(n, SPLIT_KSTEP) + SPLIT_KSTEP; }

// ---------------------------------------------------------------------------------------
// argument semantics shared by every real entry point
// ---------------------------------------------------------------------------------------
static bool pair_start(const double *ai, int k, int n) { return ai[k] != 0.0 && k + 1 < n; }

int real_units(char howmny, const int *select, int n, const double *ai, std::vector<Unit> &units)
{
    bool do_all = howmny == 'A' || howmny == 'B';
    bool do_sel = howmny == 'S';
    units.clear();
    int m = 0;
    for (int k = 0; k < n;) {
        bool pr = pair_start(ai, k, n);
        bool sel = do_all;
        if (do_sel && select != nullptr) {
            sel = select[k] != 0 || (pr && select[k + 1] != 0);
        }
        int w = pr ? 2 : 1;
        if (sel) {
            units.push_back(Unit{k, w, m});
            m += w;
        }
        k += w;
    }
    return m;
}

int reference_lwork(int n, int b) { return 2 * n * (b + 1) + 4 * (b + 1) * (b + 1) + 2 * (b + 1); }

// Returns true when the caller should go on computing; sets *info / *m / work[0] exactly as
// the reference does before its first floating-point operation.
template <class T>
bool check_args(char howmny, int n, int num_sel, int mm, int *m, T *work, int lwork, int *info, bool have_work)
{
    bool do_all = howmny == 'A' || howmny == 'B';
    bool do_sel = howmny == 'S';
    if (!do_all && !do_sel) {
        *info = -2;
        return false;
    }
    if (n == 0) {
        *info = 0;
        if (m != nullptr) {
            *m = 0;
        }
        return false;
    }
    if (m != nullptr) {
        *m = num_sel;
    }
    if (mm < num_sel) {
        *info = -16;
        return false;
    }
    if (have_work) {
        int req = reference_lwork(n, 32);
        if (lwork != -1 && lwork < req) {
            for (int b = 31; b >= 1; --b) {
                req = reference_lwork(n, b);
                if (lwork >= req) {
                    break;
                }
            }
        }
        if (lwork == -1) {
            work[0] = T(reference_lwork(n, 32));
            *info = 0;
            return false;
        }
        if (lwork < req) {
            *info = -19;
            return false;
        }
    }
    *info = 0;
    return true;
}
template bool check_args<double>(char, int, int, int, int *, double *, int, int *, bool);
template bool check_args<float>(char, int, int, int, int *, float *, int, int *, bool);
template bool check_args<std::complex<double>>(char, int, int, int, int *, std::complex<double> *, int, int *, bool);
template bool check_args<std::complex<float>>(char, int, int, int, int *, std::complex<float> *, int, int *, bool);

// Shift coefficients for one eigenvalue (dtgevc3.cpp:716-725; with aI = 0 this is :647-655).
// safemin is the smallest normal number of the interface precision (FLT_MIN for stgevc3,
// stgevc3.cpp:619-624, :688-693, :1219), which also decides that an eigenvalue is infinite.
static void shift_real(double aR, double aI, double bv, double ascale, double bscale, double &a, double &bR,
                       double &bI, bool pairform, double safemin = std::numeric_limits<double>::min())
{
    double num = pairform ? std::max(std::max(std::fabs(aR) * ascale + std::fabs(aI) * ascale, std::fabs(bv) * bscale), safemin)
                          : std::max(std::max(std::fabs(aR) * ascale, std::fabs(bv) * bscale), safemin);
    double t = 1.0 / num;
    a = (t * bv * bscale) * ascale;
    bR = (t * aR * ascale) * bscale;
    bI = pairform ? (t * aI * ascale) * bscale : 0.0;
    if (std::fabs(bv) <= safemin && (std::fabs(aR) + std::fabs(aI)) > safemin) {
        a = 0.0;
        bR = 1.0;
        bI = 0.0;
    }
}

// Block partition of the working pencil from the bottom, never splitting a 2x2 block
// (the same snapping rule as dtgevc3.cpp:32-47, with a nominal height of 64 rows).
static void partition(int n, const std::vector<double> &subw, std::vector<int> &i0s, std::vector<int> &nbs)
{
    i0s.clear();
    nbs.clear();
    int cur = n;
    while (cur > 0) {
        int i = std::max(0, cur - NB_REAL);
        if (i > 0 && subw[i] != 0.0) {
            --i;
        }
        i0s.push_back(i);
        nbs.push_back(cur - i);
        cur = i;
    }
}

// First column of the far-update region of each row of the working pencil: the bottom edge g1 of
// the row's group of GROUP_STEPS blocks.  A row only meets far updates of groups below its own,
// whose K columns lie in [g1, n).
static void far_columns(int n, const std::vector<double> &subw, std::vector<int> &cfar)
{
    std::vector<int> i0s, nbs;
    partition(n, subw, i0s, nbs);
    cfar.assign(n, n);
    for (int s = 0; s < (int)i0s.size(); ++s) {
        const int gs = s / GROUP_STEPS * GROUP_STEPS;
        const int g1 = i0s[gs] + nbs[gs];
        for (int r = i0s[s]; r < i0s[s] + nbs[s]; ++r) {
            cfar[r] = g1;
        }
    }
}

// Units of the real solve: one per real eigenvalue column, one per complex pair (p, p+1).
static void make_units(const std::vector<ColDesc> &desc, std::vector<UnitDesc> &units, std::vector<int> &col2unit)
{
    units.clear();
    const int m = (int)desc.size();
    col2unit.assign(m, 0);
    for (int p = 0; p < m;) {
        const ColDesc &c = desc[p];
        UnitDesc u{};
        u.a = c.a;
        u.bR = c.bR;
        u.bI = c.bI;
        u.col = p;
        u.w = ((c.flags & 2) && c.partner == p + 1) ? 2 : 1;
        u.d = c.d;
        u.lo = c.lo;
        u.init = (c.flags >> 2) & 3;
        for (int k = 0; k < u.w; ++k) {
            col2unit[p + k] = (int)units.size();
        }
        units.push_back(u);
        p += u.w;
    }
}

// Waves of the solve kernel: SOLVE_UNITS unit indices each (-1: empty).  Units whose own block
// is the same row block form one group (the active units of a step are a union of groups);
// within a group the real units and the pairs go to separate waves, so that waves of real
// eigenvalues run in real arithmetic.  tmin/tmax: smallest / largest unit of each tile.
static void make_tiles(const std::vector<UnitDesc> &units, const std::vector<int> &rowstep, int n,
                       std::vector<int4> &tiles, std::vector<int> &tmin, std::vector<int> &tmax)
{
    static_assert(SOLVE_UNITS == 4, "tiles hold four units");
    tiles.clear();
    tmin.clear();
    tmax.clear();
    const int nu = (int)units.size();
    auto key = [&](int u) { return units[u].d < n ? rowstep[units[u].d] : -1; };
    auto emit = [&](const std::vector<int> &list) {
        for (size_t s = 0; s < list.size(); s += SOLVE_UNITS) {
            int v[SOLVE_UNITS] = {-1, -1, -1, -1};
            int lo = INT_MAX, hi = -1;
            for (int k = 0; k < SOLVE_UNITS && s + k < list.size(); ++k) {
                v[k] = list[s + k];
                lo = std::min(lo, v[k]);
                hi = std::max(hi, v[k]);
            }
            tiles.push_back(make_int4(v[0], v[1], v[2], v[3]));
            tmin.push_back(lo);
            tmax.push_back(hi);
        }
    };
    std::vector<int> reals, pairs;
    for (int u0 = 0; u0 < nu;) {
        int u1 = u0;
        while (u1 < nu && key(u1) == key(u0)) {
            ++u1;
        }
        reals.clear();
        pairs.clear();
        for (int u = u0; u < u1; ++u) {
            (units[u].w == 2 ? pairs : reals).push_back(u);
        }
        emit(reals);
        emit(pairs);
        u0 = u1;
    }
}

// ---------------------------------------------------------------------------------------
// workspace: grow-only device buffers kept per device for the life of the process
// ---------------------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool fixed = false; // a slot of the allocation-free workspace: never reallocated
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) {
            return hipSuccess;
        }
        if (fixed) {
            return hipErrorOutOfMemory;
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

// Pinned host memory of a workspace (grows; hipHostMalloc, never hipHostRegister).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) {
            return hipSuccess;
        }
        if (p) {
            hipError_t e = hipHostFree(p);
            if (e != hipSuccess) {
                return e;
            }
            p = nullptr;
            cap = 0;
        }
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) {
            cap = bytes;
        }
        return e;
    }
};

struct RealWork {
    HostBuf hnorms;                // column norms and subdiagonal read back to the host
    hipEvent_t ev_norms = nullptr; // their copy
    DevBuf X, W, sig, colexp, blkexp, flag, rowexp, zexp, wgrp, xrel, xfin, desc, outcol, kmax, rowstep, vcol, vw, Sw, Pw, Zt, norms;
    DevBuf kcol, coef, Vt, Y, LamS, bVp, bnfk, btiles, sig_grp, bkpre, bovf, sfac, srow;
    // backward of left / back-transformed vectors: working pencil, permuted operands, products
    DevBuf bfS, bfP, bfV, bfG, bfGc, bfGt, bfVt, bfZt, bfW, bfGh, bfdS, bfdP, bfk;
    DevBuf Ut, YB, part, tcol, tw; // eigenvalue condition numbers
    DevBuf pk, pksteps;            // packed block and panel images of the sweep's steps
    DevBuf colo;                   // adjoint sweep: first row of each column's system
    DevBuf bdg;                    // backward: the diagonal (alpha, beta) terms of each working column
    DevBuf Sq, Pq, Wg, wgexp, wbexp, xbexp, sstart, Xq, Zq, xexp, cfar; // split-int8 digits and exponents
    // split-int8 form: side stream (default priority) for the digit splitting of the pencil
    // and of Z, overlapped with the sweep; events order it against the caller's stream
    hipStream_t side = nullptr;
    hipEvent_t ev_start = nullptr, ev_pencil = nullptr, ev_z = nullptr, ev_y = nullptr, ev_rhs = nullptr;
    // far-update lookahead: the rows above the next group run on `fs`
    hipStream_t fs = nullptr;
    hipEvent_t ev_wready[2] = {nullptr, nullptr}, ev_fdone[2] = {nullptr, nullptr};
    hipEvent_t ev_zsrc = nullptr, ev_zcopy = nullptr; // copy of Z on `fs`: its start and end
    hipEvent_t ev_last = nullptr; // end of the previous call's work on this workspace
    // allocation-free workspace (buffers are fixed slots of a static device array): every
    // launch and copy stays on the caller's stream, no stream or event is created
    bool fixed = false;
};

// Orders a call after the previous call's work on the device's workspace, whatever stream
// that ran on (calls on one device then share it safely), and records the end of this call
// when it goes out of scope.
struct WsOrder {
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
    ~WsOrder()
    {
        if (on && *ev != nullptr) {
            (void)hipEventRecord(*ev, st);
        }
    }
};

static std::mutex g_ws_mutex;
static RealWork g_ws[64];

static thread_local std::string g_err;
void set_error(const std::string &s) { g_err = s; }
const char *last_error() { return g_err.c_str(); }

#define ENSURE(buf, bytes) XTG_CHECK((buf).ensure(bytes))
// a nonzero return code of a call is returned at once
#define XTG_CHECK_RC(x)                                                                            \
    do {                                                                                           \
        const int rc_ = (x);                                                                       \
        if (rc_ != 0) {                                                                            \
            return rc_;                                                                            \
        }                                                                                          \
    } while (0)

// ---------------------------------------------------------------------------------------
// One side of the real problem on the device.
//   Sw/Pw: working pencil (right: S, P; left: J S^T J, J P^T J), leading dims ldsw/ldpw
//   units: selected eigenvalue units (ascending eigenvalue index), m columns in total
//   left: working coordinates are reversed (row r <-> eigen row n-1-r)
//   Vout: output (device, column-major, ldv); for back-transform Z (device, ldz) is read
// ---------------------------------------------------------------------------------------
struct RealSideArgs {
    int n;
    const double *Sw;
    long ldsw;
    const double *Pw;
    long ldpw;
    const double *ar, *ai, *be; // host
    double ascale, bscale;
    bool left;
    bool back;
    const double *Z; // back-transform matrix (Z or Q), device
    long ldz;
    double *V; // output, device
    long ldv;
    const std::vector<double> *subw; // working subdiagonal, host
    int prec;                        // 0: FP64, 1: FP32-product updates and back-transform
    const ZFeed *zf;                 // Z arriving on another stream (null: Z is ready)
    double safemin;                  // smallest normal number of the interface precision
    bool x_ready;                    // X is allocated for this side and zeroed on the stream
};

// ---------------------------------------------------------------------------------------
// The right-looking sweep over block rows, bottom to top, on the working pencil.
//   desc: packed working columns.  Forward (eigenvector) columns have d ascending and are
//   active from the step holding their own diagonal block upward (a suffix of the list);
//   adjoint columns (d >= n, lo ascending) are active while their row range [lo, n) meets
//   the block (a prefix).  X (row-major, ldx) holds the right-hand sides on entry (zero for
//   eigenvectors) and the scaled-consistent solutions on exit; colexp[p] then holds the
//   power-of-two exponent e with X[:, p] = 2^e * (true solution).
// ---------------------------------------------------------------------------------------
static hipError_t far_stream_init(RealWork &ws);

int sweep_real(RealWork &ws, int n, const double *Sw, long ldsw, const double *Pw, long ldpw,
               const std::vector<double> &subw, const std::vector<ColDesc> &desc, bool adjoint, long ldx,
               hipStream_t st, int prec, hipEvent_t pencil_ready, bool fixup = true)
{
    const int m = (int)desc.size();
    std::vector<UnitDesc> units;
    std::vector<int> col2unit;
    make_units(desc, units, col2unit);
    std::vector<int> i0s, nbs;
    partition(n, subw, i0s, nbs);
    const int nsteps = (int)i0s.size();
    std::vector<int> rowstep(n);
    for (int s = 0; s < nsteps; ++s) {
        for (int r = i0s[s]; r < i0s[s] + nbs[s]; ++r) {
            rowstep[r] = s;
        }
    }
    std::vector<int4> tiles;
    std::vector<int> tmin, tmax;
    make_tiles(units, rowstep, n, tiles, tmin, tmax);
    const long ldw = ldx;
    const int wmax = GROUP_STEPS * MAXB_REAL;
    // group operands, group scales (and their digits / exponents below) alternate between two
    // copies by group parity: the lookahead far update of one group reads its copy while the
    // next group's steps write the other
    const size_t wsz = (size_t)(2 * wmax) * ldw;
    ENSURE(ws.W, sizeof(double) * 2 * wsz);
    ENSURE(ws.sig, sizeof(double) * (size_t)m);
    ENSURE(ws.sig_grp, sizeof(double) * 2 * (size_t)m);
    ENSURE(ws.colexp, sizeof(int) * (size_t)m);
    ENSURE(ws.blkexp, sizeof(int) * (size_t)nsteps * ldx);
    ENSURE(ws.flag, sizeof(int) * 4);

    // the units in tile order (four per tile, idx = unit index, -1 and a copy of the tile's
    // first unit for an empty slot): a solve wave reads its descriptors in one load
    std::vector<UnitDesc> tdesc(4 * tiles.size());
    for (size_t t = 0; t < tiles.size(); ++t) {
        const int ut[4] = {tiles[t].x, tiles[t].y, tiles[t].z, tiles[t].w};
        for (int q = 0; q < 4; ++q) {
            tdesc[4 * t + q] = units[ut[q] >= 0 ? ut[q] : ut[0]];
            tdesc[4 * t + q].idx = ut[q];
        }
    }
    ENSURE(ws.desc, sizeof(UnitDesc) * tdesc.size());
    ENSURE(ws.rowstep, sizeof(int) * (size_t)n);
    XTG_CHECK(hipMemcpyAsync(ws.desc.p, tdesc.data(), sizeof(UnitDesc) * tdesc.size(), hipMemcpyHostToDevice, st));
    if (adjoint) {
        // first row of each column's system (ascending): the far updates skip the tiles above it
        std::vector<int> lo(m);
        for (int c = 0; c < m; ++c) {
            lo[c] = desc[c].lo;
        }
        ENSURE(ws.colo, sizeof(int) * (size_t)m);
        XTG_CHECK(hipMemcpyAsync(ws.colo.p, lo.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    }
    MixedExp mx{nullptr, nullptr, 0};
    const bool split = prec == XTGEVC3_PREC_SPLIT_I8;
    long qld = 0, qsl = 0;
    if (split) {
        // pencil digits (ws.Sq, ws.Pq, ws.rowexp) come from the side stream (pencil_ready);
        // the sweep records block exponents of the operands and of X
        qld = split_qld(n);
        qsl = (long)n * qld;
        ENSURE(ws.Wg, 2 * (size_t)SPLIT_DIGITS * m * WG_K);
        ENSURE(ws.wgexp, sizeof(int) * 2 * (size_t)m);
        ENSURE(ws.wbexp, sizeof(int) * (size_t)nsteps * ldx);
        ENSURE(ws.xbexp, sizeof(int) * (size_t)nsteps * ldx);
        ENSURE(ws.sstart, sizeof(int) * (size_t)m);
        ENSURE(ws.xexp, sizeof(int) * (size_t)m);
        std::vector<int> sstart(m);
        for (int p = 0; p < m; ++p) {
            sstart[p] = desc[p].d < n ? rowstep[desc[p].d] : 0;
        }
        XTG_CHECK(hipMemcpyAsync(ws.sstart.p, sstart.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    }
    if (prec == 1) {
        // mixed form: row exponents of the working pencil, per-column operand/solution bounds
        ENSURE(ws.rowexp, sizeof(int) * (size_t)n);
        ENSURE(ws.wgrp, sizeof(int) * 2 * (size_t)m);
        ENSURE(ws.xrel, sizeof(int) * (size_t)m);
        hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.rowexp.as<int>(), (long)n, -1000); XTG_LAUNCH_CHECK("k_fill_int");
        hipLaunchKernelGGL(k_rowexp, dim3((n + 255) / 256, (n + 63) / 64), dim3(256), 0, st, Sw, ldsw, Pw, ldpw, n, n,
                           ws.rowexp.as<int>(), (const int *)nullptr); XTG_LAUNCH_CHECK("k_rowexp");
        hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.xrel.as<int>(), (long)m, -1000000); XTG_LAUNCH_CHECK("k_fill_int");
        mx = MixedExp{ws.wgrp.as<int>(), ws.xrel.as<int>(), 1};
    }
    XTG_CHECK(hipMemcpyAsync(ws.rowstep.p, rowstep.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    // operand images of every step: the diagonal block and the near update's panel (the
    // group's rows above the block), packed in one launch
    std::vector<StepPack> spk(nsteps);
    long pko = 0;
    int pk_items = PK_BLK;
    for (int s = 0; s < nsteps; ++s) {
        const int ge = std::min(nsteps, s / GROUP_STEPS * GROUP_STEPS + GROUP_STEPS);
        StepPack &q = spk[s];
        q.i0 = i0s[s];
        q.nb = nbs[s];
        q.g0 = i0s[ge - 1];
        q.nrt = (q.i0 - q.g0 + 15) / 16;
        q.boff = pko;
        pko += PK_BLK;
        q.poff = pko;
        pko += (long)q.nrt * ((q.nb + 3) / 4) * 64;
        pk_items = std::max(pk_items, PK_BLK + q.nrt * ((q.nb + 3) / 4) * 128);
    }
    ENSURE(ws.pk, sizeof(double2) * (size_t)pko);
    ENSURE(ws.pksteps, sizeof(StepPack) * (size_t)nsteps);
    XTG_CHECK(hipMemcpyAsync(ws.pksteps.p, spk.data(), sizeof(StepPack) * nsteps, hipMemcpyHostToDevice, st));
    const double2 *pk = ws.pk.as<double2>();
    hipLaunchKernelGGL(k_pack_steps, dim3((pk_items + 255) / 256, nsteps), dim3(256), 0, st, Sw, ldsw, Pw, ldpw,
                       (const StepPack *)ws.pksteps.as<StepPack>(), ws.pk.as<double2>()); XTG_LAUNCH_CHECK("k_pack_steps");
    double *X = ws.X.as<double>();
    double *W = ws.W.as<double>();
    hipLaunchKernelGGL(k_zero_int, dim3(256), dim3(256), 0, st, ws.blkexp.as<int>(), (long)nsteps * ldx); XTG_LAUNCH_CHECK("k_zero_int");
    hipLaunchKernelGGL(k_zero_int, dim3(64), dim3(256), 0, st, ws.colexp.as<int>(), (long)m); XTG_LAUNCH_CHECK("k_zero_int");
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, ws.flag.as<int>(), 4L); XTG_LAUNCH_CHECK("k_zero_int");

    // Steps go bottom-up in groups of GROUP_STEPS row blocks [g0, g1).  Within a group each
    // step updates only the group's rows above it ("near" update, K = 2 nb); after the group
    // one "far" update applies all of its blocks to the rows above g0 (K = 2 (g1 - g0)).
    // The group's operand W holds rows [0, g1-g0) (S part) and [g1-g0, 2(g1-g0)) (P part).
    int c_begin = adjoint ? 0 : m;
    int c_end = adjoint ? 0 : m;
    static const bool lookahead_env = std::getenv("XTGEVC3_NO_LOOKAHEAD") == nullptr;
    const bool lookahead = lookahead_env && !ws.fixed;
    if (lookahead) {
        XTG_CHECK(far_stream_init(ws));
    }
    bool f_pending = false; // a far update on ws.fs has not been waited for yet
    int last_par = 0;       // its group parity
    for (int gs = 0; gs < nsteps; gs += GROUP_STEPS) {
        const int ge = std::min(nsteps, gs + GROUP_STEPS);
        const int g1 = i0s[gs] + nbs[gs];
        const int g0 = i0s[ge - 1];
        const int wr = g1 - g0;
        bool any_active = false;
        // this group's copies of the operand, group scales and their digits / exponents
        const int par = (gs / GROUP_STEPS) & 1;
        if (f_pending && last_par == par) {
            XTG_CHECK(hipStreamWaitEvent(st, ws.ev_fdone[last_par], 0));
            f_pending = false;
        }
        double *Wb = W + (size_t)par * wsz;
        double *sgb = ws.sig_grp.as<double>() + (size_t)par * m;
        int *wgexpb = split ? ws.wgexp.as<int>() + (size_t)par * m : nullptr;
        signed char *Wgb = split ? ws.Wg.as<signed char>() + (size_t)par * SPLIT_DIGITS * m * WG_K : nullptr;
        MixedExp mxb = mx;
        if (mx.on) {
            mxb.wgrp = mx.wgrp + (size_t)par * m;
        }
        // the group's blocks top-down
        GroupInfo grp{};
        grp.wr = wr;
        {
            int off = 0;
            for (int jb = 0; jb < 4; ++jb) {
                const int sj = ge - 1 - jb;
                grp.off[jb] = off;
                grp.step[jb] = sj >= gs ? sj : ge - 1;
                if (sj >= gs) {
                    off += nbs[sj];
                }
            }
            for (int jb = 4; jb <= GROUP_MAX_BLOCKS; ++jb) {
                grp.off[jb] = off;
            }
        }
        for (int s = gs; s < ge; ++s) {
            const int i0 = i0s[s], nb = nbs[s];
            if (adjoint) {
                // adjoint columns (lo ascending) are active while their rows [lo, n) meet the block
                while (c_end < m && desc[c_end].lo < i0 + nb) {
                    ++c_end;
                }
                while (c_end > 0 && desc[c_end - 1].lo >= i0 + nb) {
                    --c_end;
                }
            }
            else {
                while (c_begin > 0 && desc[c_begin - 1].d >= i0) {
                    --c_begin;
                }
            }
            if (c_begin >= c_end) {
                continue;
            }
            any_active = true;
            const int u_begin = col2unit[c_begin], u_end = col2unit[c_end - 1] + 1;
            // tiles meeting [u_begin, u_end): groups are ordered, so the range is contiguous
            int t_begin = 0;
            while (t_begin < (int)tiles.size() && tmax[t_begin] < u_begin) {
                ++t_begin;
            }
            int t_end = (int)tiles.size();
            while (t_end > t_begin && tmin[t_end - 1] >= u_end) {
                --t_end;
            }
            const int ntl = t_end - t_begin;
            const int N = c_end - c_begin;
            const int woff = i0 - g0;
            grp.last = s == ge - 1 ? 1 : 0;
            // 2x2 bumps of S in the block: bit r-1 <=> rows (r-1, r)
            unsigned long long bumps = 0;
            for (int r = 1; r < nb; ++r) {
                if (subw[i0 + r] != 0.0) {
                    bumps |= 1ull << (r - 1);
                }
            }
            const int wpb = std::min(SOLVE_WAVES_H, ntl);
            // near update (fused into the solve): the group's rows [r0, i0) above this block
            const int r0 = adjoint ? std::max(g0, std::min(desc[c_begin].lo, i0)) : g0;
            const int M = i0 - r0;
            int pr = prof_begin(st);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_solve_real<MAXB_REAL>),
                               dim3((ntl + wpb - 1) / wpb), dim3(wpb * WAVE), 0, st, pk + spk[s].boff,
                               M > 0 ? pk + spk[s].poff : (const double2 *)nullptr, g0, r0, i0, nb, bumps, X, ldx, Wb,
                               ldw, ws.desc.as<UnitDesc>(), t_begin, t_end, u_begin, u_end, ws.sig.as<double>(),
                               ws.colexp.as<int>(), ws.blkexp.as<int>() + (long)s * ldx, ws.flag.as<int>(), woff,
                               wr + woff, sgb, s == gs ? 1 : 0, grp, mxb,
                               split ? SplitOut{ws.wbexp.as<int>() + (long)s * ldx, wgexpb,
                                                ws.xbexp.as<int>() + (long)s * ldx}
                                     : SplitOut{nullptr, nullptr, nullptr}); XTG_LAUNCH_CHECK("k_solve_real");
            prof_end(pr, st, 0, 2.0 * nb * nb * N + 2.0 * std::max(M, 0) * N * (2.0 * nb),
                     8.0 * (4.0 * nb * N + 2.0 * nb * nb + 2.0 * std::max(M, 0) * N + 2.0 * nb * std::max(M, 0)));
        }
        if (!any_active) {
            continue;
        }
        // far update: rows above the group, all of its blocks at once.  With the lookahead the
        // rows of the next group (which its solves need next) run on st and the rows above it
        // on ws.fs, overlapping the next group's steps; the next far update, which writes those
        // rows again, waits for them.
        const int N = c_end - c_begin;
        const int r0 = adjoint ? std::min(desc[c_begin].lo, g0) : 0;
        const int M = g0 - r0;
        if (f_pending) {
            XTG_CHECK(hipStreamWaitEvent(st, ws.ev_fdone[last_par], 0));
            f_pending = false;
        }
        if (M <= 0 || N <= 0) {
            continue;
        }
        const int gen = std::min(nsteps, ge + GROUP_STEPS);
        const int rsplit = lookahead && ge < nsteps ? std::max(r0, i0s[gen - 1]) : r0;
        int klo = 0, kp1 = 0;
        signed char *wg = nullptr;
        const long wslice = (long)m * WG_K, wkb = (long)m * SPLIT_KSTEP;
        if (split) {
            // digits of the group operand (exponent per column over both parts, from the
            // group's last solve), each part shifted by off so the K range starts at the K step
            // boundary klo <= g0
            klo = g0 / SPLIT_KSTEP * SPLIT_KSTEP;
            const int off = g0 - klo;
            kp1 = round_up(off + wr, SPLIT_KSTEP);
            if (pencil_ready != nullptr) {
                XTG_CHECK(hipStreamWaitEvent(st, pencil_ready, 0));
                pencil_ready = nullptr;
            }
            int pr = prof_begin(st);
            wg = Wgb + (long)c_begin * SPLIT_KSTEP;
            hipLaunchKernelGGL(k_slice_digits, dim3((N + 63) / 64, (2 * kp1 + 15) / 16), dim3(256), 0, st,
                               (const double *)(Wb + c_begin), ldw, off, wr, (const double *)(Wb + (long)wr * ldw + c_begin),
                               ldw, kp1 + off, wr, N, 2 * kp1, (const int *)(wgexpb + c_begin), wg, wslice,
                               wkb, 0, (signed char *)nullptr); XTG_LAUNCH_CHECK("k_slice_digits");
            prof_end(pr, st, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 8.0 * (2.0 * wr * N) + SPLIT_DIGITS * 2.0 * kp1 * N);
        }
        // the product for rows [ra, rb) on stream fst
        auto far = [&](int ra, int rb, hipStream_t fst) -> int {
            const int Mp = rb - ra;
            if (Mp <= 0) {
                return 0;
            }
            int pr = prof_begin(fst);
            if (split) {
                const long kbs = (long)n * SPLIT_KSTEP, ko = (long)(klo / SPLIT_KSTEP) * kbs + (long)ra * SPLIT_KSTEP;
                XTG_CHECK((hipError_t)launch_gemm_i8(fst, Mp, N, kp1, kp1, ws.Sq.as<signed char>() + ko,
                                                     ws.Pq.as<signed char>() + ko, qsl, kbs, wg, wslice, wkb,
                                                     ws.rowexp.as<int>() + ra, wgexpb + c_begin,
                                                     X + (long)ra * ldx + c_begin, ldx, 1, nullptr, 1, sgb + c_begin,
                                                     nullptr));
                prof_end(pr, fst, XTGEVC3_PROF_UPDATE_SPLIT, 2.0 * Mp * N * (2.0 * wr),
                         8.0 * (2.0 * Mp * N) + SPLIT_DIGITS * (2.0 * kp1 * Mp + 2.0 * kp1 * N));
            }
            else {
                const int gt = launch_gemm_f64(0, fst, Mp, N, wr, wr, Sw + ra + (long)g0 * ldsw, ldsw,
                                   Pw + ra + (long)g0 * ldpw, ldpw, (const double *)(Wb + c_begin), ldw,
                                   X + (long)ra * ldx + c_begin, ldx, (const int *)nullptr, 1,
                                   (const double *)(sgb + c_begin), (const int *)nullptr,
                                   (const double *)nullptr, prec, mx.on ? ws.rowexp.as<int>() + ra : nullptr,
                                   mx.on ? mxb.wgrp + c_begin : nullptr, 1, 0, 0,
                                   adjoint ? ws.colo.as<int>() + c_begin : nullptr, ra); XTG_LAUNCH_CHECK("k_gemm_f64<0>");
                prof_end(pr, fst, update_kind(gt), 2.0 * Mp * N * (2.0 * wr),
                         8.0 * (2.0 * Mp * N + 2.0 * wr * Mp + 2.0 * wr * N));
            }
            return 0;
        };
        if (const int frc = far(rsplit, g0, st)) {
            return frc;
        }
        if (rsplit > r0) {
            XTG_CHECK(hipEventRecord(ws.ev_wready[par], st));
            XTG_CHECK(hipStreamWaitEvent(ws.fs, ws.ev_wready[par], 0));
            if (const int frc = far(r0, rsplit, ws.fs)) {
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
    if (pencil_ready != nullptr) {
        XTG_CHECK(hipStreamWaitEvent(st, pencil_ready, 0)); // the side stream's digits are in use until here
    }
    // apply the per-column scale history to the solved blocks (fixup == false: the caller's
    // pass over X applies it, k_slice_xfix)
    if (fixup) {
        // the row range of each step (steps run bottom-up)
        std::vector<int> srow(2 * (size_t)nsteps);
        for (int s = 0; s < nsteps; ++s) {
            srow[2 * s] = i0s[s];
            srow[2 * s + 1] = i0s[s] + nbs[s];
        }
        ENSURE(ws.srow, sizeof(int) * 2 * (size_t)nsteps);
        XTG_CHECK(hipMemcpyAsync(ws.srow.p, srow.data(), sizeof(int) * 2 * nsteps, hipMemcpyHostToDevice, st));
        dim3 grid((m + 255) / 256, nsteps);
        hipLaunchKernelGGL(k_fixup_steps, grid, dim3(256), 0, st, X, ldx, m, (const int *)ws.srow.as<int>(),
                           (const int *)ws.blkexp.as<int>(), ldx, (const int *)ws.colexp.as<int>(),
                           (const int *)ws.flag.as<int>()); XTG_LAUNCH_CHECK("k_fixup_steps");
    }
    if (split) {
        // exponents of the final columns of X from the per-step block exponents
        hipLaunchKernelGGL(k_xexp_final, dim3((m + 255) / 256), dim3(256), 0, st, (const int *)ws.xbexp.as<int>(),
                           (const int *)ws.blkexp.as<int>(), ldx, (const int *)ws.colexp.as<int>(),
                           (const int *)ws.sstart.as<int>(), nsteps, m, ws.xexp.as<int>()); XTG_LAUNCH_CHECK("k_xexp_final");
    }
    XTG_CHECK(hipGetLastError());
    return 0;
}

// Creates the split form's side stream and events once per device workspace.
static hipError_t split_side_init(RealWork &ws)
{
    if (ws.side == nullptr) {
        hipError_t e = hipStreamCreateWithFlags(&ws.side, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ws.side = nullptr;
            return e;
        }
        for (hipEvent_t *ev : {&ws.ev_start, &ws.ev_pencil, &ws.ev_z, &ws.ev_y, &ws.ev_rhs}) {
            e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (e != hipSuccess) {
                return e;
            }
        }
    }
    return hipSuccess;
}

// Creates the far-update lookahead stream and its events once per device workspace.
static hipError_t far_stream_init(RealWork &ws)
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
        for (hipEvent_t *ev : {&ws.ev_zsrc, &ws.ev_zcopy}) {
            e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (e != hipSuccess) {
                return e;
            }
        }
    }
    return hipSuccess;
}

// Copies Z (n x n, leading dimension ldz) into ws.Zt on ws.fs after the work queued so far on
// st, so that the copy runs beside the sweep; ws.ev_zcopy marks its end.
static hipError_t zcopy_async(RealWork &ws, const double *Z, long ldz, int n, hipStream_t st)
{
    hipError_t e = far_stream_init(ws);
    if (e == hipSuccess) {
        e = hipEventRecord(ws.ev_zsrc, st);
    }
    if (e == hipSuccess) {
        e = hipStreamWaitEvent(ws.fs, ws.ev_zsrc, 0);
    }
    if (e == hipSuccess) {
        e = hipMemcpy2DAsync(ws.Zt.p, sizeof(double) * n, Z, sizeof(double) * ldz, sizeof(double) * n, n,
                             hipMemcpyDeviceToDevice, ws.fs);
    }
    if (e == hipSuccess) {
        e = hipEventRecord(ws.ev_zcopy, ws.fs);
    }
    return e;
}

// The column 1-norms of S and P and the subdiagonal of S (3n values in ws.norms, written on st)
// read back to out.  The copy goes to pinned memory and the host waits for it alone, after
// `pre` has queued on the streams work that does not need the values (the device keeps busy
// while the host builds the launch descriptors from them); the allocation-free workspace
// waits for the stream.
static int read_norms(RealWork &ws, int n, hipStream_t st, std::vector<double> &out, const std::function<int()> &pre)
{
    const size_t bytes = sizeof(double) * 3 * (size_t)n;
    out.resize(3 * (size_t)n);
    if (ws.fixed) {
        XTG_CHECK(hipMemcpyAsync(out.data(), ws.norms.p, bytes, hipMemcpyDeviceToHost, st));
        XTG_CHECK(hipStreamSynchronize(st));
        return pre ? pre() : 0;
    }
    XTG_CHECK(ws.hnorms.ensure(bytes));
    if (ws.ev_norms == nullptr) {
        XTG_CHECK(hipEventCreateWithFlags(&ws.ev_norms, hipEventDisableTiming));
    }
    XTG_CHECK(hipMemcpyAsync(ws.hnorms.p, ws.norms.p, bytes, hipMemcpyDeviceToHost, st));
    XTG_CHECK(hipEventRecord(ws.ev_norms, st));
    if (pre) {
        if (const int rc = pre()) {
            return rc;
        }
    }
    XTG_CHECK(hipEventSynchronize(ws.ev_norms));
    const double *h = static_cast<const double *>(ws.hnorms.p);
    std::copy(h, h + 3 * (size_t)n, out.begin());
    return 0;
}

// Digits of the rows of A (and B, with the exponent of each row taken over both): E[r] from the
// exact row maxima, then QA (QB) = k-blocked digits over [0, kp) (k-block pitch 32 rows).
// With cfirst, E[r] bounds only the columns k >= cfirst[r] (the far-update region of a pencil
// row); digits left of it may saturate and must only meet zero digits of the other operand.
static hipError_t split_rows(hipStream_t st, int rows, int cols, int kp, const double *A, long lda, const double *B,
                             long ldb, int *E, signed char *QA, signed char *QB, const int *cfirst = nullptr)
{
    hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, E, (long)rows, -1000);
    hipLaunchKernelGGL(k_rowexp, dim3((rows + 255) / 256, (cols + 63) / 64), dim3(256), 0, st, A, lda, B, ldb, rows,
                       cols, E, cfirst);
    const dim3 grid((rows + 63) / 64, (kp + 15) / 16);
    const long qs = (long)rows * kp, kbp = (long)rows * SPLIT_KSTEP;
    hipLaunchKernelGGL(k_slice_digits, grid, dim3(256), 0, st, A, lda, 0, cols, (const double *)nullptr, 0L, 0, 0, rows,
                       kp, (const int *)E, QA, qs, kbp, 0, (signed char *)nullptr);
    if (B != nullptr) {
        hipLaunchKernelGGL(k_slice_digits, grid, dim3(256), 0, st, B, ldb, 0, cols, (const double *)nullptr, 0L, 0, 0,
                           rows, kp, (const int *)E, QB, qs, kbp, 0, (signed char *)nullptr);
    }
    return hipGetLastError();
}

static int run_real_side(RealWork &ws, const RealSideArgs &A, const std::vector<Unit> &units, int m, hipStream_t st)
{
    const int n = A.n;
    // --- column descriptors in working order -------------------------------------------
    std::vector<ColDesc> desc(m);
    std::vector<int> outcol(m), kmaxv(m);
    std::vector<int> vcol, vw;
    int p = 0;
    for (size_t u = 0; u < units.size(); ++u) {
        const Unit &un = A.left ? units[units.size() - 1 - u] : units[u];
        const int k = un.k;
        double a, bR, bI;
        if (un.w == 1) {
            shift_real(A.ar[k], 0.0, A.be[k], A.ascale, A.bscale, a, bR, bI, false, A.safemin);
            ColDesc c{};
            c.a = a;
            c.bR = bR;
            c.bI = 0.0;
            c.d = A.left ? n - 1 - k : k;
            c.lo = 0;
            c.partner = p;
            c.flags = INIT_UNIT << 2;
            desc[p] = c;
            outcol[p] = un.outc;
            kmaxv[p] = c.d + 1;
            ++p;
        }
        else {
            shift_real(A.ar[k], A.ai[k], A.be[k], A.ascale, A.bscale, a, bR, bI, true, A.safemin);
            ColDesc c{};
            c.a = a;
            c.bR = bR;
            c.bI = A.left ? -bI : bI; // left eigenvectors use conj(bcoeff) (dtgevc3.cpp:1584-1591)
            c.d = A.left ? n - 2 - k : k;
            c.lo = 0;
            c.flags = 2 | ((A.left ? INIT_PAIR_ROW2 : INIT_PAIR_ROW1) << 2);
            ColDesc ci = c;
            c.partner = p + 1;
            ci.partner = p;
            ci.flags |= 1;
            desc[p] = c;
            desc[p + 1] = ci;
            outcol[p] = un.outc;     // real part
            outcol[p + 1] = un.outc + 1; // imaginary part
            kmaxv[p] = kmaxv[p + 1] = c.d + 2;
            p += 2;
        }
    }
    for (const Unit &un : units) {
        vcol.push_back(un.outc);
        vw.push_back(un.w);
    }
    const long ldx = ((long)m + 7) / 8 * 8;
    ENSURE(ws.outcol, sizeof(int) * (size_t)m);
    ENSURE(ws.kmax, sizeof(int) * (size_t)m);
    ENSURE(ws.vcol, sizeof(int) * vcol.size());
    ENSURE(ws.vw, sizeof(int) * vw.size());
    XTG_CHECK(hipMemcpyAsync(ws.outcol.p, outcol.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.kmax.p, kmaxv.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.vcol.p, vcol.data(), sizeof(int) * vcol.size(), hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.vw.p, vw.data(), sizeof(int) * vw.size(), hipMemcpyHostToDevice, st));
    ENSURE(ws.X, sizeof(double) * (size_t)n * ldx);
    if (!A.x_ready) {
        hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, ws.X.as<double>(), (long)n * ldx); XTG_LAUNCH_CHECK("k_zero");
    }
    const bool split = A.prec == XTGEVC3_PREC_SPLIT_I8;
    if (split) {
        // digits of the working pencil (and of Z for the back-transform) on the side stream,
        // after everything the caller's stream holds so far
        const long qld = split_qld(n), qsl = (long)n * qld;
        const int kb = round_up(n, SPLIT_KSTEP);
        ENSURE(ws.rowexp, sizeof(int) * (size_t)n);
        ENSURE(ws.Sq, (size_t)SPLIT_DIGITS * qsl);
        ENSURE(ws.Pq, (size_t)SPLIT_DIGITS * qsl);
        if (A.back) {
            ENSURE(ws.zexp, sizeof(int) * (size_t)n);
            ENSURE(ws.Zq, (size_t)SPLIT_DIGITS * n * kb);
        }
        std::vector<int> cfar;
        far_columns(n, *A.subw, cfar);
        ENSURE(ws.cfar, sizeof(int) * (size_t)n);
        XTG_CHECK(hipMemcpyAsync(ws.cfar.p, cfar.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        XTG_CHECK(split_side_init(ws));
        XTG_CHECK(hipEventRecord(ws.ev_start, st));
        XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_start, 0));
        int pr = prof_begin(ws.side);
        XTG_CHECK(split_rows(ws.side, n, n, (int)qld, A.Sw, A.ldsw, A.Pw, A.ldpw, ws.rowexp.as<int>(),
                             ws.Sq.as<signed char>(), ws.Pq.as<signed char>(), ws.cfar.as<int>()));
        prof_end(pr, ws.side, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 8.0 * 4.0 * n * (double)n + 2.0 * SPLIT_DIGITS * qsl);
        XTG_CHECK(hipEventRecord(ws.ev_pencil, ws.side));
        if (A.back) {
            const double *Rp = A.left ? A.Z + (long)(n - 1) * A.ldz : A.Z;
            const long ldr = A.left ? -A.ldz : A.ldz;
            pr = prof_begin(ws.side);
            XTG_CHECK(split_rows(ws.side, n, n, kb, Rp, ldr, nullptr, 0, ws.zexp.as<int>(), ws.Zq.as<signed char>(),
                                 nullptr));
            prof_end(pr, ws.side, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 8.0 * 2.0 * n * (double)n + SPLIT_DIGITS * (double)n * kb);
            XTG_CHECK(hipEventRecord(ws.ev_z, ws.side));
        }
    }
    // the split back-transform makes the digits of X with the scale fix-up in the same pass
    const bool fused_fix = split && A.back;
    int rc = sweep_real(ws, n, A.Sw, A.ldsw, A.Pw, A.ldpw, *A.subw, desc, false, ldx, st, A.prec,
                        split ? ws.ev_pencil : nullptr, !fused_fix);
    if (rc) {
        return rc;
    }
    double *X = ws.X.as<double>();

    // --- output: back-transform or copy, then normalise ---------------------------------
    if (A.back && A.zf != nullptr) {
        const hipEvent_t zready = A.zf->ready(A.left);
        if (zready == nullptr) {
            return -1000;
        }
        XTG_CHECK(hipStreamWaitEvent(st, zready, 0));
    }
    if (A.back) {
        // V[:, outcol(p)] = Zw * X[:, p]; for left eigenvectors Zw = Q J (reversed columns)
        const double *Rp = A.left ? A.Z + (long)(n - 1) * A.ldz : A.Z;
        long ldr = A.left ? -A.ldz : A.ldz;
        double ksum = 0.0;
        int kmx = 0;
        for (int q = 0; q < m; ++q) {
            ksum += kmaxv[q];
            kmx = std::max(kmx, kmaxv[q]);
        }
        if (A.prec == 1) {
            // mixed form: exponents of the final X columns and of the rows of Z (Q)
            ENSURE(ws.xfin, sizeof(int) * (size_t)m);
            ENSURE(ws.zexp, sizeof(int) * (size_t)n);
            hipLaunchKernelGGL(k_xfin, dim3((m + 255) / 256), dim3(256), 0, st, (const int *)ws.colexp.as<int>(),
                               (const int *)ws.xrel.as<int>(), m, ws.xfin.as<int>()); XTG_LAUNCH_CHECK("k_xfin");
            hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.zexp.as<int>(), (long)n, -1000); XTG_LAUNCH_CHECK("k_fill_int");
            hipLaunchKernelGGL(k_rowexp, dim3((n + 255) / 256, (n + 63) / 64), dim3(256), 0, st, A.Z, A.ldz,
                               (const double *)nullptr, 0L, n, n, ws.zexp.as<int>(), (const int *)nullptr); XTG_LAUNCH_CHECK("k_rowexp");
        }
        if (split) {
            // digits of the columns of X (exponents from the sweep); Z's come from the side stream
            const int kb = round_up(n, SPLIT_KSTEP);
            ENSURE(ws.Xq, (size_t)SPLIT_DIGITS * m * kb);
            int pr = prof_begin(st);
            hipLaunchKernelGGL(k_slice_xfix, dim3((m + 63) / 64, (kb + 15) / 16), dim3(256), 0, st, (const double *)X,
                               ldx, n, m, kb, (const int *)ws.rowstep.as<int>(), (const int *)ws.blkexp.as<int>(), ldx,
                               (const int *)ws.colexp.as<int>(), (const int *)ws.xexp.as<int>(),
                               ws.Xq.as<signed char>(), (long)m * kb, (long)m * SPLIT_KSTEP); XTG_LAUNCH_CHECK("k_slice_xfix");
            prof_end(pr, st, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 8.0 * (double)n * m + SPLIT_DIGITS * (double)m * kb);
            XTG_CHECK(hipStreamWaitEvent(st, ws.ev_z, 0));
            pr = prof_begin(st);
            XTG_CHECK((hipError_t)launch_gemm_i8(st, m, n, kb, 0, ws.Xq.as<signed char>(), ws.Xq.as<signed char>(),
                                                 (long)m * kb, (long)m * SPLIT_KSTEP, ws.Zq.as<signed char>(),
                                                 (long)n * kb, (long)n * SPLIT_KSTEP,
                                                 ws.xexp.as<int>(), ws.zexp.as<int>(), A.V, A.ldv, 1,
                                                 ws.outcol.as<int>(), 0, nullptr, ws.kmax.as<int>()));
            prof_end(pr, st, XTGEVC3_PROF_BACK_SPLIT, 2.0 * n * ksum,
                     8.0 * (double)n * m + SPLIT_DIGITS * ((double)n * kmx + ksum));
        }
        else if (A.zf != nullptr && A.zf->done && A.zf->chunks > 1) {
            // pieces of consecutive working units (about equal column counts), each normalised
            // and handed to the caller as soon as it is queued; a piece's output columns are
            // contiguous (the units of a side are in output order, reversed for the left side)
            const int U = (int)units.size(), C = std::min(A.zf->chunks, U);
            std::vector<int> ustart(U + 1, 0);
            for (int u = 0; u < U; ++u) {
                ustart[u + 1] = ustart[u] + (A.left ? units[U - 1 - u] : units[u]).w;
            }
            for (int i = 0; i < C; ++i) {
                const int u0 = (int)((long)U * i / C), u1 = (int)((long)U * (i + 1) / C);
                const int p0 = ustart[u0], p1 = ustart[u1];
                int c0 = n, c1 = 0;
                double ks = 0.0;
                int kx = 0;
                for (int q = p0; q < p1; ++q) {
                    c0 = std::min(c0, outcol[q]);
                    c1 = std::max(c1, outcol[q] + 1);
                    ks += kmaxv[q];
                    kx = std::max(kx, kmaxv[q]);
                }
                int pr = prof_begin(st);
                launch_gemm_f64(1, st, p1 - p0, n, n, 0, (const double *)X + p0, ldx, (const double *)nullptr, 0L, Rp,
                                ldr, A.V, A.ldv, ws.outcol.as<int>() + p0, 0, (const double *)nullptr,
                                ws.kmax.as<int>() + p0, (const double *)nullptr, A.prec,
                                A.prec == 1 ? ws.xfin.as<int>() + p0 : nullptr,
                                A.prec == 1 ? ws.zexp.as<int>() : nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<1>");
                prof_end(pr, st, 2, 2.0 * n * ks, 8.0 * ((double)n * kx + ks + (double)n * (p1 - p0)));
                const int v0 = A.left ? U - u1 : u0; // the pieces' units in vcol / vw order
                hipLaunchKernelGGL(k_normalize, dim3((unsigned)(u1 - u0)), dim3(256), 0, st, A.V, A.ldv, n,
                                   ws.vcol.as<int>() + v0, ws.vw.as<int>() + v0); XTG_LAUNCH_CHECK("k_normalize");
                if (float *o = static_cast<float *>(A.zf->out32[A.left ? 1 : 0])) {
                    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)(c1 - c0)), dim3(256), 0, st,
                                       (const double *)A.V + (long)c0 * A.ldv, A.ldv, (long)n, o + (long)c0 * n,
                                       (long)n); XTG_LAUNCH_CHECK("k_round_f32");
                }
                if (const int drc = A.zf->done(A.left, c0, c1, st)) {
                    return drc;
                }
            }
            XTG_CHECK(hipGetLastError());
            return 0;
        }
        else {
            int pr = prof_begin(st);
            launch_gemm_f64(1, st, m, n, n, 0, (const double *)X, ldx, (const double *)nullptr, 0L, Rp, ldr, A.V,
                            A.ldv, ws.outcol.as<int>(), 0, (const double *)nullptr, ws.kmax.as<int>(),
                            (const double *)nullptr, A.prec, A.prec == 1 ? ws.xfin.as<int>() : nullptr,
                            A.prec == 1 ? ws.zexp.as<int>() : nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<1>");
            prof_end(pr, st, 2, 2.0 * n * ksum, 8.0 * ((double)n * kmx + ksum + (double)n * m));
        }
    }
    else {
        dim3 grid((m + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(k_copy_out, grid, dim3(32, 8), 0, st, (const double *)X, ldx, n, m, A.V, A.ldv,
                           ws.outcol.as<int>(), A.left ? 1 : 0); XTG_LAUNCH_CHECK("k_copy_out");
    }
    hipLaunchKernelGGL(k_normalize, dim3((unsigned)vcol.size()), dim3(256), 0, st, A.V, A.ldv, n, ws.vcol.as<int>(),
                       ws.vw.as<int>()); XTG_LAUNCH_CHECK("k_normalize");
    XTG_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------------------
// Device-pointer entry for the real interface.  S, P, VL, VR are device arrays; select,
// alphar, alphai, beta are host arrays.  No argument checks beyond the reference's: call
// check_args first.
// ---------------------------------------------------------------------------------------
static int real_device_ws(RealWork &ws, char side, char howmny, const int *select, int n, const double *S, long lds,
                          const double *P, long ldp, const double *ar, const double *ai, const double *be, double *VL,
                          long ldvl, double *VR, long ldvr, hipStream_t st, int prec, const ZFeed *zf,
                          double safemin);

int real_device(char side, char howmny, const int *select, int n, const double *S, long lds, const double *P,
                long ldp, const double *ar, const double *ai, const double *be, double *VL, long ldvl, double *VR,
                long ldvr, hipStream_t st, int prec, const ZFeed *zf, double safemin)
{
    if (zf != nullptr && prec == XTGEVC3_PREC_SPLIT_I8) {
        set_error("Z delivered beside the sweep is not supported in the split form");
        return -1000;
    }
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    RealWork &ws = g_ws[dev & 63];
    WsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    return real_device_ws(ws, side, howmny, select, n, S, lds, P, ldp, ar, ai, be, VL, ldvl, VR, ldvr, st, prec, zf,
                          safemin);
}

// ---------------------------------------------------------------------------------------
// Allocation-free variant of the FP64 device entry: the workspace is carved from one static
// device array per GPU (`scratch`, NOALLOC_SCRATCH_BYTES, held until the process ends), so a
// call makes no hipMalloc / hipFree and creates no stream or event; everything runs on the
// caller's stream.  Sizes up to NOALLOC_MAX_N.  The array's device address comes from
// hipGetSymbolAddress once per device (first call on it, with the device current); host code
// uses only that address.  Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------------------
constexpr int NOALLOC_MAX_N = XTGEVC3_NOALLOC_MAX_N;
constexpr size_t na_slot(size_t bytes) { return (bytes + 255) / 256 * 256; }
constexpr size_t NA_N = NOALLOC_MAX_N;
constexpr size_t NA_LDX = (NA_N + 7) / 8 * 8;
constexpr size_t NA_STEPS = NA_N / 32 + 2; // row blocks have at least 63 rows
// slot sizes (bytes) of the buffers the FP64 path uses, at the maximum size
constexpr size_t NA_BYTES[] = {
    sizeof(double) * 3 * NA_N,                                 // norms
    sizeof(double) * 2 * (2 * GROUP_STEPS * 65) * NA_LDX,      // W (two group copies)
    sizeof(double) * NA_N,                                     // sig
    sizeof(double) * 2 * NA_N,                                 // sig_grp
    sizeof(int) * NA_N,                                        // colexp
    sizeof(int) * NA_STEPS * NA_LDX,                           // blkexp
    sizeof(int) * 4,                                           // flag
    sizeof(UnitDesc) * 4 * 2 * NA_N,                           // desc (tile order: four per tile)
    sizeof(int) * NA_N,                                        // rowstep
    sizeof(int) * 2 * NA_STEPS,                                // srow
    // pk: a block image per step, and panels of at most 13 + 9 + 5 row tiles per group of 4 steps
    sizeof(double2) * (NA_STEPS * PK_BLK + (NA_STEPS + 3) / 4 * 27 * (PK_KSMAX / 4) * 64),
    sizeof(StepPack) * NA_STEPS,                               // pksteps
    sizeof(int) * NA_N,                                        // outcol
    sizeof(int) * NA_N,                                        // kmax
    sizeof(int) * NA_N,                                        // vcol
    sizeof(int) * NA_N,                                        // vw
    sizeof(double) * NA_N * NA_LDX,                            // X
    sizeof(double) * NA_N * NA_N,                              // Zt
    sizeof(double) * NA_N * NA_N,                              // Sw
    sizeof(double) * NA_N * NA_N,                              // Pw
};
constexpr size_t na_total()
{
    size_t t = 0;
    for (size_t b : NA_BYTES) {
        t += na_slot(b);
    }
    return t;
}
constexpr size_t NOALLOC_SCRATCH_BYTES = na_total();
static_assert(NOALLOC_SCRATCH_BYTES <= (size_t)XTGEVC3_NOALLOC_SCRATCH_BYTES, "documented scratch size");

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_na_mutex;
static RealWork g_na_ws[64];
static void *g_na_ptr[64]; // device address of `scratch` on each device

static hipError_t noalloc_setup(int dev)
{
    if (g_na_ptr[dev] != nullptr) {
        return hipSuccess;
    }
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        return e;
    }
    if (ptr == nullptr) {
        return hipErrorInvalidSymbol;
    }
    RealWork &ws = g_na_ws[dev];
    DevBuf *slots[] = {&ws.norms, &ws.W,      &ws.sig,  &ws.sig_grp, &ws.colexp, &ws.blkexp,
                       &ws.flag,  &ws.desc,   &ws.rowstep, &ws.srow, &ws.pk,     &ws.pksteps, &ws.outcol, &ws.kmax,
                       &ws.vcol,  &ws.vw,     &ws.X,    &ws.Zt,      &ws.Sw,     &ws.Pw};
    static_assert(sizeof(slots) / sizeof(slots[0]) == sizeof(NA_BYTES) / sizeof(NA_BYTES[0]), "one slot per buffer");
    size_t off = 0;
    for (size_t i = 0; i < sizeof(NA_BYTES) / sizeof(NA_BYTES[0]); ++i) {
        slots[i]->p = static_cast<unsigned char *>(ptr) + off;
        slots[i]->cap = NA_BYTES[i];
        slots[i]->fixed = true;
        off += na_slot(NA_BYTES[i]);
    }
    // every other buffer of the workspace is a zero-capacity fixed slot: using one is an error
    DevBuf *unused[] = {&ws.rowexp, &ws.zexp, &ws.wgrp, &ws.xrel, &ws.xfin, &ws.kcol, &ws.coef,
                        &ws.Vt,     &ws.Y,    &ws.LamS, &ws.bVp, &ws.bnfk, &ws.btiles, &ws.bkpre, &ws.Ut,  &ws.YB,   &ws.part,
                        &ws.tcol,   &ws.tw,   &ws.Sq,   &ws.Pq,   &ws.Wg,    &ws.wgexp, &ws.wbexp, &ws.xbexp,
                        &ws.sstart, &ws.Xq,   &ws.Zq,   &ws.xexp, &ws.cfar, &ws.colo, &ws.bdg};
    for (DevBuf *b : unused) {
        b->fixed = true;
    }
    ws.fixed = true;
    g_na_ptr[dev] = ptr;
    return hipSuccess;
}

int real_device_noalloc(char side, char howmny, const int *select, int n, const double *S, long lds,
                        const double *P, long ldp, const double *ar, const double *ai, const double *be, double *VL,
                        long ldvl, double *VR, long ldvr, hipStream_t st)
{
    if (n > NOALLOC_MAX_N) {
        set_error("n = " + std::to_string(n) + " exceeds the allocation-free entry's maximum of " +
                  std::to_string(NOALLOC_MAX_N));
        return XTGEVC3_INFO_UNSUPPORTED;
    }
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_na_mutex);
    XTG_CHECK(noalloc_setup(dev & 63));
    return real_device_ws(g_na_ws[dev & 63], side, howmny, select, n, S, lds, P, ldp, ar, ai, be, VL, ldvl, VR, ldvr,
                          st, XTGEVC3_PREC_FP64, nullptr, kDblMin);
}

static int real_device_ws(RealWork &ws, char side, char howmny, const int *select, int n, const double *S, long lds,
                          const double *P, long ldp, const double *ar, const double *ai, const double *be, double *VL,
                          long ldvl, double *VR, long ldvr, hipStream_t st, int prec, const ZFeed *zf,
                          double safemin)
{
    bool right = side == 'R' || side == 'B';
    bool left = side == 'L' || side == 'B';
    bool back = howmny == 'B';
    std::vector<Unit> units;
    int m = real_units(howmny, select, n, ai, units);
    if (m == 0 || (!(right && VR) && !(left && VL))) {
        return 0;
    }

    // global scales from the 1-norms of S and P over the upper Hessenberg part
    // column norms of S and P and the subdiagonal of S (block snapping) in one buffer
    ENSURE(ws.norms, sizeof(double) * 3 * (size_t)n);
    hipLaunchKernelGGL(k_colnorms, dim3(n), dim3(256), 0, st, S, lds, P, ldp, n, 1, ws.norms.as<double>()); XTG_LAUNCH_CHECK("k_colnorms");
    hipLaunchKernelGGL(k_subdiag, dim3((n + 255) / 256), dim3(256), 0, st, S, lds, n, ws.norms.as<double>() + 2 * (long)n); XTG_LAUNCH_CHECK("k_subdiag");
    // queued before the host waits for the norms: the zeroed X of the first side and the
    // back-transform's copy of Z for the right side (the output overwrites the input VR)
    const bool right_side = right && VR != nullptr;
    const bool zcopy_early = right_side && back && zf == nullptr && prec != XTGEVC3_PREC_SPLIT_I8 && !ws.fixed;
    std::vector<double> nrm;
    XTG_CHECK_RC(read_norms(ws, n, st, nrm, [&]() -> int {
        const long ldx = ((long)m + 7) / 8 * 8;
        ENSURE(ws.X, sizeof(double) * (size_t)n * ldx);
        hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, ws.X.as<double>(), (long)n * ldx); XTG_LAUNCH_CHECK("k_zero");
        if (zcopy_early) {
            ENSURE(ws.Zt, sizeof(double) * (size_t)n * n);
            XTG_CHECK(zcopy_async(ws, VR, ldvr, n, st));
        }
        return 0;
    }));
    // sub[i] = S[i, i-1]
    std::vector<double> sub(nrm.begin() + 2 * (long)n, nrm.end());
    double an = 0.0, bn = 0.0;
    for (int c = 0; c < n; ++c) {
        an = std::max(an, nrm[c]);
        bn = std::max(bn, nrm[n + c]);
    }
    double ascale = 1.0 / std::max(an, safemin);
    double bscale = 1.0 / std::max(bn, safemin);
    // the back-transform's own copy of Z (the output overwrites VR / VL) arrives on ws.fs
    ZFeed zown{};
    zown.ready = [&ws](bool) { return ws.ev_zcopy; };

    if (right && VR != nullptr) {
        RealSideArgs A{};
        A.n = n;
        A.Sw = S;
        A.ldsw = lds;
        A.Pw = P;
        A.ldpw = ldp;
        A.ar = ar;
        A.ai = ai;
        A.be = be;
        A.ascale = ascale;
        A.bscale = bscale;
        A.left = false;
        A.back = back;
        A.subw = &sub;
        A.prec = prec;
        A.V = VR;
        A.ldv = ldvr;
        A.zf = zf;
        A.safemin = safemin;
        A.x_ready = true;
        if (back && zf != nullptr) {
            A.Z = static_cast<const double *>(zf->zr);
            A.ldz = n;
        }
        else if (back && prec == XTGEVC3_PREC_SPLIT_I8) {
            // the split form reads Z only through its digits, made before the output is written
            A.Z = VR;
            A.ldz = ldvr;
        }
        else if (back) {
            // the input VR holds Z; keep a copy since the output overwrites it
            ENSURE(ws.Zt, sizeof(double) * (size_t)n * n);
            if (ws.fixed) {
                XTG_CHECK(hipMemcpy2DAsync(ws.Zt.p, sizeof(double) * n, VR, sizeof(double) * ldvr,
                                           sizeof(double) * n, n, hipMemcpyDeviceToDevice, st));
            }
            else {
                if (!zcopy_early) {
                    XTG_CHECK(zcopy_async(ws, VR, ldvr, n, st));
                }
                A.zf = &zown;
            }
            A.Z = ws.Zt.as<double>();
            A.ldz = n;
        }
        int rc = run_real_side(ws, A, units, m, st);
        if (rc) {
            return rc;
        }
    }
    if (left && VL != nullptr) {
        ENSURE(ws.Sw, sizeof(double) * (size_t)n * n);
        ENSURE(ws.Pw, sizeof(double) * (size_t)n * n);
        dim3 grid((n + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(k_reverse_transpose, grid, dim3(32, 8), 0, st, S, lds, n, ws.Sw.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        hipLaunchKernelGGL(k_reverse_transpose, grid, dim3(32, 8), 0, st, P, ldp, n, ws.Pw.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        // working subdiagonal: Sw[i, i-1] = S[n-i, n-1-i]
        std::vector<double> subw(n, 0.0);
        for (int i = 1; i < n; ++i) {
            subw[i] = sub[n - i];
        }
        RealSideArgs A{};
        A.n = n;
        A.Sw = ws.Sw.as<double>();
        A.ldsw = n;
        A.Pw = ws.Pw.as<double>();
        A.ldpw = n;
        A.ar = ar;
        A.ai = ai;
        A.be = be;
        A.ascale = ascale;
        A.bscale = bscale;
        A.left = true;
        A.back = back;
        A.subw = &subw;
        A.prec = prec;
        A.V = VL;
        A.ldv = ldvl;
        A.zf = zf;
        A.safemin = safemin;
        A.x_ready = !right_side; // after a right side, X is zeroed again for this side
        if (back && zf != nullptr) {
            A.Z = static_cast<const double *>(zf->zl);
            A.ldz = n;
        }
        else if (back && prec == XTGEVC3_PREC_SPLIT_I8) {
            A.Z = VL;
            A.ldz = ldvl;
        }
        else if (back) {
            // queued after the right side's back-transform, which reads the same workspace
            ENSURE(ws.Zt, sizeof(double) * (size_t)n * n);
            if (ws.fixed) {
                XTG_CHECK(hipMemcpy2DAsync(ws.Zt.p, sizeof(double) * n, VL, sizeof(double) * ldvl,
                                           sizeof(double) * n, n, hipMemcpyDeviceToDevice, st));
            }
            else {
                XTG_CHECK(zcopy_async(ws, VL, ldvl, n, st));
                A.zf = &zown;
            }
            A.Z = ws.Zt.as<double>();
            A.ldz = n;
        }
        int rc = run_real_side(ws, A, units, m, st);
        if (rc) {
            return rc;
        }
    }
    return 0;
}


// ---------------------------------------------------------------------------------------
// Backward of the selected right eigenvectors (no back-transform): dL/dS and dL/dP for
// nbatch cotangent batches G[:, b*m + c] of the normalised output V (n x m, as the forward
// returns it).  alpha_k = S_kk and beta_k = P_kk for the selected real eigenvalues.
// Returns -1001 through *info when a selected eigenvalue is complex or infinite.
// ---------------------------------------------------------------------------------------
static int real_backward_core(RealWork &ws, int n, const double *S, long lds, const double *P, long ldp,
                              const double *ar, const double *ai, const double *be, const int *select, const double *V,
                              long ldv, const double *G, long ldg, int nbatch, double *dS, long ldds, double *dP,
                              long lddp, int *info, hipStream_t st, int prec, bool rhs_given);

int real_backward_device(int n, const double *S, long lds, const double *P, long ldp, const double *ar,
                         const double *ai, const double *be, const int *select, const double *V, long ldv,
                         const double *G, long ldg, int nbatch, double *dS, long ldds, double *dP, long lddp, int *info,
                         hipStream_t st, int prec)
{
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    RealWork &ws = g_ws[dev & 63];
    WsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    return real_backward_core(ws, n, S, lds, P, ldp, ar, ai, be, select, V, ldv, G, ldg, nbatch, dS, ldds, dP, lddp,
                              info, st, prec, false);
}

// The backward core.  rhs_given: G already holds the adjoint right-hand sides (the cotangent of
// the unnormalised vector expressed in the Schur basis), V the Schur-basis vectors they belong
// to; otherwise V is the normalised output and the normalisation's correction is applied to G.
static int real_backward_core(RealWork &ws, int n, const double *S, long lds, const double *P, long ldp,
                              const double *ar, const double *ai, const double *be, const int *select, const double *V,
                              long ldv, const double *G, long ldg, int nbatch, double *dS, long ldds, double *dP,
                              long lddp, int *info, hipStream_t st, int prec, bool rhs_given)
{
    std::vector<Unit> units;
    int m = real_units(select ? 'S' : 'A', select, n, ai, units);
    *info = 0;
    for (const Unit &u : units) {
        if (u.w != 1 || std::fabs(be[u.k]) <= std::numeric_limits<double>::min()) {
            *info = -1001;
            return 0;
        }
    }
    if (m == 0 || nbatch <= 0) {
        return 0;
    }
    // the working columns (k descending, batches adjacent) and their buffers need no scales
    std::vector<int> kcol(m);
    for (int c = 0; c < m; ++c) {
        kcol[c] = units[c].k;
    }
    const int M = m * nbatch;
    const long ldx = ((long)M + 7) / 8 * 8;
    const long ldt = ((long)m + 7) / 8 * 8;
    const long npad = (n + 1) & ~1; // rows of the outer products' operands (16-byte pairs)
    ENSURE(ws.norms, sizeof(double) * 3 * (size_t)n);
    ENSURE(ws.kcol, sizeof(int) * (size_t)m);
    ENSURE(ws.coef, sizeof(double) * 4 * (size_t)m);
    ENSURE(ws.Sw, sizeof(double) * (size_t)n * n);
    ENSURE(ws.Pw, sizeof(double) * (size_t)n * n);
    ENSURE(ws.X, sizeof(double) * (size_t)n * ldx);
    ENSURE(ws.Vt, sizeof(double) * (size_t)n * ldt);
    ENSURE(ws.Y, sizeof(double) * (size_t)n * ldt);
    ENSURE(ws.LamS, sizeof(double) * (size_t)M * npad);
    ENSURE(ws.bVp, sizeof(double) * (size_t)npad * m);
    double *X = ws.X.as<double>();
    const bool split = prec == XTGEVC3_PREC_SPLIT_I8;
    // dS_b = V diag(-a) Lam_b, dP_b = V diag(b) Lam_b  (as D[j][i] = sum_c s_c V[j][c] Lam[c][i]),
    // all batches and both products in one launch.  v_c is zero below row kcol[c] and kcol
    // ascends, so with the columns taken in reverse order (negative strides) row j of the
    // product needs only the first kpre[j] = #{c : kcol[c] >= j} of them: a per-row K bound.
    // Only the support is formed (dS: rows i <= j + 1, dP: i <= j; zeros elsewhere), where entry
    // (i, j) needs the columns with kcol[c] >= max(j, i + 1) (lambda_c[i] = 0 for kcol[c] <= i).
    std::vector<int> kpre(n);
    for (int j = 0, c = 0; j < n; ++j) {
        while (c < m && kcol[c] < j) {
            ++c;
        }
        kpre[j] = m - c;
    }
    // tiles of the products: those with work by descending K range (the hardware deals
    // consecutive workgroups to the XCDs in turn), then, as a second list, the tiles wholly below
    // the band, which only write zeros and run on the side stream beside the sweep
    std::vector<int> tiles;
    int nwork = 0, nzero = 0;
    if (n >= 2) {
        const int ta = (n + 127) / 128, tb = (n + 63) / 64;
        std::vector<std::pair<int, int>> work; // (-K, packed tile)
        for (int a = 0; a < ta; ++a) {
            for (int b = 0; b < tb; ++b) {
                const int a0 = 128 * a, b0 = 64 * b;
                if (b0 > a0 + 128) {
                    continue;
                }
                const int kk = std::min(kpre[a0], b0 + 1 < n ? kpre[b0 + 1] : 0);
                work.push_back({-kk, (a << 16) | b});
            }
        }
        std::stable_sort(work.begin(), work.end(),
                         [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first < y.first; });
        for (const auto &w : work) {
            tiles.push_back(w.second);
        }
        nwork = (int)tiles.size();
        for (int a = 0; a < ta; ++a) {
            for (int b = 0; b < tb; ++b) {
                if (64 * b > 128 * a + 128) {
                    tiles.push_back((a << 16) | b);
                }
            }
        }
        nzero = (int)tiles.size() - nwork;
        ENSURE(ws.btiles, sizeof(int) * tiles.size());
        XTG_CHECK(hipMemcpyAsync(ws.btiles.p, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, st));
    }
    // beside the sweep, on the side stream (inputs only): Y = P V (the alpha / beta terms need
    // lambda^T P v) and the forward vectors with rows padded to npad (zero past n, 16-byte
    // aligned columns) for the outer products; then (zero_work) the zero tiles of dS and dP.
    // Both wait for the adjoint right-hand sides (ev_rhs on st), so that they meet the
    // latency-bound sweep rather than the bandwidth-bound passes before it; ev_y marks their end
    auto side_work = [&]() -> int {
        hipLaunchKernelGGL(k_to_rowmajor, dim3((m + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, ws.side, V, ldv, n, m,
                           ws.Vt.as<double>(), ldt); XTG_LAUNCH_CHECK("k_to_rowmajor");
        const int pr = prof_begin(ws.side);
        launch_gemm_f64(2, ws.side, n, m, n, 0, P, ldp, (const double *)nullptr, 0L,
                           (const double *)ws.Vt.as<double>(), ldt, ws.Y.as<double>(), ldt, (const int *)nullptr, 0,
                           (const double *)nullptr, (const int *)nullptr, (const double *)nullptr, 0, nullptr, nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<2>");
        prof_end(pr, ws.side, 3, 2.0 * n * n * (double)m, 8.0 * ((double)n * n + 2.0 * n * m));
        hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, ws.side, ws.bVp.as<double>(), npad * m); XTG_LAUNCH_CHECK("k_zero");
        hipLaunchKernelGGL(k_bwd_permute, dim3((n + 255) / 256, m), dim3(256), 0, ws.side, V, ldv, n, m, 1, 0, 0,
                           ws.bVp.as<double>(), npad); XTG_LAUNCH_CHECK("k_bwd_permute");
        return 0;
    };
    auto zero_work = [&]() -> int {
        XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_rhs, 0));
        if (nzero > 0) {
            // below-band tiles: zeros only, no operand is read
            hipLaunchKernelGGL(k_bwd_zero_tiles, dim3((unsigned)nzero, nbatch), dim3(256), 0, ws.side, n, dS, ldds, dP,
                               lddp, (const int *)ws.btiles.as<int>() + nwork, (long)n * ldds, (long)n * lddp);
            XTG_LAUNCH_CHECK("k_bwd_zero_tiles");
        }
        XTG_CHECK(hipEventRecord(ws.ev_y, ws.side));
        return 0;
    };
    hipLaunchKernelGGL(k_colnorms, dim3(n), dim3(256), 0, st, S, lds, P, ldp, n, 1, ws.norms.as<double>()); XTG_LAUNCH_CHECK("k_colnorms");
    hipLaunchKernelGGL(k_subdiag, dim3((n + 255) / 256), dim3(256), 0, st, S, lds, n, ws.norms.as<double>() + 2 * (long)n); XTG_LAUNCH_CHECK("k_subdiag");
    // queued before the host waits for the norms: the working pencil, the adjoint right-hand
    // sides and (FP64) the side-stream work
    std::vector<double> nrm;
    XTG_CHECK_RC(read_norms(ws, n, st, nrm, [&]() -> int {
        // the working pencil (reversed transposes) on the side stream beside the right-hand
        // sides; ev_z marks it, and the sweep waits for it
        XTG_CHECK(split_side_init(ws));
        XTG_CHECK(hipEventRecord(ws.ev_start, st));
        XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_start, 0));
        dim3 g32((n + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, ws.side, S, lds, n, ws.Sw.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, ws.side, P, ldp, n, ws.Pw.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        XTG_CHECK(hipEventRecord(ws.ev_z, ws.side));
        XTG_CHECK(hipMemcpyAsync(ws.kcol.p, kcol.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        // the right-hand sides: every entry of the rows < n and columns < M is written (zero
        // below each column's system), and nothing reads the padding columns of X
        const int pr = prof_begin(st);
        if (M >= 4096) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bwd_rhs<16>), dim3((M + 15) / 16), dim3(256), 0, st, V, ldv, G, ldg, n,
                               m, nbatch, ws.kcol.as<int>(), X, ldx, rhs_given ? 0 : 1); XTG_LAUNCH_CHECK("k_bwd_rhs");
        }
        else {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bwd_rhs<4>), dim3((M + 3) / 4), dim3(256), 0, st, V, ldv, G, ldg, n, m,
                               nbatch, ws.kcol.as<int>(), X, ldx, rhs_given ? 0 : 1); XTG_LAUNCH_CHECK("k_bwd_rhs");
        }
        prof_end(pr, st, 3, 0.0, 8.0 * 3.0 * n * M);
        XTG_CHECK(hipEventRecord(ws.ev_rhs, st));
        if (!split) {
            // the side work starts after the right-hand sides (bandwidth-bound, on the chain)
            // and runs beside the latency-bound sweep
            XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_rhs, 0));
            XTG_CHECK_RC(side_work());
            XTG_CHECK_RC(zero_work());
        }
        return 0;
    }));
    std::vector<double> sub(nrm.begin() + 2 * (long)n, nrm.end());
    double an = 0.0, bn = 0.0;
    for (int c = 0; c < n; ++c) {
        an = std::max(an, nrm[c]);
        bn = std::max(bn, nrm[n + c]);
    }
    const double safemin = std::numeric_limits<double>::min();
    const double ascale = 1.0 / std::max(an, safemin), bscale = 1.0 / std::max(bn, safemin);
    std::vector<double> subw(n, 0.0);
    for (int i = 1; i < n; ++i) {
        subw[i] = sub[n - i];
    }
    // per forward column: a, b, s = a / beta, b / a
    std::vector<double> coef(4 * (size_t)m);
    std::vector<ColDesc> desc(M);
    for (int c = 0; c < m; ++c) {
        const int k = kcol[c];
        double a, bR, bI;
        shift_real(ar[k], 0.0, be[k], ascale, bscale, a, bR, bI, false);
        coef[4 * c + 0] = a;
        coef[4 * c + 1] = bR;
        coef[4 * c + 2] = a / be[k];
        coef[4 * c + 3] = (a != 0.0) ? bR / a : 0.0;
        for (int b = 0; b < nbatch; ++b) {
            int q = (m - 1 - c) * nbatch + b;
            ColDesc d{};
            d.a = a;
            d.bR = bR;
            d.bI = 0.0;
            d.d = n;
            d.lo = n - k;
            d.partner = q;
            d.flags = INIT_NONE << 2;
            desc[q] = d;
        }
    }
    XTG_CHECK(hipMemcpyAsync(ws.coef.p, coef.data(), sizeof(double) * 4 * m, hipMemcpyHostToDevice, st));
    int pr = 0;
    if (split) {
        // digits of the working pencil on the side stream, after its reversed transposes, then
        // the side-stream work
        const long qld = split_qld(n), qsl = (long)n * qld;
        ENSURE(ws.rowexp, sizeof(int) * (size_t)n);
        ENSURE(ws.Sq, (size_t)SPLIT_DIGITS * qsl);
        ENSURE(ws.Pq, (size_t)SPLIT_DIGITS * qsl);
        std::vector<int> cfar;
        far_columns(n, subw, cfar);
        ENSURE(ws.cfar, sizeof(int) * (size_t)n);
        XTG_CHECK(hipMemcpyAsync(ws.cfar.p, cfar.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        XTG_CHECK(split_side_init(ws));
        XTG_CHECK(hipEventRecord(ws.ev_start, st));
        XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_start, 0));
        pr = prof_begin(ws.side);
        XTG_CHECK(split_rows(ws.side, n, n, (int)qld, ws.Sw.as<double>(), (long)n, ws.Pw.as<double>(), (long)n,
                             ws.rowexp.as<int>(), ws.Sq.as<signed char>(), ws.Pq.as<signed char>(), ws.cfar.as<int>()));
        prof_end(pr, ws.side, XTGEVC3_PROF_SPLIT_DIGITS, 0.0, 8.0 * 4.0 * n * (double)n + 2.0 * SPLIT_DIGITS * qsl);
        XTG_CHECK(hipEventRecord(ws.ev_pencil, ws.side));
        XTG_CHECK_RC(side_work());
        XTG_CHECK_RC(zero_work());
    }
    XTG_CHECK(hipStreamWaitEvent(st, ws.ev_z, 0)); // the working pencil
    // (no scale fix-up pass: k_bwd_lambda_t and k_bwd_diag read each row at its step's exponent)
    int rc = sweep_real(ws, n, ws.Sw.as<double>(), n, ws.Pw.as<double>(), n, subw, desc, true, ldx, st, prec,
                        split ? ws.ev_pencil : nullptr, false);
    if (rc) {
        return rc;
    }
    XTG_CHECK(hipStreamWaitEvent(st, ws.ev_y, 0));
    ENSURE(ws.bovf, sizeof(int) * (size_t)M);
    ENSURE(ws.bnfk, sizeof(int) * (size_t)nbatch * npad);
    hipLaunchKernelGGL(k_zero_int, dim3((M + 255) / 256), dim3(256), 0, st, ws.bovf.as<int>(), (long)M); XTG_LAUNCH_CHECK("k_zero_int");
    hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, st, ws.bnfk.as<int>(), (long)nbatch * npad, INT_MAX); XTG_LAUNCH_CHECK("k_fill_int");
    hipLaunchKernelGGL(k_bwd_lambda_t, dim3((M + 31) / 32, (int)((npad + 31) / 32)), dim3(32, 8), 0, st,
                       (const double *)X, ldx, n, m, nbatch, ws.kcol.as<int>(), (const int *)ws.rowstep.as<int>(),
                       (const int *)ws.blkexp.as<int>(), ws.LamS.as<double>(), npad, ws.bovf.as<int>(),
                       ws.bnfk.as<int>()); XTG_LAUNCH_CHECK("k_bwd_lambda_t");
    // the diagonal (alpha, beta) terms on the side stream, beside the outer products (they read
    // X and Y only); added into dS_kk, dP_kk after them
    ENSURE(ws.bdg, sizeof(double) * 2 * (size_t)M);
    XTG_CHECK(hipEventRecord(ws.ev_start, st));
    XTG_CHECK(hipStreamWaitEvent(ws.side, ws.ev_start, 0));
    hipLaunchKernelGGL(k_bwd_diag, dim3((M + 15) / 16), dim3(256), 0, ws.side, (const double *)X, ldx,
                       (const double *)ws.Y.as<double>(), ldt, n, m, nbatch, ws.kcol.as<int>(), ws.coef.as<double>(),
                       (const int *)ws.rowstep.as<int>(), (const int *)ws.blkexp.as<int>(),
                       ws.bdg.as<double>()); XTG_LAUNCH_CHECK("k_bwd_diag");
    XTG_CHECK(hipEventRecord(ws.ev_y, ws.side));
    double kwork = 0.0; // sum over the dS support of the exact K range of each entry
    for (int j = 0; j < n; ++j) {
        kwork += (double)j * kpre[j] + (j + 1 < n ? kpre[j + 1] : 0) + (j + 2 < n ? kpre[j + 2] : 0);
    }
    ENSURE(ws.bkpre, sizeof(int) * (size_t)n);
    XTG_CHECK(hipMemcpyAsync(ws.bkpre.p, kpre.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    {
        // dS_b = V diag(-a) Lam_b and dP_b = V diag(b) Lam_b in one pass over V and Lam_b (the
        // columns in reverse order, so that row j needs only the first kpre[j] of them)
        // s1 = sf[0, m), s2 = sf[ms, ms + m), each followed by zeros (the kernel reads whole K steps,
        // 16 bytes per lane: ms is even, so both start 16-byte aligned)
        const int ms = (m + 17) & ~1;
        std::vector<double> sf(2 * (size_t)ms, 0.0);
        for (int k = 0; k < m; ++k) {
            const int c = m - 1 - k;
            sf[k] = -coef[4 * c + 0];
            sf[ms + k] = coef[4 * c + 1];
        }
        ENSURE(ws.sfac, sizeof(double) * 2 * (size_t)ms);
        XTG_CHECK(hipMemcpyAsync(ws.sfac.p, sf.data(), sizeof(double) * 2 * ms, hipMemcpyHostToDevice, st));
        const double *Vrev = ws.bVp.as<double>() + (long)(m - 1) * npad;
        const long lrev = (long)(m - 1) * npad;
        pr = prof_begin(st);
        if (n < 2) {
            // lambda has no rows (k = 0): the products vanish, only the diagonal terms remain
            XTG_CHECK(hipMemset2DAsync(dS, sizeof(double) * ldds, 0, sizeof(double) * n, (size_t)nbatch * n, st));
            XTG_CHECK(hipMemset2DAsync(dP, sizeof(double) * lddp, 0, sizeof(double) * n, (size_t)nbatch * n, st));
        }
        else {
            // the tiles with work (the zero tiles ran on the side stream)
            const dim3 grid((unsigned)nwork, nbatch);
            hipLaunchKernelGGL(k_bwd_outer, grid, dim3(256), 0, st, n, m, Vrev, -npad,
                               (const double *)(ws.LamS.as<double>() + lrev), -npad,
                               (const double *)ws.sfac.as<double>(), (const double *)(ws.sfac.as<double>() + ms), dS,
                               ldds, dP, lddp, (const int *)ws.bkpre.as<int>(), (const int *)ws.bnfk.as<int>(),
                               (const int *)ws.btiles.as<int>(), (long)m * npad, npad, (long)n * ldds,
                               (long)n * lddp); XTG_LAUNCH_CHECK("k_bwd_outer");
        }
        prof_end(pr, st, 3, 4.0 * kwork * nbatch,
                 8.0 * nbatch * ((double)n * m + (double)m * n + 2.0 * (double)n * n));
    }
    XTG_CHECK(hipStreamWaitEvent(st, ws.ev_y, 0));
    hipLaunchKernelGGL(k_bwd_diag_add, dim3((M + 255) / 256), dim3(256), 0, st, (const double *)ws.bdg.as<double>(), n,
                       m, nbatch, ws.kcol.as<int>(), dS, ldds, dP, lddp); XTG_LAUNCH_CHECK("k_bwd_diag_add");
    XTG_CHECK(hipGetLastError());
    // info > 0: the number of (batch, column) adjoints whose lambda exceeds the double range;
    // the gradient entries that depend on them are NaN, every other entry is finite and exact
    std::vector<int> ovf(M);
    XTG_CHECK(hipMemcpyAsync(ovf.data(), ws.bovf.p, sizeof(int) * M, hipMemcpyDeviceToHost, st));
    XTG_CHECK(hipStreamSynchronize(st));
    int novf = 0;
    for (int q = 0; q < M; ++q) {
        novf += ovf[q] != 0;
    }
    *info = novf;
    return 0;
}

// ---------------------------------------------------------------------------------------
// Backward of selected real eigenvectors for either side, with or without the back-transform.
//   side 'R': V the right vectors; side 'L': the left vectors (y^H (beta S - alpha P) = 0).
//   Zb (optional): the back-transform matrix the vectors were formed with (Z for right, Q for
//   left vectors: V = normalise(Zb x)); then dZ (optional) receives dL/dZb for every batch.
// Left vectors are right vectors of the working pencil J S^T J, J P^T J (rows reversed; the
// eigenvalues in reverse order): the right backward runs there and dS = J dS_w^T J.  For
// back-transformed vectors v = Zb x / s (s the normalisation): with g'' = g - (g . v) sign(v_p) e_p,
// the adjoint right-hand side is Zb^T g'', the Schur-basis vector w = Zb^T v (= x / s, zero below
// its eigenvalue's row), dS and dP follow from w as for untransformed vectors, and
// dL/dZb = G'' W^T.
// ---------------------------------------------------------------------------------------
int real_backward_full_device(char side, int n, const double *S, long lds, const double *P, long ldp,
                              const double *ar, const double *ai, const double *be, const int *select,
                              const double *Zb, long ldz, const double *V, long ldv, const double *G, long ldg,
                              int nbatch, double *dS, long ldds, double *dP, long lddp, double *dZ, long lddz,
                              int *info, hipStream_t st)
{
    const bool left = side == 'L' || side == 'l';
    if (!left && Zb == nullptr) {
        return real_backward_device(n, S, lds, P, ldp, ar, ai, be, select, V, ldv, G, ldg, nbatch, dS, ldds, dP,
                                    lddp, info, st, XTGEVC3_PREC_FP64);
    }
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    RealWork &ws = g_ws[dev & 63];
    WsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    std::vector<Unit> units;
    const int m = real_units(select ? 'S' : 'A', select, n, ai, units);
    *info = 0;
    for (const Unit &u : units) {
        if (u.w != 1 || std::fabs(be[u.k]) <= std::numeric_limits<double>::min()) {
            *info = -1001;
            return 0;
        }
    }
    if (m == 0 || nbatch <= 0) {
        return 0;
    }
    const int M = m * nbatch;
    // working problem: the pencil, eigenvalues, selection and the operands in its coordinates
    const double *Sw = S, *Pw = P;
    long ldsw = lds, ldpw = ldp;
    std::vector<double> arw(ar, ar + n), aiw(ai, ai + n), bew(be, be + n);
    std::vector<int> selw;
    const int *selp = select;
    const double *Vw = V, *Gw = G;
    long ldvw = ldv, ldgw = ldg;
    const double *Zw = Zb;
    long ldzw = ldz;
    if (left) {
        ENSURE(ws.bfS, sizeof(double) * (size_t)n * n);
        ENSURE(ws.bfP, sizeof(double) * (size_t)n * n);
        dim3 g32((n + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, st, S, lds, n, ws.bfS.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, st, P, ldp, n, ws.bfP.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_reverse_transpose");
        Sw = ws.bfS.as<double>();
        Pw = ws.bfP.as<double>();
        ldsw = ldpw = n;
        for (int i = 0; i < n; ++i) {
            arw[i] = ar[n - 1 - i];
            aiw[i] = ai[n - 1 - i];
            bew[i] = be[n - 1 - i];
        }
        if (select != nullptr) {
            selw.assign(select, select + n);
            std::reverse(selw.begin(), selw.end());
            selp = selw.data();
        }
        // columns in reverse eigenvalue order; the rows too unless back-transformed (then the
        // rows are components in the original basis and Q J carries the reversal)
        ENSURE(ws.bfV, sizeof(double) * (size_t)n * m);
        ENSURE(ws.bfG, sizeof(double) * (size_t)n * M);
        const int rr = Zb == nullptr ? 1 : 0;
        hipLaunchKernelGGL(k_bwd_permute, dim3((n + 255) / 256, m), dim3(256), 0, st, V, ldv, n, m, 1, rr, 1,
                           ws.bfV.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_bwd_permute");
        hipLaunchKernelGGL(k_bwd_permute, dim3((n + 255) / 256, M), dim3(256), 0, st, G, ldg, n, m, nbatch, rr, 1,
                           ws.bfG.as<double>(), (long)n); XTG_LAUNCH_CHECK("k_bwd_permute");
        Vw = ws.bfV.as<double>();
        Gw = ws.bfG.as<double>();
        ldvw = ldgw = n;
        if (Zb != nullptr) {
            Zw = Zb + (long)(n - 1) * ldz; // Q J: columns reversed
            ldzw = -ldz;
        }
    }
    double *dSw = dS, *dPw = dP;
    long ldsg = ldds, ldpg = lddp;
    if (left) {
        ENSURE(ws.bfdS, sizeof(double) * (size_t)n * n * nbatch);
        ENSURE(ws.bfdP, sizeof(double) * (size_t)n * n * nbatch);
        dSw = ws.bfdS.as<double>();
        dPw = ws.bfdP.as<double>();
        ldsg = ldpg = n;
    }
    int rc = 0;
    if (Zw == nullptr) {
        rc = real_backward_core(ws, n, Sw, ldsw, Pw, ldpw, arw.data(), aiw.data(), bew.data(), selp, Vw, ldvw, Gw,
                                ldgw, nbatch, dSw, ldsg, dPw, ldpg, info, st, XTGEVC3_PREC_FP64, false);
    }
    else {
        // the eigenvalue row of each working column (ascending), for the truncation of w
        std::vector<Unit> uw;
        real_units(selp ? 'S' : 'A', selp, n, aiw.data(), uw);
        std::vector<int> kc(m);
        for (int c = 0; c < m; ++c) {
            kc[c] = uw[c].k;
        }
        ENSURE(ws.bfk, sizeof(int) * (size_t)m);
        XTG_CHECK(hipMemcpyAsync(ws.bfk.p, kc.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        const long ldt = ((long)m + 7) / 8 * 8, ldtM = ((long)M + 7) / 8 * 8;
        ENSURE(ws.bfGc, sizeof(double) * (size_t)n * M);
        ENSURE(ws.bfGt, sizeof(double) * (size_t)n * ldtM);
        ENSURE(ws.bfVt, sizeof(double) * (size_t)n * ldt);
        ENSURE(ws.bfZt, sizeof(double) * (size_t)n * n);
        ENSURE(ws.bfW, sizeof(double) * (size_t)n * m);
        ENSURE(ws.bfGh, sizeof(double) * (size_t)n * M);
        double *Gc = ws.bfGc.as<double>(), *Gt = ws.bfGt.as<double>(), *Vt = ws.bfVt.as<double>();
        double *Zt = ws.bfZt.as<double>(), *W = ws.bfW.as<double>(), *Gh = ws.bfGh.as<double>();
        // g'' (the cotangent of the unnormalised vector, in the output basis)
        hipLaunchKernelGGL(k_bwd_gcorr, dim3(M), dim3(64), 0, st, Vw, ldvw, Gw, ldgw, n, m, Gc, (long)n); XTG_LAUNCH_CHECK("k_bwd_gcorr");
        // row-major copies: V, G'' (n x m, n x M) and Zb (n x n, its rows)
        hipLaunchKernelGGL(k_to_rowmajor, dim3((m + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, st, Vw, ldvw, n, m, Vt, ldt); XTG_LAUNCH_CHECK("k_to_rowmajor");
        hipLaunchKernelGGL(k_to_rowmajor, dim3((M + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, st, (const double *)Gc, (long)n, n, M, Gt, ldtM); XTG_LAUNCH_CHECK("k_to_rowmajor");
        hipLaunchKernelGGL(k_to_rowmajor, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, st, Zw, ldzw, n, n, Zt, (long)n); XTG_LAUNCH_CHECK("k_to_rowmajor");
        // W = Zb^T V and Gh = Zb^T G'' (column-major, ld n): D[c][j] = sum_k V(k, c) Zb(k, j)
        launch_gemm_f64(2, st, m, n, n, 0, (const double *)Vt, ldt, (const double *)nullptr, 0L, (const double *)Zt,
                        (long)n, W, (long)n, (const int *)nullptr, 0, (const double *)nullptr, (const int *)nullptr,
                        (const double *)nullptr, 0, nullptr, nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<2>");
        launch_gemm_f64(2, st, M, n, n, 0, (const double *)Gt, ldtM, (const double *)nullptr, 0L, (const double *)Zt,
                        (long)n, Gh, (long)n, (const int *)nullptr, 0, (const double *)nullptr, (const int *)nullptr,
                        (const double *)nullptr, 0, nullptr, nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<2>");
        hipLaunchKernelGGL(k_bwd_trunc, dim3((n + 255) / 256, m), dim3(256), 0, st, W, (long)n, n, m,
                           (const int *)ws.bfk.as<int>()); XTG_LAUNCH_CHECK("k_bwd_trunc");
        rc = real_backward_core(ws, n, Sw, ldsw, Pw, ldpw, arw.data(), aiw.data(), bew.data(), selp, W, (long)n, Gh,
                                (long)n, nbatch, dSw, ldsg, dPw, ldpg, info, st, XTGEVC3_PREC_FP64, true);
        if (rc == 0 && dZ != nullptr) {
            // dL/dZb_b = G''_b W^T: D[j][i] = sum_c W(j, c) G''_b(i, c), written to dZ(i, j) (left:
            // to column n-1-j, as Zb J = Q J)
            double *Cz = left ? dZ + (long)(n - 1) * lddz : dZ;
            const long ldcz = left ? -lddz : lddz;
            launch_gemm_f64(2, st, n, n, m, 0, (const double *)W, (long)n, (const double *)nullptr, 0L,
                            (const double *)Gc, (long)n, Cz, ldcz, (const int *)nullptr, 0, (const double *)nullptr,
                            (const int *)nullptr, (const double *)nullptr, 0, nullptr, nullptr, nbatch, (long)m * n,
                            (long)n * lddz); XTG_LAUNCH_CHECK("k_gemm_f64<2>");
        }
    }
    if (rc != 0 || *info < 0) {
        return rc;
    }
    if (left) {
        // dS_b = J dSw_b^T J (the same map takes the working pencil back)
        dim3 g32((n + 31) / 32, (n + 31) / 32);
        for (int b = 0; b < nbatch; ++b) {
            hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, st, (const double *)dSw + (long)b * n * n,
                               (long)n, n, dS + (long)b * n * ldds, ldds); XTG_LAUNCH_CHECK("k_reverse_transpose");
            hipLaunchKernelGGL(k_reverse_transpose, g32, dim3(32, 8), 0, st, (const double *)dPw + (long)b * n * n,
                               (long)n, n, dP + (long)b * n * lddp, lddp); XTG_LAUNCH_CHECK("k_reverse_transpose");
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Eigenvalue condition numbers s_k (LAPACK dtgsna, JOB = 'E', HOWMNY = 'S' / 'A') from the
// left and right eigenvectors of the pencil (S, P) as dtgevc3 returns them (device arrays, the
// same basis for S, P and the vectors): s_k = |(u^H S v, u^H P v)| / (|u| |v|), a complex pair
// sharing one value over its two columns.  Two FP64 products (S V, P V) and a column reduction
// in fixed order.  s: device array of m values.
// ---------------------------------------------------------------------------------------
int real_tgsna_device(char howmny, const int *select, int n, const double *S, long lds, const double *P, long ldp,
                      const double *ai, const double *VL, long ldvl, const double *VR, long ldvr, int m, double *s,
                      hipStream_t st)
{
    int dev = 0;
    XTG_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    RealWork &ws = g_ws[dev & 63];
    WsOrder order{&ws.ev_last, st};
    XTG_CHECK(order.begin());
    std::vector<Unit> units;
    const int mc = real_units(howmny, select, n, ai, units);
    if (mc != m) {
        set_error("m = " + std::to_string(m) + " does not match the " + std::to_string(mc) +
                  " columns the selection describes");
        return -1;
    }
    if (m == 0) {
        return 0;
    }
    const long ldt = ((long)m + 7) / 8 * 8;
    ENSURE(ws.Vt, sizeof(double) * (size_t)n * ldt);
    ENSURE(ws.Ut, sizeof(double) * (size_t)n * ldt);
    ENSURE(ws.Y, sizeof(double) * (size_t)n * ldt);
    ENSURE(ws.YB, sizeof(double) * (size_t)n * ldt);
    const int nu = (int)units.size();
    std::vector<int> ucol(nu), uw(nu);
    for (int q = 0; q < nu; ++q) {
        ucol[q] = units[q].outc;
        uw[q] = units[q].w;
    }
    ENSURE(ws.tcol, sizeof(int) * (size_t)nu);
    ENSURE(ws.tw, sizeof(int) * (size_t)nu);
    XTG_CHECK(hipMemcpyAsync(ws.tcol.p, ucol.data(), sizeof(int) * nu, hipMemcpyHostToDevice, st));
    XTG_CHECK(hipMemcpyAsync(ws.tw.p, uw.data(), sizeof(int) * nu, hipMemcpyHostToDevice, st));
    const dim3 gt((m + 31) / 32, (n + 31) / 32);
    hipLaunchKernelGGL(k_to_rowmajor, gt, dim3(32, 8), 0, st, VR, ldvr, n, m, ws.Vt.as<double>(), ldt); XTG_LAUNCH_CHECK("k_to_rowmajor");
    hipLaunchKernelGGL(k_to_rowmajor, gt, dim3(32, 8), 0, st, VL, ldvl, n, m, ws.Ut.as<double>(), ldt); XTG_LAUNCH_CHECK("k_to_rowmajor");
    for (int part = 0; part < 2; ++part) {
        launch_gemm_f64(2, st, n, m, n, 0, part == 0 ? S : P, part == 0 ? lds : ldp, (const double *)nullptr, 0L,
                        (const double *)ws.Vt.as<double>(), ldt, part == 0 ? ws.Y.as<double>() : ws.YB.as<double>(),
                        ldt, (const int *)nullptr, 0, (const double *)nullptr, (const int *)nullptr,
                        (const double *)nullptr, 0, nullptr, nullptr); XTG_LAUNCH_CHECK("k_gemm_f64<2>");
    }
    constexpr int ROWS = 128;
    const int nchunk = (n + ROWS - 1) / ROWS;
    ENSURE(ws.part, sizeof(double) * 12 * (size_t)nchunk * nu);
    hipLaunchKernelGGL(k_tgsna_part, dim3((nu + 255) / 256, nchunk), dim3(256), 0, st,
                       (const double *)ws.Ut.as<double>(), (const double *)ws.Vt.as<double>(),
                       (const double *)ws.Y.as<double>(), (const double *)ws.YB.as<double>(), ldt, n, nu,
                       (const int *)ws.tcol.as<int>(), (const int *)ws.tw.as<int>(), ROWS, ws.part.as<double>()); XTG_LAUNCH_CHECK("k_tgsna_part");
    hipLaunchKernelGGL(k_tgsna_fin, dim3((nu + 255) / 256), dim3(256), 0, st, (const double *)ws.part.as<double>(),
                       nchunk, nu, (const int *)ws.tcol.as<int>(), (const int *)ws.tw.as<int>(), s); XTG_LAUNCH_CHECK("k_tgsna_fin");
    XTG_CHECK(hipGetLastError());
    return 0;
}

} // namespace xtg

