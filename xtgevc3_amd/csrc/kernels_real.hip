// gfx950 kernels of the real (FP64) xTGEVC3 path.
//
//   k_solve_real    -- one block-row step of the right-looking blocked back substitution:
//                      per-column shifted quasi-triangular solve of a diagonal block,
//                      one lane per eigenvector column, S/P block staged in LDS,
//                      complex pairs on two lanes, per-column power-of-two overflow scaling.
//                      (replaces dlalsr/dlalsl + dlauhs/dlau2s/dlalhs/dlal2s,
//                       reference src/dtgevc3.cpp:132-1073)
//   k_gemm_f64      -- FP64 MFMA (v_mfma_f64_16x16x4) GEMM used for the off-diagonal RHS
//                      update (dtgevc3.cpp:1432-1433, :1624-1625) and for the back-transform
//                      (dtgevc3.cpp:1446, :1638).
//   k_fixup_scale   -- applies the per-column scale history to the solved blocks.
//   k_copy_out      -- X (row-major working layout) -> V (column-major), zero fill.
//   k_normalize     -- LAPACK-style normalisation of output vectors.
//   k_reverse_transpose -- builds the working pencil J S^T J for left eigenvectors.
//   k_colnorms      -- 1-norm column sums of S and P over the upper Hessenberg part.
#include <hip/hip_runtime.h>
#include <cfloat>
#include <climits>
#include <cmath>
#include <type_traits>

#include "common.h"

namespace xtg {

struct cpx {
    double r, i;
};

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ cpx cmul(cpx a, cpx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ double abs1(cpx a) { return fabs(a.r) + fabs(a.i); }

// Reciprocal of a complex number, Smith's ordering (no overflow of |c|^2); c != 0.
__device__ __forceinline__ cpx crecip(cpx c)
{
    const bool sw = fabs(c.i) > fabs(c.r);
    const double big = sw ? c.i : c.r, sml = sw ? c.r : c.i;
    const double t = sml / big;
    const double rd = 1.0 / (big + sml * t);
    const double tr = t * rd;
    return sw ? cpx{tr, -rd} : cpx{rd, -tr};
}

constexpr double kEps = 2.220446049250313e-16;     // DBL_EPSILON
constexpr double kSafeMin = 2.2250738585072014e-308; // DBL_MIN
constexpr double kScaleTrigger = 0x1p958;          // a solution entry above this triggers a rescale
constexpr int kScaleTarget = 446;                  // rescaled entries stay below 2^446 (2^512 headroom)

// Exponent e >= 0 with 2^-e * num * fac <= 2^kScaleTarget (abs1 bounds, no product formed).
__device__ __forceinline__ int rescale_exponent(double num, double fac)
{
    if (!(num > 0.0) || !(fac > 0.0)) {
        return 0;
    }
    const int e = ilogb(num) + ilogb(fac) + 2 - kScaleTarget;
    return e > 0 ? e : 0;
}

// ---------------------------------------------------------------------------------------
// Row broadcast: every lane of each 16-lane row receives lane `src` of its row (DPP
// row_newbcast, gfx90a and later).
// ---------------------------------------------------------------------------------------
constexpr int QL = 16; // lanes per eigenvalue unit

template <int SRC> __device__ __forceinline__ int row_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + SRC, 0xF, 0xF, false);
}

template <int SRC> __device__ __forceinline__ double row_bcast_d(double v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + SRC, 0xF, 0xF, false); // one v_mov_b64_dpp
}

#define XTG_RB_CASES(F)                                                                                          \
    switch (src) {                                                                                               \
    case 0: return F<0>(v);                                                                                      \
    case 1: return F<1>(v);                                                                                      \
    case 2: return F<2>(v);                                                                                      \
    case 3: return F<3>(v);                                                                                      \
    case 4: return F<4>(v);                                                                                      \
    case 5: return F<5>(v);                                                                                      \
    case 6: return F<6>(v);                                                                                      \
    case 7: return F<7>(v);                                                                                      \
    case 8: return F<8>(v);                                                                                      \
    case 9: return F<9>(v);                                                                                      \
    case 10: return F<10>(v);                                                                                    \
    case 11: return F<11>(v);                                                                                    \
    case 12: return F<12>(v);                                                                                    \
    case 13: return F<13>(v);                                                                                    \
    case 14: return F<14>(v);                                                                                    \
    default: return F<15>(v);                                                                                    \
    }

__device__ __forceinline__ double qbd(double v, int src) { XTG_RB_CASES(row_bcast_d) }
__device__ __forceinline__ int qbi(int v, int src) { XTG_RB_CASES(row_bcast_i) }
#undef XTG_RB_CASES

__device__ __forceinline__ cpx qbc(cpx v, int src) { return {qbd(v.r, src), qbd(v.i, src)}; }

// Packed LDS layout of a staged diagonal block: column c keeps the 16-row blocks at and above its
// diagonal, rows [0, 16 (c / 16 + 1)), at sp[sp_col(c) + row]; the subdiagonal entry (c + 1, c)
// of every column is kept apart.  2640 {S,P} entries for 65 columns (42 KB).
__host__ __device__ constexpr int sp_col(int c) { return 16 * ((c >> 4) + 1) * (8 * (c >> 4) + (c & 15)); }
static_assert(sp_col(65) + 65 <= PK_BLK, "block image");

// Column c and row l of packed LDS item it = sp_col(c) + l (c < 80): a 16-column block qb
// starts at 128 qb (qb + 1) and holds 16 (qb + 1) rows per column.
__device__ __forceinline__ void sp_decode(int it, int &c, int &l)
{
    const int qb = (it >= 256) + (it >= 768) + (it >= 1536) + (it >= 2560);
    const int off = it - 128 * qb * (qb + 1);
    // off / (16 (qb + 1)) for off < 256 (qb + 1): (off / 16) * ceil(2^16 / (qb + 1)) / 2^16
    const int mg = qb == 0 ? 65536 : (qb == 1 ? 32768 : (qb == 2 ? 21846 : (qb == 3 ? 16384 : 13108)));
    const int cc = ((off >> 4) * mg) >> 16;
    l = off - 16 * (qb + 1) * cc;
    c = 16 * qb + cc;
}

// Rotations inside 16-lane rows (DPP row_ror): ror<1> gives lane j the value of lane j - 1
// (lane 0 that of lane 15), ror<15> the value of lane j + 1 (lane 15 that of lane 0).
template <int K> __device__ __forceinline__ double row_ror_d(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x120 + K, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x120 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int K> __device__ __forceinline__ double ror(double v) { return row_ror_d<K>(v); }
template <int K> __device__ __forceinline__ cpx ror(cpx v) { return {row_ror_d<K>(v.r), row_ror_d<K>(v.i)}; }
template <int K> __device__ __forceinline__ double2 ror(double2 v)
{
    return make_double2(row_ror_d<K>(v.x), row_ror_d<K>(v.y));
}

// Arithmetic of a unit's vector entries: double for waves of real eigenvalues only, cpx when
// the wave holds a complex pair (real units then carry a zero imaginary part).
template <class V> struct Ar;
template <> struct Ar<double> {
    static __device__ __forceinline__ double mk(double r, double) { return r; }
    static __device__ __forceinline__ double re(double v) { return v; }
    static __device__ __forceinline__ double im(double) { return 0.0; }
    static __device__ __forceinline__ double mul(double x, double y) { return x * y; }
    static __device__ __forceinline__ double sub(double x, double y) { return x - y; }
    static __device__ __forceinline__ double recip(double c) { return 1.0 / c; }
    static __device__ __forceinline__ double abs1(double x) { return fabs(x); }
    static __device__ __forceinline__ double amax(double x) { return fabs(x); }
    static __device__ __forceinline__ double bc(double v, int src) { return qbd(v, src); }
    static __device__ __forceinline__ double scal(double v, double f) { return v * f; }
    static __device__ __forceinline__ double fmar(double s, double x, double acc) { return fma(s, x, acc); }
    // shifted entry a*S - b*P (b real for these waves)
    static __device__ __forceinline__ double shift(double a, double bR, double, double2 e) { return a * e.x - bR * e.y; }
    static __device__ __forceinline__ double bmul(double bR, double, double x) { return bR * x; }
    // v - S_lr (a x) + P_lr (b x)
    static __device__ __forceinline__ double upd(double v, double2 e, double ax, double bx)
    {
        return fma(e.y, bx, fma(-e.x, ax, v));
    }
};
template <> struct Ar<cpx> {
    static __device__ __forceinline__ cpx mk(double r, double i) { return {r, i}; }
    static __device__ __forceinline__ double re(cpx v) { return v.r; }
    static __device__ __forceinline__ double im(cpx v) { return v.i; }
    static __device__ __forceinline__ cpx mul(cpx x, cpx y) { return cmul(x, y); }
    static __device__ __forceinline__ cpx sub(cpx x, cpx y) { return csub(x, y); }
    static __device__ __forceinline__ cpx recip(cpx c) { return crecip(c); }
    static __device__ __forceinline__ double abs1(cpx x) { return xtg::abs1(x); }
    static __device__ __forceinline__ double amax(cpx x) { return fmax(fabs(x.r), fabs(x.i)); }
    static __device__ __forceinline__ cpx bc(cpx v, int src) { return qbc(v, src); }
    static __device__ __forceinline__ cpx scal(cpx v, double f) { return {v.r * f, v.i * f}; }
    static __device__ __forceinline__ cpx fmar(double s, cpx x, cpx acc) { return {fma(s, x.r, acc.r), fma(s, x.i, acc.i)}; }
    static __device__ __forceinline__ cpx shift(double a, double bR, double bI, double2 e)
    {
        return {a * e.x - bR * e.y, -bI * e.y};
    }
    static __device__ __forceinline__ cpx bmul(double bR, double bI, cpx x)
    {
        return {bR * x.r - bI * x.i, bR * x.i + bI * x.r};
    }
    static __device__ __forceinline__ cpx upd(cpx v, double2 e, cpx ax, cpx bx)
    {
        return {fma(e.y, bx.r, fma(-e.x, ax.r, v.r)), fma(e.y, bx.i, fma(-e.x, ax.i, v.i))};
    }
};

// Element-wise select (a ?: on structs would select addresses and keep arrays of them out of
// registers).
__device__ __forceinline__ double vsel(bool c, double x, double y) { return c ? x : y; }
__device__ __forceinline__ cpx vsel(bool c, cpx x, cpx y) { return {c ? x.r : y.r, c ? x.i : y.i}; }
__device__ __forceinline__ double2 vsel(bool c, double2 x, double2 y) { return make_double2(c ? x.x : y.x, c ? x.y : y.y); }

// Per-lane description of one unit in one block step (see k_solve_real).
struct UnitStep {
    const double2 *sp;  // staged {S,P} block (packed columns, sp_col), LDS
    const double2 *spd; // its subdiagonal {S,P}_{c+1,c}, LDS
    int nb, j, dl, lol, init;
    unsigned long long bumps;
    double a, bR, bI, bn1;
};

// Right-hand side of one unit's rows in a block (real, imaginary part per row slot).
template <int QR> struct RawRHS {
    double2 v[QR];
};

// Writes of one unit after its block sweep: X rows and operand rows of the block, the group
// bookkeeping of W and the column scale state (see k_solve_real).
struct UnitOut {
    bool act, pair;
    int ucols, col, i0;
    int ce0;     // colexp of the unit's columns at entry (loaded at kernel start)
    double sgp;  // sig_grp of the lane's column at entry
    double *X;
    long ldx;
    double *W;
    long ldw;
    double *sig;
    int *colexp, *blkexp, *flag;
    int w_s, w_p;
    double *sig_grp;
    int first_in_group;
    GroupInfo grp;
    MixedExp mx;
    SplitOut sq;
    // fused near update (wl == nullptr: none): the block's operand rows of the unit's columns
    // into LDS, wl[k * WL_LD + slot] (S part k < ks, P part ks + k), and the column scale of
    // each slot into wsig (pair units use slots wslot and wslot + 1)
    double *wl;
    double *wsig;
    int wslot, ks;
};
constexpr int WL_LD = 32; // LDS pitch of the near update's operand (columns of a workgroup)

// One unit's block step: factor the lane's rows, sweep the block bottom-up starting from the
// right-hand side `raw`, then write the block's solution, operands and scale state.
//
// Rows are handled as positions p = 16 i + j (row slot i of lane j).  The factor gives every
// position a pivot reciprocal f and, at the lower row of a 2x2 bump of S, an elimination
// multiplier g, so that each row step is the same: t = v_p - g v_{p-1}, x_p = t f.  A pivoted
// bump exchanges the equations of its two rows: the position keeps its unknown, takes the other
// row's right-hand side and updates with the other row's {S,P} entries (eqd = +-1).  The sweep
// runs slot by slot: inside a slot every row step updates the slot's own rows; the slots above
// take the slot's rows 1..15 in one batched update before its row 0 (whose 2x2 bump partner
// may be the slot above's row 15), then row 0 like a single row step.
template <class V, int QR, int MAXB>
__device__ __forceinline__ void unit_block(const UnitStep u, const RawRHS<QR> raw, const UnitOut o)
{
    typedef Ar<V> A;
    const double2 *sp = u.sp;
    const double2 *spd = u.spd;
    const int nb = u.nb, j = u.j, dl = u.dl, lol = u.lol;
    const double a = u.a, bR = u.bR, bI = u.bI, bn1 = u.bn1;
    const unsigned long long bumps = u.bumps;

    // ---- factor: f, g and the equation offset of every position
    V f[QR], g[QR];
    int eqd[QR];
    bool anysw = false;
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = QL * i + j;
        f[i] = A::mk(0.0, 0.0);
        g[i] = A::mk(0.0, 0.0);
        eqd[i] = 0;
        if (r < nb) {
            const bool lower = r >= 1 && ((bumps >> (r - 1)) & 1ull);
            const bool upper = r + 1 < nb && ((bumps >> r) & 1ull);
            if (!lower && !upper) {
                const double2 ee = sp[sp_col(r) + r];
                V c = A::shift(a, bR, bI, ee);
                const double dmin = fmax(kEps * (fabs(a) * fabs(ee.x) + bn1 * fabs(ee.y)), kSafeMin);
                if (A::abs1(c) < dmin) {
                    c = A::mk(dmin, 0.0);
                }
                const bool solve = r < dl && r >= lol;
                f[i] = vsel(solve, A::recip(c), A::mk(r == dl ? 1.0 : 0.0, 0.0));
            }
            else {
                const int r1 = lower ? r - 1 : r, r2 = r1 + 1;
                const double2 e11 = sp[sp_col(r1) + r1], e12 = sp[sp_col(r2) + r1];
                const double2 e21 = spd[r1], e22 = sp[sp_col(r2) + r2];
                V c11 = A::shift(a, bR, bI, e11), c12 = A::shift(a, bR, bI, e12);
                V c21 = A::shift(a, bR, bI, e21), c22 = A::shift(a, bR, bI, e22);
                const double dm1 = fmax(kEps * (fabs(a) * (fabs(e11.x) + fabs(e21.x)) + bn1 * fabs(e11.y)), kSafeMin);
                const double dm2 = fmax(kEps * (fabs(a) * (fabs(e12.x) + fabs(e22.x)) + bn1 * fabs(e22.y)), kSafeMin);
                const bool s1 = r1 < dl && r1 >= lol, s2 = r2 < dl && r2 >= lol;
                bool sw = false;
                V mlt = A::mk(0.0, 0.0), iu = A::mk(0.0, 0.0), ip = A::mk(0.0, 0.0);
                if (r1 == dl) {
                    // the unit's own 2x2 block: both entries preset
                    iu = A::mk(1.0, 0.0);
                    ip = A::mk(1.0, 0.0);
                }
                else if (s1) {
                    // pivoted LU of [c11 c12; c21 c22]: pivot row rp = sw ? r2 : r1
                    sw = A::abs1(c21) > A::abs1(c11);
                    V pv = vsel(sw, c21, c11), sv = vsel(sw, c11, c21), tv = vsel(sw, c12, c22);
                    const V q = vsel(sw, c22, c12);
                    if (A::abs1(pv) < dm1) {
                        pv = A::mk(dm1, 0.0);
                    }
                    ip = A::recip(pv);
                    mlt = A::mul(sv, ip);
                    V uv = A::sub(tv, A::mul(mlt, q));
                    if (A::abs1(uv) < dm2) {
                        uv = A::mk(dm2, 0.0);
                    }
                    iu = A::recip(uv);
                }
                else if (s2) {
                    // upper row outside the system (zero): x2 = rhs2 / c22
                    if (A::abs1(c22) < dm2) {
                        c22 = A::mk(dm2, 0.0);
                    }
                    iu = A::recip(c22);
                }
                f[i] = vsel(lower, iu, ip);
                g[i] = vsel(lower, mlt, A::mk(0.0, 0.0));
                eqd[i] = sw ? (lower ? -1 : 1) : 0;
                anysw = anysw || sw;
            }
        }
    }

    // ---- right-hand sides by equation row: pivoted bumps exchange their two rows' values
    double2 rv[QR];
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        rv[i] = raw.v[i];
    }
    if (__any(anysw)) {
        double2 up[QR], dn[QR]; // values of lane j + 1 (ror 15) and lane j - 1 (ror 1)
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            up[i] = ror<15>(raw.v[i]);
            dn[i] = ror<1>(raw.v[i]);
        }
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            const double2 zero = make_double2(0.0, 0.0);
            const double2 un = i + 1 < QR ? up[i + 1 < QR ? i + 1 : i] : zero; // slot below, lane 0
            const double2 dp = i > 0 ? dn[i > 0 ? i - 1 : 0] : zero;            // slot above, lane 15
            const double2 vu = vsel(j == QL - 1, un, up[i]);
            const double2 vd = vsel(j == 0, dp, dn[i]);
            rv[i] = vsel(eqd[i] > 0, vu, vsel(eqd[i] < 0, vd, rv[i]));
        }
    }
    V vr[QR];
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        vr[i] = A::mk(rv[i].x, rv[i].y);
    }

    // starting vector on the unit's own block
    if (dl >= 0 && dl < nb) {
        V x0 = A::mk(0.0, 0.0), x1v = A::mk(0.0, 0.0);
        if (u.init == INIT_UNIT) {
            x0 = A::mk(1.0, 0.0);
        }
        else if (u.init == INIT_PAIR_ROW1 && dl + 1 < nb) {
            const double2 e11 = sp[sp_col(dl) + dl], e12 = sp[sp_col(dl + 1) + dl];
            x0 = A::mk(-(a * e12.x - bR * e12.y), bI * e12.y);
            x1v = A::mk(a * e11.x - bR * e11.y, -bI * e11.y);
        }
        else if (u.init == INIT_PAIR_ROW2 && dl + 1 < nb) {
            const double2 e22 = sp[sp_col(dl + 1) + dl + 1], e21 = spd[dl];
            x0 = A::mk(a * e22.x - bR * e22.y, -bI * e22.y);
            x1v = A::mk(-a * e21.x, 0.0);
        }
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            const int r = QL * i + j;
            vr[i] = vsel(r == dl, x0, vsel(r == dl + 1 && u.init != INIT_UNIT, x1v, vr[i]));
        }
    }

    // ---- sweep
    int my_exp = 0;
#pragma clang loop unroll(full)
    for (int i = QR - 1; i >= 0; --i) {
        if (QL * i >= nb) {
            continue;
        }
        // the slot's own coefficients: {S,P} of the lane's equation row at column 16 i + m,
        // zero where the row takes no update from the column (m <= j: the lane's own and lower
        // positions; the unit's preset row dl from column dl + 1)
        const int pos = QL * i + j;
        const int eqr = pos + eqd[i];
        double2 dc[QL];
#pragma clang loop unroll(full)
        for (int m = 0; m < QL; ++m) {
            dc[m] = make_double2(0.0, 0.0);
            if (QL * i + m < MAXB) {
                const int c = QL * i + m;
                const double2 e = sp[sp_col(c) + eqr];
                dc[m] = vsel(m > j && !(pos == dl && c == dl + 1), e, dc[m]);
            }
            // materialise the masked entries here, once per slot (the masks would otherwise
            // be re-evaluated in every row step that uses them)
            asm volatile("" : "+v"(dc[m].x), "+v"(dc[m].y));
        }
        // the slot's solution entries (broadcast) times the shift coefficients, a x and b x, for
        // the batched update: the products with S and P are formed from the scaled entries, as
        // in a row step (S x or P x alone can exceed the double range on graded pencils where
        // S (a x) and P (b x) do not)
        V xa[QL], xb[QL];
        int xe[QL]; // rescale exponent of the slot at the time each was stored
        int es = 0;
#pragma clang loop unroll(full)
        for (int m = 0; m < QL; ++m) {
            xa[m] = A::mk(0.0, 0.0);
            xb[m] = A::mk(0.0, 0.0);
            xe[m] = 0;
        }
        // the next row's owner lane keeps its updated value in tn (the select that stores a
        // row's solution in its owner lane then stays off the dependent chain)
        V tn = A::mk(0.0, 0.0);
#pragma clang loop unroll(full)
        for (int jr = QL - 1; jr >= 0; --jr) {
            const int r = QL * i + jr;
            // rows nb <= r < MAXB run too: their factors, right-hand sides and {S,P} entries are
            // zero, so they solve to zero and change nothing (no per-row test on nb)
            if (r >= MAXB) {
                continue;
            }
            if (jr == 0 && i > 0 && r + 1 < nb) {
                // slots above: rows 1..15 of this slot at once,
                // v -= a (S x) - b (P x) over the slot's columns
                if (__any(es != 0)) {
#pragma clang loop unroll(full)
                    for (int m = 1; m < QL; ++m) {
                        const double sc = ldexp(1.0, xe[m] - es);
                        xa[m] = A::scal(xa[m], sc);
                        xb[m] = A::scal(xb[m], sc);
                    }
                }
#pragma clang loop unroll(full)
                for (int k = 0; k < QR; ++k) {
                    if (k < i) {
                        const int eqk = QL * k + j + eqd[k];
                        V as = A::mk(0.0, 0.0), ap = A::mk(0.0, 0.0);
#pragma clang loop unroll(full)
                        for (int m = 1; m < QL; ++m) {
                            const double2 e = sp[sp_col(QL * i + m) + eqk];
                            as = A::fmar(e.x, xa[m], as);
                            ap = A::fmar(e.y, xb[m], ap);
                        }
                        vr[k] = A::upd(vr[k], make_double2(1.0, 1.0), as, ap);
                    }
                }
            }
            V t = (jr < QL - 1 && r + 1 < MAXB) ? tn : vr[i];
            if (__builtin_expect(r >= 1 && ((bumps >> (r - 1)) & 1ull), 0)) {
                // lower row of a 2x2 bump: eliminate with the upper row's current value
                const V prev = (jr > 0 || i == 0) ? ror<1>(vr[i]) : ror<1>(vr[i > 0 ? i - 1 : 0]);
                t = A::sub(t, A::mul(g[i], prev));
            }
            const V tf = A::mul(t, f[i]);
            V x = A::bc(tf, jr);
            // tested on the owner lane before the broadcast (every lane of the row receives its value)
            const bool bad = j == jr && !(A::abs1(tf) <= kScaleTrigger);
            if (__builtin_expect(__any(bad), 0)) {
                const int eo = bad ? rescale_exponent(A::abs1(t), A::abs1(f[i])) : 0;
                const int e = qbi(eo, jr);
                const double sc = ldexp(1.0, -e);
#pragma clang loop unroll(full)
                for (int k = 0; k < QR; ++k) {
                    vr[k] = A::scal(vr[k], sc);
                }
                t = A::scal(t, sc);
                my_exp -= e;
                es += e;
                x = A::bc(A::mul(t, f[i]), jr);
            }
            const V ax = A::scal(x, a), bx = A::bmul(bR, bI, x);
            xa[jr] = ax;
            xb[jr] = bx;
            xe[jr] = es;
            tn = A::upd(vr[i], dc[jr], ax, bx);
            vr[i] = vsel(j == jr, x, tn);
            if (jr == 0) {
                // row 16 i into the slots above (not into the preset row dl from column dl + 1)
#pragma clang loop unroll(full)
                for (int k = 0; k < QR; ++k) {
                    if (k < i) {
                        const int pk = QL * k + j;
                        const double2 e = sp[sp_col(r) + pk + eqd[k]];
                        const double2 ek = vsel(pk == dl && r == dl + 1, make_double2(0.0, 0.0), e);
                        vr[k] = A::upd(vr[k], ek, ax, bx);
                    }
                }
            }
        }
    }

    // ---- writes ----
    const UnitStep &us = u;
    const bool act = o.act, pair = o.pair;
    const int ucols = o.ucols, col = o.col, i0 = o.i0;
    double *__restrict__ X = o.X;
    double *__restrict__ W = o.W;
    double *__restrict__ sig = o.sig;
    int *__restrict__ colexp = o.colexp;
    int *__restrict__ blkexp = o.blkexp;
    int *__restrict__ flag = o.flag;
    double *__restrict__ sig_grp = o.sig_grp;
    const long ldx = o.ldx, ldw = o.ldw;
    const int w_s = o.w_s, w_p = o.w_p, first_in_group = o.first_in_group;
    const GroupInfo &grp = o.grp;
    const MixedExp &mx = o.mx;
    // X and W of this block: rows w_s + r (S part: -a x) and w_p + r (P part: b x)
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = QL * i + j;
        if (act && r < nb) {
            const V v = vr[i];
            const V bx = A::bmul(bR, bI, v);
            double *xr = X + (long)(i0 + r) * ldx + col;
            double *ws = W + (long)(w_s + r) * ldw + col;
            double *wp = W + (long)(w_p + r) * ldw + col;
            xr[0] = A::re(v);
            ws[0] = -a * A::re(v);
            wp[0] = A::re(bx);
            if (pair) {
                xr[1] = A::im(v);
                ws[1] = -a * A::im(v);
                wp[1] = A::im(bx);
            }
        }
    }
    if (o.wl != nullptr) {
        // the same operand values into LDS for the fused near update; rows [nb, ks) and
        // inactive units are zero
        constexpr bool CW = std::is_same<V, cpx>::value;
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            const int r = QL * i + j;
            if (r < o.ks) {
                const bool live = act && r < nb;
                const V v = vr[i];
                const V bx = A::bmul(bR, bI, v);
                double *w0 = o.wl + r * WL_LD + o.wslot;
                double *w1 = o.wl + (o.ks + r) * WL_LD + o.wslot;
                w0[0] = live ? -a * A::re(v) : 0.0;
                w1[0] = live ? A::re(bx) : 0.0;
                if (CW) {
                    w0[1] = live ? -a * A::im(v) : 0.0;
                    w1[1] = live ? A::im(bx) : 0.0;
                }
            }
        }
        if (j == 0) {
            const double sg = ldexp(1.0, my_exp);
            o.wsig[o.wslot] = sg;
            if (CW) {
                o.wsig[o.wslot + 1] = sg;
            }
        }
    }
    const bool starts_here = dl >= 0 && dl < nb;
    if (act) {
        const int ce = o.ce0 + my_exp; // pair columns share their exponents
        if (starts_here && !first_in_group) {
            for (int k = w_s + nb + j; k < grp.wr; k += QL) {
                double *w0 = W + (long)k * ldw + col;
                double *w1 = W + (long)(grp.wr + k) * ldw + col;
                w0[0] = 0.0;
                w1[0] = 0.0;
                if (pair) {
                    w0[1] = 0.0;
                    w1[1] = 0.0;
                }
            }
        }
        if (grp.last && !starts_here) {
            // lower blocks' operand rows to the final scale: all loads of a block are issued
            // before its stores (a load-store chain per element serialised the round trips)
            constexpr int KR = (MAXB + QL - 1) / QL;
#pragma clang loop unroll(full)
            for (int jb = 1; jb < 4; ++jb) {
                const int k0 = grp.off[jb], k1 = grp.off[jb + 1];
                if (k0 < k1) {
                    const int e = ce - blkexp[(long)(grp.step[jb] - grp.step[0]) * ldx + col];
                    if (e != 0) {
                        double v[KR][4];
#pragma clang loop unroll(full)
                        for (int t = 0; t < KR; ++t) {
                            const int k = k0 + j + QL * t;
                            if (k < k1) {
                                const double *w0 = W + (long)k * ldw + col;
                                const double *w1 = W + (long)(grp.wr + k) * ldw + col;
                                v[t][0] = w0[0];
                                v[t][1] = w1[0];
                                v[t][2] = pair ? w0[1] : 0.0;
                                v[t][3] = pair ? w1[1] : 0.0;
                            }
                        }
#pragma clang loop unroll(full)
                        for (int t = 0; t < KR; ++t) {
                            const int k = k0 + j + QL * t;
                            if (k < k1) {
                                double *w0 = W + (long)k * ldw + col;
                                double *w1 = W + (long)(grp.wr + k) * ldw + col;
                                w0[0] = ldexp(v[t][0], e);
                                w1[0] = ldexp(v[t][1], e);
                                if (pair) {
                                    w0[1] = ldexp(v[t][2], e);
                                    w1[1] = ldexp(v[t][3], e);
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    if (mx.on) {
        // magnitude bounds of this block's solution entries and operands (mixed form)
        double mxv = 0.0;
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            if (QL * i + j < nb) {
                mxv = fmax(mxv, A::amax(vr[i]));
            }
        }
#pragma clang loop unroll(full)
        for (int off = 1; off < QL; off <<= 1) {
            mxv = fmax(mxv, __shfl_xor(mxv, off));
        }
        const double mw = mxv * fmax(fabs(a), 2.0 * us.bn1);
        // zero blocks: a sentinel far below any exponent
        const int ex = mxv > 0.0 ? ilogb(mxv) + 1 : -1000000;
        const int ew = mw > 0.0 ? ilogb(mw) + 1 : -1000000;
        if (act && j < ucols) {
            const int p = col + j;
            const int ce = o.ce0 + my_exp;
            const bool fresh = first_in_group || starts_here;
            if (grp.last) {
                mx.wgrp[p] = min(max(fresh ? ew : max(ew, ce + mx.wgrp[p]), -1000), 1000);
            }
            else {
                mx.wgrp[p] = fresh ? ew - ce : max(mx.wgrp[p], ew - ce);
            }
            mx.xrel[p] = max(mx.xrel[p], ex - ce);
        }
    }
    if (o.sq.bexp != nullptr) {
        // split-int8 form: exponent of each column's block operand (exact maximum over the
        // block's S and P parts) and, at the group's last step, of its group operand
        const SplitOut &sq = o.sq;
        double m0 = 0.0, m1 = 0.0, x0 = 0.0, x1 = 0.0;
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            if (QL * i + j < nb) {
                const V v = vr[i];
                const V bx = A::bmul(bR, bI, v);
                m0 = fmax(m0, fmax(fabs(a * A::re(v)), fabs(A::re(bx))));
                m1 = fmax(m1, fmax(fabs(a * A::im(v)), fabs(A::im(bx))));
                x0 = fmax(x0, fabs(A::re(v)));
                x1 = fmax(x1, fabs(A::im(v)));
            }
        }
#pragma clang loop unroll(full)
        for (int off = 1; off < QL; off <<= 1) {
            m0 = fmax(m0, __shfl_xor(m0, off));
            m1 = fmax(m1, __shfl_xor(m1, off));
            x0 = fmax(x0, __shfl_xor(x0, off));
            x1 = fmax(x1, __shfl_xor(x1, off));
        }
        if (act && j < ucols) {
            const int p = col + j;
            const double mj = j == 0 ? m0 : m1;
            const double xj = j == 0 ? x0 : x1;
            const int ej = mj > 0.0 ? min(max(ilogb(mj) + 1, -1000), 1000) : -1000;
            sq.bexp[p] = ej;
            sq.xbexp[p] = xj > 0.0 ? min(max(ilogb(xj) + 1, -1000), 1000) : -1000;
            if (grp.last) {
                // the lower blocks the column reaches, each brought to the final scale by
                // 2^(ce - blkexp) above
                const int ce = o.ce0 + my_exp;
                const int d = dl + i0;
                int eg = ej;
                for (int jb = 1; jb < GROUP_MAX_BLOCKS; ++jb) {
                    if (grp.off[jb] < grp.off[jb + 1] && i0 + grp.off[jb] <= d) {
                        const long so = (long)(grp.step[jb] - grp.step[0]) * ldx + p;
                        const int eb = sq.bexp[so];
                        if (eb > -1000) {
                            eg = max(eg, eb + ce - blkexp[so]);
                        }
                    }
                }
                sq.gexp[p] = min(max(eg, -1000), 1000);
            }
        }
    }
    if (act && j < ucols) {
        // lane 0: real part column, lane 1: imaginary part column
        const int p = col + j;
        const double sg = ldexp(1.0, my_exp);
        sig[p] = sg;
        // cumulative scale over the group's steps (the group's far update applies it);
        // a unit starting in this block has no earlier steps in the group
        sig_grp[p] = (first_in_group || starts_here) ? sg : o.sgp * sg;
        const int ce = o.ce0 + my_exp;
        colexp[p] = ce;
        blkexp[p] = ce;
        if (my_exp != 0) {
            *flag = 1;
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_solve_real: one step (row block [i0, i0+nb)) for all active eigenvalue units.
//
// X is the working matrix in row-major layout X[row * ldx + col] (rows = vector components,
// columns = packed working columns).  Rows above the current block hold the running
// right-hand sides, rows of the block hold RHS on entry and the solution on exit.
// W receives the update operands of the block's group, row-major W[k * ldw + col]:
// rows w_s + r hold -a * x_r, rows w_p + r hold (b x)_r, so that
// X_above := sig * X_above + [S_w | P_w] * W.
//
// A unit is one eigenvalue: a real one (one column) or a complex pair (two columns, real
// and imaginary part).  A 16-lane row serves one unit: lane j keeps rows r = 16 i + j of the
// block.  A wave holds four units (tdesc[4 t + q], their descriptors in tile order, idx = -1
// for an empty slot); waves whose active units are all real sweep in real arithmetic, the
// others in complex arithmetic (real units then carry a zero imaginary part).  Before the
// sweep each lane factors its own rows independently: 1x1 rows get the reciprocal of the
// shifted diagonal (a S_rr - b P_rr, perturbed to the rounding level when tiny), 2x2 bumps of
// S (rows r-1, r: bit r-1 of `bumps`) a pivoted LU whose upper row takes 1 / pivot and whose
// lower row takes 1 / U22 and the multiplier (unit_block).  Rows outside the unit's system get
// zero coefficients; the unit's own block gets unit coefficients, its entries being preset to
// the starting vector.  The sweep then runs bottom-up: the owner lane turns the right-hand side
// into the solution, broadcasts it in the row (DPP), and the lanes update the rows above: the
// row slot's own rows at every step, the slots above once per slot.
// A solution entry above 2^958 makes the unit rescale by a power of two (down to 2^446),
// recorded as an exponent (blkexp: the column's exponent after this step).
// Group bookkeeping of the operands W (rows [0, wr): S part, [wr, 2 wr): P part; the group's
// blocks top-down at local rows grp.off[j], solved in steps grp.step[j]):
//   * a unit starting in this block (not the group's first step) zeros its operand rows of
//     the group's lower blocks, which it never wrote;
//   * in the group's last (top) step each unit brings its operand rows of the lower blocks
//     to its final scale: W *= 2^(colexp - blkexp[step of the block]).
// The block {S,P} comes from its packed image pkb (k_pack_steps).  With pkp != nullptr the
// launch also applies the block to the group's rows above it (the near update), for rows
// [r0n, i0) of the workgroup's columns:
//   X[row][c] := sig[c] X[row][c] + sum_k S[row][i0+k] W_s[k][c] + P[row][i0+k] W_p[k][c]
// with W_s = -a x, W_p = b x of this block (the values written to W) and sig the column's
// rescale of this step.  The workgroup's columns (4 slots per real wave, 8 per complex wave,
// at most 32) take the products on 16x16x4 FP64 MFMA over the panel image pkp (rows from g0n
// in tiles of 16, the panel's k-steps: the S part over ks = nb rounded up to 4, then the P
// part), the operand from LDS; each wave takes every nw-th row tile.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack_steps(const double *__restrict__ S, long lds, const double *__restrict__ P,
                                                    long ldp, const StepPack *__restrict__ steps,
                                                    double2 *__restrict__ pk)
{
    const StepPack st = steps[blockIdx.y];
    const int it = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i0 = st.i0, nb = st.nb;
    if (it < PK_BLK) {
        constexpr int NI = sp_col(65);
        double2 v = make_double2(0.0, 0.0);
        if (it < NI) {
            int c, l;
            sp_decode(it, c, l);
            if (c < nb && l < nb) {
                v = make_double2(S[(i0 + l) + (long)(i0 + c) * lds], P[(i0 + l) + (long)(i0 + c) * ldp]);
            }
        }
        else if (it < NI + 65) {
            const int c = it - NI;
            if (c + 1 < nb) {
                v = make_double2(S[(i0 + c + 1) + (long)(i0 + c) * lds], P[(i0 + c + 1) + (long)(i0 + c) * ldp]);
            }
        }
        pk[st.boff + it] = v;
        return;
    }
    const int q = it - PK_BLK; // a double of the panel image
    const int kq = (nb + 3) / 4, per_rt = kq * 128;
    if (q >= st.nrt * per_rt) {
        return;
    }
    const int rt = q / per_rt, rem = q - rt * per_rt;
    const int p = rem >> 7, lane = (rem >> 1) & 63, h = rem & 1;
    const int kk = 2 * p + h;
    const int fr = lane & 15, fk = lane >> 4;
    const bool spart = kk < kq;
    const int kc = 4 * (spart ? kk : kk - kq) + fk;
    const int row = st.g0 + 16 * rt + fr;
    double v = 0.0;
    if (row < i0 && kc < nb) {
        v = spart ? S[row + (long)(i0 + kc) * lds] : P[row + (long)(i0 + kc) * ldp];
    }
    reinterpret_cast<double *>(pk + st.poff)[q] = v;
}

template <int MAXB>
__global__ __launch_bounds__(256) void k_solve_real(
    const double2 *__restrict__ pkb, const double2 *__restrict__ pkp, int g0n, int r0n, int i0, int nb,
    unsigned long long bumps, double *__restrict__ X, long ldx, double *__restrict__ W, long ldw,
    const UnitDesc *__restrict__ tdesc, int t_begin, int t_end, int u_begin,
    int u_end, double *__restrict__ sig, int *__restrict__ colexp, int *__restrict__ blkexp, int *__restrict__ flag,
    int w_s, int w_p, double *__restrict__ sig_grp, int first_in_group, GroupInfo grp, MixedExp mx, SplitOut sq)
{
    static_assert(MAXB <= 65, "block layout: the staging decode covers 5 column blocks");
    constexpr int QR = (MAXB + QL - 1) / QL; // rows per lane
    constexpr int NI = sp_col(MAXB);
    __shared__ double2 sp[sp_col(MAXB)];     // sp[sp_col(c) + l] = {S_lc, P_lc}, l < 16 (c / 16 + 1)
    __shared__ double2 spd[MAXB];            // spd[c] = {S, P}_{c+1, c}
    __shared__ double wl[2 * PK_KSMAX * WL_LD];
    __shared__ double wsig[WL_LD];
    __shared__ int wtyp[4];
    __shared__ int ucol[4][4]; // working column of each unit (-1: inactive), bit 30: a pair

    const int wave = threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nw = blockDim.x / WAVE;
    const int j = lane & (QL - 1);
    const int cq = lane / QL;
    const int t = min(t_begin + (int)(blockIdx.x * (blockDim.x / WAVE)) + wave, t_end - 1);
    const bool wave_live = t_begin + (int)(blockIdx.x * (blockDim.x / WAVE)) + wave < t_end;

    // the block image: one 16-byte load per item, issued first (it needs no descriptor)
    constexpr int PER = 11;
    static_assert(PER * 256 >= NI + MAXB, "one pass of 256 threads");
    double2 stg[PER];
#pragma clang loop unroll(full)
    for (int k = 0; k < PER; ++k) {
        const int it = (int)threadIdx.x + k * (int)blockDim.x;
        stg[k] = it < NI + MAXB ? pkb[it] : make_double2(0.0, 0.0);
    }

    const UnitDesc ud = tdesc[4 * t + cq];
    const int u = ud.idx;
    const bool act = wave_live && u >= 0 && u >= u_begin && u < u_end;
    const int col = ud.col;
    const bool pair = ud.w == 2;
    const bool cpxw = __any(act && pair);
    // the unit's column exponent and group scale, read now rather than after the sweep
    const int ce0 = act ? colexp[col] : 0;
    const double sgp = (act && j < ud.w) ? sig_grp[col + j] : 1.0;
    UnitStep us;
    us.sp = sp;
    us.spd = spd;
    us.nb = nb;
    us.j = j;
    us.dl = act ? ud.d - i0 : -1 - MAXB; // inactive units: every row is a zero row
    us.lol = ud.lo - i0;
    us.init = ud.init;
    us.bumps = bumps;
    us.a = ud.a;
    us.bR = ud.bR;
    us.bI = ud.bI;
    us.bn1 = fabs(ud.bR) + fabs(ud.bI);

    // right-hand side of the unit's rows in the block
    RawRHS<QR> raw;
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = QL * i + j;
        const bool ld = r < nb && r < us.dl && r >= us.lol;
        const double *xp = X + (long)(i0 + r) * ldx + col;
        raw.v[i].x = ld ? xp[0] : 0.0;
        raw.v[i].y = (ld && pair) ? xp[1] : 0.0;
    }

    // block image into LDS (items past the staged columns: the subdiagonal)
#pragma clang loop unroll(full)
    for (int k = 0; k < PER; ++k) {
        const int it = (int)threadIdx.x + k * (int)blockDim.x;
        if (it < NI) {
            sp[it] = stg[k];
        }
        else if (it < NI + MAXB) {
            spd[it - NI] = stg[k];
        }
    }
    if (blockDim.x < 256) {
        // fewer than 4 waves: the rest of the image
        for (int it = (int)threadIdx.x + PER * (int)blockDim.x; it < NI + MAXB; it += (int)blockDim.x) {
            const double2 e = pkb[it];
            if (it < NI) {
                sp[it] = e;
            }
            else {
                spd[it - NI] = e;
            }
        }
    }
    const bool near = pkp != nullptr;
    if (near) {
        if (lane == 0) {
            wtyp[wave] = cpxw ? 8 : 4;
        }
        if (j == 0) {
            ucol[wave][cq] = act ? col | (pair ? 1 << 30 : 0) : -1;
        }
    }
    __syncthreads();

    UnitOut o;
    o.act = act;
    o.pair = pair;
    o.ce0 = ce0;
    o.sgp = sgp;
    o.ucols = ud.w;
    o.col = col;
    o.i0 = i0;
    o.X = X;
    o.ldx = ldx;
    o.W = W;
    o.ldw = ldw;
    o.sig = sig;
    o.colexp = colexp;
    o.blkexp = blkexp;
    o.flag = flag;
    o.w_s = w_s;
    o.w_p = w_p;
    o.sig_grp = sig_grp;
    o.first_in_group = first_in_group;
    o.grp = grp;
    o.mx = mx;
    o.sq = sq;
    o.wl = nullptr;
    o.wsig = wsig;
    o.wslot = 0;
    o.ks = (nb + 3) & ~3;
    // near update: slots of the workgroup's columns (a real wave 4, a complex wave 8: two per
    // unit), this wave's from base; cidx[ct] = the working column of slot 16 ct + (lane % 16)
    int nslot = 0;
    int cidx[2] = {-1, -1};
    if (near) {
        int base = 0;
        for (int w = 0; w < nw; ++w) {
            base += w < wave ? wtyp[w] : 0;
            nslot += wtyp[w];
        }
        o.wl = wl;
        o.wslot = base + (cpxw ? 2 : 1) * cq;
#pragma clang loop unroll(full)
        for (int ct = 0; ct < 2; ++ct) {
            int sl = 16 * ct + (lane & 15);
            for (int w = 0; w < nw; ++w) {
                const int tw = wtyp[w];
                if (sl >= 0 && sl < tw) {
                    const int per = tw / 4, uc = ucol[w][sl / per];
                    const int c = uc & ~(1 << 30);
                    cidx[ct] = uc < 0 ? -1 : (sl % per == 0 ? c : ((uc >> 30) & 1 ? c + 1 : -1));
                }
                sl -= tw;
            }
        }
    }
    if (!cpxw) {
        unit_block<double, QR, MAXB>(us, raw, o);
    }
    else {
        unit_block<cpx, QR, MAXB>(us, raw, o);
    }
    if (!near) {
        return;
    }

    // ---- near update on MFMA: the wave's row tiles rt of the panel (every nw-th), the panel's
    // k-step pairs and the outputs' current X loaded one tile ahead of the products
    const int ks = o.ks, kq = ks / 4, nks = 2 * kq;
    const int nrt = (i0 - g0n + 15) / 16;
    const int nct = (nslot + 15) / 16;
    const int fr = lane & 15, fk = lane >> 4;
    constexpr int KQM = PK_KSMAX / 4;
    auto load_tile = [&](int rt, double2 (&af)[KQM], double (&cv)[2][4]) {
        const double2 *pa = pkp + (long)rt * kq * 64 + lane;
#pragma clang loop unroll(full)
        for (int p = 0; p < KQM; ++p) {
            af[p] = p < kq ? pa[p * 64] : make_double2(0.0, 0.0);
        }
#pragma clang loop unroll(full)
        for (int v = 0; v < 4; ++v) {
            const int row = g0n + 16 * rt + fk + 4 * v;
            const bool ok = row >= r0n && row < i0;
#pragma clang loop unroll(full)
            for (int ct = 0; ct < 2; ++ct) {
                cv[ct][v] = (ok && cidx[ct] >= 0) ? X[(long)row * ldx + cidx[ct]] : 0.0;
            }
        }
    };
    double csg[2];
    auto tile = [&](int rt, const double2 (&af)[KQM], const double (&cv)[2][4]) {
        double4_t acc[2];
#pragma clang loop unroll(full)
        for (int ct = 0; ct < 2; ++ct) {
#pragma clang loop unroll(full)
            for (int v = 0; v < 4; ++v) {
                acc[ct][v] = 0.0;
            }
        }
#pragma clang loop unroll(full)
        for (int kk = 0; kk < 2 * KQM; ++kk) {
            if (kk < nks) {
                const double av = (kk & 1) ? af[kk >> 1].y : af[kk >> 1].x;
                const double *wr = wl + (4 * kk + fk) * WL_LD + fr;
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wr[0], acc[0], 0, 0, 0);
                if (nct > 1) {
                    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wr[16], acc[1], 0, 0, 0);
                }
            }
        }
#pragma clang loop unroll(full)
        for (int v = 0; v < 4; ++v) {
            const int row = g0n + 16 * rt + fk + 4 * v;
#pragma clang loop unroll(full)
            for (int ct = 0; ct < 2; ++ct) {
                if (row >= r0n && row < i0 && cidx[ct] >= 0 && ct < nct) {
                    X[(long)row * ldx + cidx[ct]] = fma(csg[ct], cv[ct][v], acc[ct][v]);
                }
            }
        }
    };
    // the wave's tiles rt = rtw + nw k, k < cnt, taken from k = blockIdx.x mod cnt on (every
    // workgroup reads the whole panel: started at one place they would all meet in the same L2
    // channels)
    const int rtw = (r0n - g0n) / 16 + wave;
    const int cnt = rtw < nrt ? (nrt - rtw + nw - 1) / nw : 0;
    const int k0 = cnt > 0 ? (int)(blockIdx.x % (unsigned)cnt) : 0;
    auto rt_of = [&](int k) { return rtw + nw * (k0 + k < cnt ? k0 + k : k0 + k - cnt); };
    double2 afA[KQM], afB[KQM];
    double cvA[2][4], cvB[2][4];
    if (cnt > 0) {
        load_tile(rt_of(0), afA, cvA); // the panel and X do not depend on the other waves' solves
    }
    __syncthreads();
#pragma clang loop unroll(full)
    for (int ct = 0; ct < 2; ++ct) {
        csg[ct] = wsig[16 * ct + fr];
    }
    for (int k = 0; k < cnt; k += 2) {
        if (k + 1 < cnt) {
            load_tile(rt_of(k + 1), afB, cvB);
        }
        tile(rt_of(k), afA, cvA);
        if (k + 2 < cnt) {
            load_tile(rt_of(k + 2), afA, cvA);
        }
        if (k + 1 < cnt) {
            tile(rt_of(k + 1), afB, cvB);
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_gemm_f64: D[a][b] = sum_k L[k][a] * R[k][b] on FP64 MFMA, 16x16x4 tiles.
//   L: two K segments, L[a + k*ldl] (segment 1, K1 rows) then L2[a + (k-K1)*ldl2]
//   R: R[k*ldr + b] (segment 1), R2[(k-K1)*ldr + b] (segment 2; R + K1*ldr when null)
//   C: row a starts at crow(a) = C + (outcol ? outcol[a] : a) * ldc; element b contiguous.
//   mode 0: C = D;  mode 1: C = sig[b] * C + D.
//   kmax (optional): per-a upper bound of k (exclusive); a tile uses the max over its rows,
//   so triangular operands skip their zero part.
//   afl (optional, non-decreasing in b): rows a < afl[b] - aoff of column b are not needed (rows
//   outside the column's system); a tile whose rows all lie above afl[b0] - aoff of its first
//   column is skipped, and tiles are taken column tiles fastest (not XCD-grouped).
//   Batches (gridDim.y > 1): batch y reads R + y*bsr (and R2 + y*bsr) and writes C + y*bsc;
//   L, kmax and the other arguments are shared.
// Workgroup: 256 threads (2x2 waves), tile GT (a) x GT (b) (each wave GT/2 x GT/2), K chunks
// of GK rows double-buffered in LDS; the next chunk is fetched into registers while the
// current one feeds the MFMAs.  GT = 128, GK = 16 for large products; GT = 64, GK = 32 when
// 128-tiles would leave CUs idle (the per-step updates of a few block rows).  Each segment is padded to whole chunks with zero rows (masked at the
// LDS store, so no wait on the loads precedes the MFMAs).  Columns a >= M, b >= N are
// clamped to the last one: they only feed outputs that are never stored.
// WIDE: 16-byte loads of column pairs (M, N >= 2); otherwise one double per load.
// ---------------------------------------------------------------------------------------

// F32 (mixed form): operands rounded to FP32 at the LDS store, products and accumulation on
// v_mfma_f32_16x16x4 (FP32), the running C stays FP64 (C = sig * C + D in FP64).  To keep
// the FP32 operands in range, L[.][a] is scaled by 2^-sa[a] and R[.][b] by 2^-sb[b]
// (power-of-two exponents bounding the row / column magnitudes; null: no scaling) and D by
// 2^(sa[a] + sb[b]) in the FP64 epilogue.
template <int TAG, bool WIDE, int GTA, int GTB, int GK, bool F32 = false>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K1, int K2, const double *__restrict__ L, long ldl,
                                                  const double *__restrict__ L2, long ldl2,
                                                  const double *__restrict__ R, long ldr, double *__restrict__ C,
                                                  long ldc, const int *__restrict__ outcol, int mode,
                                                  const double *__restrict__ sig, const int *__restrict__ kmax,
                                                  const double *__restrict__ R2, const int *__restrict__ sa,
                                                  const int *__restrict__ sb, long bsr, long bsc,
                                                  const int *__restrict__ afl, int aoff)
{
    R += (long)blockIdx.y * bsr;
    C += (long)blockIdx.y * bsc;
    if (R2 == nullptr) {
        R2 = R + (long)K1 * ldr;
    }
    else {
        R2 += (long)blockIdx.y * bsr;
    }
    typedef typename std::conditional<F32, float, double>::type el_t;
    __shared__ el_t Ls[2][GK][GTA];
    __shared__ el_t Rs[2][GK][GTB];

    // XCD-aware tile order; with a per-row K bound the heaviest row tiles come first
    int tm, tn;
    if (afl == nullptr) {
        xcd_tile(blockIdx.x, (M + GTA - 1) / GTA, (N + GTB - 1) / GTB, tm, tn);
    }
    else {
        // column tiles fastest: consecutive workgroups (dealt to the 8 XCDs in turn) take
        // consecutive column tiles, so the skipped tiles spread evenly over the XCDs
        const int tn_n = (N + GTB - 1) / GTB;
        tm = blockIdx.x / tn_n;
        tn = blockIdx.x - tm * tn_n;
    }
    const int a0 = tm * GTA;
    const int b0 = tn * GTB;
    if (afl != nullptr && aoff + a0 + GTA <= afl[b0]) {
        return;
    }
    const int tid = threadIdx.x;
    const int wave = tid / WAVE;
    const int lane = tid % WAVE;
    const int wa = (wave >> 1) * (GTA / 2); // wave offset in a
    const int wb = (wave & 1) * (GTB / 2);  // wave offset in b

    int K = K1 + K2;
    // TAG 2 with a per-row K bound (the backward's outer products): R entries that are not
    // finite are staged as zero, and nfk[b] keeps the first K index of a non-finite entry of
    // column b; an output whose row's exact K range [0, kmax[a]) includes it becomes NaN in
    // the epilogue.  Outputs whose terms are all finite are then exact, and the zeros of L past
    // a row's own bound never multiply an Inf of R (the tile's K bound is its rows' maximum).
    constexpr bool NFG = TAG == 2;
    __shared__ int nfk[NFG ? GTB : 1];
    const bool guard = NFG && kmax != nullptr;
    if (guard && tid < GTB) {
        nfk[tid] = INT_MAX;
    }
    if (kmax != nullptr) {
        __shared__ int kred;
        if (tid == 0) {
            kred = 0;
        }
        __syncthreads();
        if (tid < GTA && a0 + tid < M) {
            atomicMax(&kred, kmax[a0 + tid]);
        }
        __syncthreads();
        K = min(K, kred);
    }
    const int rows1 = min(K1, K);
    const int rows2 = max(0, min(K2, K - K1));
    const int nch1 = (rows1 + GK - 1) / GK;
    const int nchunks = nch1 + (rows2 + GK - 1) / GK;

    constexpr int FTA = GTA / 32; // 16x16 MFMA blocks per wave tile edge (wave tile GTA/2 x GTB/2)
    constexpr int FTB = GTB / 32;
    static_assert(GK % 4 == 0 && (GTA == 32 || GTA == 64 || GTA == 128) && (GTB == 64 || GTB == 128), "tile shape");
    typedef typename std::conditional<F32, float4_t, double4_t>::type acc_t;
    acc_t acc[FTA][FTB];
#pragma clang loop unroll(full)
    for (int i = 0; i < FTA; ++i) {
#pragma clang loop unroll(full)
        for (int j = 0; j < FTB; ++j) {
#pragma clang loop unroll(full)
            for (int v = 0; v < 4; ++v) {
                acc[i][j][v] = 0;
            }
        }
    }

    // register staging of one chunk.  L: column (pair) cl of the tile, rows lkl + RSL s;
    // R: column (pair) cr, rows lkr + RSR s.
    constexpr int VW = WIDE ? 2 : 1;
    constexpr int RSL = 256 / (GTA / VW), NSL = GK / RSL;
    constexpr int RSR = 256 / (GTB / VW), NSR = GK / RSR;
    static_assert(NSL * RSL == GK && NSR * RSR == GK, "staging layout");
    const int cl = VW * (tid % (GTA / VW)), lkl = tid / (GTA / VW);
    const int cr = VW * (tid % (GTB / VW)), lkr = tid / (GTB / VW);
    const int aa = min(a0 + cl, M - VW);
    const int bb = min(b0 + cr, N - VW);
    const bool odd_a = WIDE && a0 + cl == M - 1; // pair shifted left at the last column
    const bool odd_b = WIDE && b0 + cr == N - 1;
    // FP32 form: scale factors of this thread's operand columns (2^-exponent)
    double fa0 = 1.0, fa1 = 1.0, fb0 = 1.0, fb1 = 1.0;
    if constexpr (F32) {
        if (sa != nullptr) {
            fa0 = ldexp(1.0, -sa[aa]);
            fa1 = WIDE ? ldexp(1.0, -sa[aa + VW - 1]) : fa0;
        }
        if (sb != nullptr) {
            fb0 = ldexp(1.0, -sb[bb]);
            fb1 = WIDE ? ldexp(1.0, -sb[bb + VW - 1]) : fb0;
        }
    }
    typedef typename std::conditional<WIDE, double2, double>::type vec_t;
    vec_t lreg[NSL], rreg[NSR];
    int valid = 0; // rows of the staged chunk inside their segment
    int kbase = 0; // K index of the staged chunk's first row
    auto load_chunk = [&](int ch) {
        const bool s1 = ch < nch1;
        const int k0 = (s1 ? ch : ch - nch1) * GK;
        const int rows = s1 ? rows1 : rows2;
        const long ll = s1 ? ldl : ldl2;
        const double *lb = (s1 ? L : L2) + aa;
        const double *rb = (s1 ? R : R2) + bb;
        valid = rows - k0;
        kbase = (s1 ? 0 : K1) + k0;
        if (valid >= GK) {
#pragma clang loop unroll(full)
            for (int s = 0; s < NSL; ++s) {
                lreg[s] = *(const vec_t *)(lb + (long)(k0 + lkl + RSL * s) * ll);
            }
#pragma clang loop unroll(full)
            for (int s = 0; s < NSR; ++s) {
                rreg[s] = *(const vec_t *)(rb + (long)(k0 + lkr + RSR * s) * ldr);
            }
        }
        else {
#pragma clang loop unroll(full)
            for (int s = 0; s < NSL; ++s) {
                lreg[s] = *(const vec_t *)(lb + (long)min(k0 + lkl + RSL * s, rows - 1) * ll);
            }
#pragma clang loop unroll(full)
            for (int s = 0; s < NSR; ++s) {
                rreg[s] = *(const vec_t *)(rb + (long)min(k0 + lkr + RSR * s, rows - 1) * ldr);
            }
        }
    };
    auto store_chunk = [&](int buf) {
#pragma clang loop unroll(full)
        for (int s = 0; s < NSL; ++s) {
            const int r = lkl + RSL * s;
            const bool in = r < valid;
            if constexpr (WIDE) {
                double2 lv = lreg[s];
                lv.x = odd_a ? lv.y : lv.x;
                if (!in) {
                    lv = make_double2(0.0, 0.0);
                }
                if constexpr (F32) {
                    *(float2 *)&Ls[buf][r][cl] = make_float2((float)(lv.x * (odd_a ? fa1 : fa0)), (float)(lv.y * fa1));
                }
                else {
                    *(double2 *)&Ls[buf][r][cl] = lv;
                }
            }
            else {
                Ls[buf][r][cl] = (el_t)(in ? lreg[s] * fa0 : 0.0);
            }
        }
#pragma clang loop unroll(full)
        for (int s = 0; s < NSR; ++s) {
            const int r = lkr + RSR * s;
            const bool in = r < valid;
            if constexpr (WIDE) {
                double2 rv = rreg[s];
                rv.x = odd_b ? rv.y : rv.x;
                if (!in) {
                    rv = make_double2(0.0, 0.0);
                }
                if constexpr (NFG) {
                    if (guard && !(isfinite(rv.x) && isfinite(rv.y))) {
                        if (!isfinite(rv.x)) {
                            atomicMin(&nfk[cr], kbase + r);
                            rv.x = 0.0;
                        }
                        if (!isfinite(rv.y)) {
                            atomicMin(&nfk[cr + 1], kbase + r);
                            rv.y = 0.0;
                        }
                    }
                }
                if constexpr (F32) {
                    *(float2 *)&Rs[buf][r][cr] = make_float2((float)(rv.x * (odd_b ? fb1 : fb0)), (float)(rv.y * fb1));
                }
                else {
                    *(double2 *)&Rs[buf][r][cr] = rv;
                }
            }
            else {
                double rv = in ? rreg[s] * fb0 : 0.0;
                if constexpr (NFG) {
                    if (guard && !isfinite(rv)) {
                        atomicMin(&nfk[cr], kbase + r);
                        rv = 0.0;
                    }
                }
                Rs[buf][r][cr] = (el_t)rv;
            }
        }
    };

    if (nchunks > 0) {
        load_chunk(0);
        store_chunk(0);
    }
    __syncthreads();
    const int fr = lane % 16; // row/col within a 16x16 tile for fragment loads
    const int fk = lane / 16; // k within a 4-step
    for (int ch = 0; ch < nchunks; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) {
            load_chunk(ch + 1);
        }
#pragma clang loop unroll(full)
        for (int ks = 0; ks < GK; ks += 4) {
            el_t af[FTA], bf[FTB];
#pragma clang loop unroll(full)
            for (int i = 0; i < FTA; ++i) {
                af[i] = Ls[buf][ks + fk][wa + i * 16 + fr];
            }
#pragma clang loop unroll(full)
            for (int j = 0; j < FTB; ++j) {
                bf[j] = Rs[buf][ks + fk][wb + j * 16 + fr];
            }
#pragma clang loop unroll(full)
            for (int i = 0; i < FTA; ++i) {
#pragma clang loop unroll(full)
                for (int j = 0; j < FTB; ++j) {
                    if constexpr (F32) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
                    }
                    else {
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                    }
                }
            }
        }
        if (ch + 1 < nchunks) {
            store_chunk(buf ^ 1);
        }
        __syncthreads();
    }

    // epilogue: lane holds D[lane/16 + 4*v][lane%16] of each 16x16 tile (v_mfma_f64_16x16x4
    // layout), D[4*(lane/16) + v][lane%16] for v_mfma_f32_16x16x4.  The lane's row and column
    // indices are loaded first and the C reads of a row group are issued one group ahead of its
    // writes, so that loads of one kind are in flight together (a wait per output would
    // serialise their round trips).
    // (each uniform choice is made once around a straight run of loads, and the C reads are
    // unconditional at clamped addresses: per-load branches would each end in a wait)
    constexpr int NA = 4 * FTA;
    int rac[NA], rsa[NA], rkm[NA];
    long roff[NA];
#pragma clang loop unroll(full)
    for (int i = 0; i < FTA; ++i) {
#pragma clang loop unroll(full)
        for (int vv = 0; vv < 4; ++vv) {
            const int q = 4 * i + vv;
            rac[q] = min(a0 + wa + i * 16 + (F32 ? 4 * fk + vv : fk + 4 * vv), M - 1);
            roff[q] = (long)rac[q] * ldc;
            rsa[q] = 0;
            rkm[q] = 0;
        }
    }
    if (outcol != nullptr) {
#pragma clang loop unroll(full)
        for (int q = 0; q < NA; ++q) {
            roff[q] = (long)outcol[rac[q]] * ldc;
        }
    }
    if (F32 && sa != nullptr) {
#pragma clang loop unroll(full)
        for (int q = 0; q < NA; ++q) {
            rsa[q] = sa[rac[q]];
        }
    }
    if (NFG && guard) {
#pragma clang loop unroll(full)
        for (int q = 0; q < NA; ++q) {
            rkm[q] = kmax[rac[q]];
        }
    }
    double csig[FTB];
    int cbl[FTB], csb[FTB], cnf[FTB];
#pragma clang loop unroll(full)
    for (int j = 0; j < FTB; ++j) {
        cbl[j] = min(b0 + wb + j * 16 + fr, N - 1);
        csig[j] = 1.0;
        csb[j] = 0;
        cnf[j] = NFG ? nfk[cbl[j] - b0] : 0;
    }
    if (mode == 1) {
#pragma clang loop unroll(full)
        for (int j = 0; j < FTB; ++j) {
            csig[j] = sig[cbl[j]];
        }
    }
    if (F32 && sb != nullptr) {
#pragma clang loop unroll(full)
        for (int j = 0; j < FTB; ++j) {
            csb[j] = sb[cbl[j]];
        }
    }
    double cv[2][4][FTB];
#pragma clang loop unroll(full)
    for (int i = 0; i <= FTA; ++i) {
        if (mode == 1 && i < FTA) {
            // C of row group i (issued before the writes of group i - 1)
#pragma clang loop unroll(full)
            for (int vv = 0; vv < 4; ++vv) {
#pragma clang loop unroll(full)
                for (int j = 0; j < FTB; ++j) {
                    cv[i & 1][vv][j] = C[roff[4 * i + vv] + cbl[j]];
                }
            }
        }
        if (i == 0) {
            continue;
        }
        const int g = i - 1;
#pragma clang loop unroll(full)
        for (int vv = 0; vv < 4; ++vv) {
            const int q = 4 * g + vv;
            const int ar = a0 + wa + g * 16 + (F32 ? 4 * fk + vv : fk + 4 * vv);
#pragma clang loop unroll(full)
            for (int j = 0; j < FTB; ++j) {
                const int bc = b0 + wb + j * 16 + fr;
                if (ar < M && bc < N) {
                    double d = (double)acc[g][j][vv];
                    if constexpr (F32) {
                        d = ldexp(d, rsa[q] + csb[j]);
                    }
                    if (mode == 1) {
                        d = fma(csig[j], cv[g & 1][vv][j], d);
                    }
                    if constexpr (NFG) {
                        if (guard && cnf[j] < rkm[q]) {
                            d = __builtin_nan("");
                        }
                    }
                    C[roff[q] + bc] = d;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_bwd_outer: both gradient products of the backward in one pass over their shared operands,
//     D1[a][b] = sum_k s1[k] L[k][a] R[k][b],   D2[a][b] = sum_k s2[k] L[k][a] R[k][b],
// written to C1 + a ldc1 + b and C2 + a ldc2 + b for a, b < n (L[k][a] = L[a + k ldl],
// R[k][b] = R[k ldr + b]; strides may be negative, k < K).  With L = the forward vectors V,
// R = the adjoint solutions lambda and s1 = -a_c, s2 = b_c (columns in reverse order) these are
// dL/dS and dL/dP.  L and R rows hold npad = n rounded up to even entries (zero past n) at
// 16-byte aligned addresses.  kmax[a] (non-increasing in a) bounds k per row a: a row's terms
// past its own bound are exact zeros of L.  R holds finite values only: k_bwd_lambda_t stages
// a non-finite lambda entry as zero and keeps in nfk[b] the first k of such an entry in column
// b; an output whose exact range [0, kmax[a]) includes it becomes NaN.  Batches: gridDim.y,
// R + y bsr, nfk + y bsn, C1 + y bsc1, C2 + y bsc2.  tiles[blockIdx.x] = (a tile << 16) | b tile:
// the tiles with work (b0 <= a0 + 128) in descending order of their K range, so that
// consecutive workgroups (which the hardware deals to the 8 XCDs in turn) spread the work
// evenly over the XCDs, heaviest first.
// Support: only the entries the forward reads carry a gradient, C1 (dL/dS) on the upper
// Hessenberg band b <= a + 1, C2 (dL/dP) on the upper triangle b <= a; every other entry is
// written as zero (the tiles wholly below the band by k_bwd_zero_tiles).  A tile's K range
// stops at kmax[b0 + 1] = #{c : k_c > b0}: R[k][b] = lambda_c[b] is an exact zero for k_c <= b.
// Tile 128 (a) x 64 (b), four waves of 64 x 32 per product, K steps of 16 rows staged by
// direct global -> LDS loads (16 bytes per lane) into a 2-stage ring one step ahead, together
// with the step's factors s1, s2 (both arrays hold OB_GK entries past K), which multiply the A
// fragment in registers (zero past the tile's K range).  50 KB of LDS and at most 168 VGPRs, so
// that three workgroups share a compute unit: while one waits for its loads or writes its
// outputs, the other two keep the FP64 MFMA pipes busy.  Each wave skips its MFMA steps past
// its own K range.
// ---------------------------------------------------------------------------------------
constexpr int OB_GA = 128, OB_GB = 64, OB_GK = 16, OB_NST = 2;
__global__ __launch_bounds__(256, 3) void k_bwd_outer(int n, int K, const double *__restrict__ L, long ldl,
                                                   const double *__restrict__ R, long ldr,
                                                   const double *__restrict__ s1, const double *__restrict__ s2,
                                                   double *__restrict__ C1, long ldc1, double *__restrict__ C2,
                                                   long ldc2, const int *__restrict__ kmax,
                                                   const int *__restrict__ nfk, const int *__restrict__ tiles,
                                                   long bsr, long bsn, long bsc1, long bsc2)
{
    constexpr int GTA = OB_GA, GTB = OB_GB, GK = OB_GK;
    constexpr int FTA = GTA / 32, FTB = GTB / 32;
    constexpr int FOFF = GK * (GTA + GTB); // the step's factors s1, s2 after the operand rows
    constexpr int STAGE = FOFF + 128;      // doubles per ring stage
    constexpr int IPW = (GK * GTA / 128 + GK * GTB / 128) / 4; // 16-byte load instructions per wave and stage
    static_assert(IPW == 6, "staging layout: 4 L rows and 2 R row pairs per wave");
    // (tile, batch) of this workgroup: with a multiple of 8 batches each XCD (the hardware deals
    // consecutive workgroups to the 8 XCDs in turn) takes whole batches, tile-major, so that its
    // L2 holds the lambda of its own batches only
    int bx = blockIdx.x, by = blockIdx.y;
    if (gridDim.y % 8 == 0) {
        const long id = (long)blockIdx.y * gridDim.x + blockIdx.x;
        const long pos = id / 8;
        const int bpx = gridDim.y / 8;
        bx = (int)(pos / bpx);
        by = (int)(id % 8) + 8 * (int)(pos % bpx);
    }
    R += (long)by * bsr;
    nfk += (long)by * bsn;
    C1 += (long)by * bsc1;
    C2 += (long)by * bsc2;
    __shared__ __attribute__((aligned(16))) double lds[OB_NST * STAGE];
    const int npad = (n + 1) & ~1;
    const int tp = tiles[bx];
    const int a0 = (tp >> 16) * GTA, b0 = (tp & 0xffff) * GTB;
    const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
    const int wa = (wave >> 1) * (GTA / 2), wb = (wave & 1) * (GTB / 2);
    const int KK = min(min(K, kmax[a0]), b0 + 1 < n ? kmax[b0 + 1] : 0);
    const int nst = (KK + GK - 1) / GK;
    // this wave's own K range (its first row and first column bound every term of its block;
    // the terms of k in [KW, KK) are exact zeros for its outputs), zero when every entry of its
    // block lies below the band; its MFMA steps past KW are skipped
    int KW = a0 + wa < n ? min(KK, kmax[a0 + wa]) : 0;
    KW = min(KW, b0 + wb + 1 < n ? kmax[b0 + wb + 1] : 0);
    if (b0 + wb > a0 + wa + 64) {
        KW = 0;
    }

    // per-lane sources: L rows 4 u + wave (u < 4), lane pair 2 lane of the row; R row pairs
    // 2 t, 2 t + 1 with t = 4 u + wave (u < 2), lanes 0-31 the first row, 32-63 the second.
    // Pairs past npad are clamped (they only feed outputs that are never stored).
    const double *lsrc = L + min(a0 + 2 * lane, npad - 2);
    const double *rsrc = R + min(b0 + 2 * (lane & 31), npad - 2);
    auto issue = [&](int st) {
        const int k0 = st * GK;
        double *stage = lds + (st % OB_NST) * STAGE;
#pragma clang loop unroll(full)
        for (int u = 0; u < 4; ++u) {
            const int r = 4 * u + wave;
            const double *g = lsrc + (long)min(k0 + r, K - 1) * ldl;
            __builtin_amdgcn_global_load_lds((const void *)g,
                                             (__attribute__((address_space(3))) void *)(stage + r * GTA), 16, 0, 0);
        }
#pragma clang loop unroll(full)
        for (int u = 0; u < 2; ++u) {
            const int t = 4 * u + wave;
            const int r = 2 * t + (lane >> 5);
            const double *g = rsrc + (long)min(k0 + r, K - 1) * ldr;
            __builtin_amdgcn_global_load_lds((const void *)g,
                                             (__attribute__((address_space(3))) void *)(stage + GK * GTA + 2 * t * GTB),
                                             16, 0, 0);
        }
        if (wave == 0) {
            // s1[k0, k0 + 16) by lanes 0-7, s2[k0, k0 + 16) by lanes 8-15 (both arrays hold GK
            // entries past K); lanes 16-63 repeat lanes 0-15 into the padding
            const int l = lane % 16;
            const double *g = (l < 8 ? s1 : s2) + k0 + 2 * (l % 8);
            __builtin_amdgcn_global_load_lds((const void *)g, (__attribute__((address_space(3))) void *)(stage + FOFF),
                                             16, 0, 0);
        }
    };

    double4_t acc1[FTA][FTB], acc2[FTA][FTB];
#pragma clang loop unroll(full)
    for (int i = 0; i < FTA; ++i) {
#pragma clang loop unroll(full)
        for (int j = 0; j < FTB; ++j) {
#pragma clang loop unroll(full)
            for (int v = 0; v < 4; ++v) {
                acc1[i][j][v] = 0.0;
                acc2[i][j][v] = 0.0;
            }
        }
    }
    if (nst > 0) {
        issue(0);
    }
    const int fr = lane % 16, fk = lane / 16;
    for (int st = 0; st < nst; ++st) {
        wait_vm<0>();
        // every wave's loads of this stage have landed and every wave is done with the stage
        // refilled next
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        if (st + 1 < nst) {
            issue(st + 1);
        }
        const double *Ls = lds + (st % OB_NST) * STAGE;
        const double *Rs = Ls + GK * GTA;
        const double *Fs = Ls + FOFF;
        const int k0 = st * GK;
#pragma clang loop unroll(full)
        for (int ks = 0; ks < GK; ks += 4) {
            if (k0 + ks < KW) {
                // the lane's factors for k = k0 + ks + fk, zero past the tile's K range (the
                // staged rows there repeat row K - 1)
                const bool in = k0 + ks + fk < KK;
                const double f1 = in ? Fs[ks + fk] : 0.0;
                const double f2 = in ? Fs[GK + ks + fk] : 0.0;
                double bf[FTB];
#pragma clang loop unroll(full)
                for (int j = 0; j < FTB; ++j) {
                    bf[j] = Rs[(ks + fk) * GTB + wb + j * 16 + fr];
                }
#pragma clang loop unroll(full)
                for (int i = 0; i < FTA; ++i) {
                    const double af = Ls[(ks + fk) * GTA + wa + i * 16 + fr];
                    const double a1 = af * f1, a2 = af * f2;
#pragma clang loop unroll(full)
                    for (int j = 0; j < FTB; ++j) {
                        acc1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bf[j], acc1[i][j], 0, 0, 0);
                        acc2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bf[j], acc2[i][j], 0, 0, 0);
                    }
                }
            }
        }
    }
    // epilogue: the first non-finite k of each of the lane's columns, then the outputs
    int cnf[FTB];
#pragma clang loop unroll(full)
    for (int j = 0; j < FTB; ++j) {
        cnf[j] = nfk[min(b0 + wb + j * 16 + fr, n - 1)];
    }
#pragma clang loop unroll(full)
    for (int i = 0; i < FTA; ++i) {
#pragma clang loop unroll(full)
        for (int vv = 0; vv < 4; ++vv) {
            const int ar = a0 + wa + i * 16 + fk + 4 * vv;
            if (ar >= n) {
                continue;
            }
            const int kr = kmax[ar];
#pragma clang loop unroll(full)
            for (int j = 0; j < FTB; ++j) {
                const int bc = b0 + wb + j * 16 + fr;
                if (bc < n) {
                    const bool bad = cnf[j] < kr;
                    const double v1 = bad ? __builtin_nan("") : acc1[i][j][vv];
                    const double v2 = bad ? __builtin_nan("") : acc2[i][j][vv];
                    C1[(long)ar * ldc1 + bc] = bc <= ar + 1 ? v1 : 0.0;
                    C2[(long)ar * ldc2 + bc] = bc <= ar ? v2 : 0.0;
                }
            }
        }
    }
}

// k_bwd_zero_tiles: the tiles of k_bwd_outer's grid that lie wholly below the band (b0 > a0 + 128)
// hold zeros only, in both gradients; this kernel writes them.  It holds no LDS and few
// registers, so that its workgroups fit on the compute units beside the block solves of the
// adjoint sweep that it runs next to (on a second stream), instead of keeping them out as
// k_bwd_outer's 50 KB operand ring would.  tiles[blockIdx.x] = (a tile << 16) | b tile as in
// k_bwd_outer; batches gridDim.y, C1 + y bsc1, C2 + y bsc2.
__global__ __launch_bounds__(256) void k_bwd_zero_tiles(int n, double *__restrict__ C1, long ldc1,
                                                        double *__restrict__ C2, long ldc2,
                                                        const int *__restrict__ tiles, long bsc1, long bsc2)
{
    C1 += (long)blockIdx.y * bsc1;
    C2 += (long)blockIdx.y * bsc2;
    const int tp = tiles[blockIdx.x];
    const int a0 = (tp >> 16) * OB_GA, b0 = (tp & 0xffff) * OB_GB;
    const int bc = b0 + (int)threadIdx.x % OB_GB;
    if (bc >= n) {
        return;
    }
    const int a1 = min(a0 + OB_GA, n);
    for (int ar = a0 + (int)threadIdx.x / OB_GB; ar < a1; ar += 256 / OB_GB) {
        C1[(long)ar * ldc1 + bc] = 0.0;
        C2[(long)ar * ldc2 + bc] = 0.0;
    }
}

// out[r] = max(out[r], e) with e the exponent bounding max_k max(|A[r + k lda]|, |B[r + k ldb]|)
// (B may be null) over the column chunk [64 blockIdx.y, +64) of k < ncols, and k >= cfirst[r]
// when cfirst is given, clamped to [-1000, 1000]; out is filled with -1000 beforehand.
// Threads run over rows (coalesced).
constexpr int ROWEXP_CHUNK = 64;
__global__ void k_rowexp(const double *__restrict__ A, long lda, const double *__restrict__ B, long ldb, int n,
                         int ncols, int *__restrict__ out, const int *__restrict__ cfirst)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = blockIdx.y * ROWEXP_CHUNK;
    if (r >= n) {
        return;
    }
    const int k1 = min(ncols, k0 + ROWEXP_CHUNK);
    double mx = 0.0;
    for (int k = cfirst != nullptr ? max(k0, cfirst[r]) : k0; k < k1; ++k) {
        mx = fmax(mx, fabs(A[r + (long)k * lda]));
        if (B != nullptr) {
            mx = fmax(mx, fabs(B[r + (long)k * ldb]));
        }
    }
    if (mx > 0.0) {
        atomicMax(out + r, min(max(ilogb(mx) + 1, -1000), 1000));
    }
}

// out[p] = clamp(colexp[p] + xrel[p]): exponent bounding the final column p of X.
__global__ void k_xfin(const int *__restrict__ colexp, const int *__restrict__ xrel, int m, int *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) {
        out[p] = min(max(colexp[p] + xrel[p], -1000), 1000);
    }
}

__global__ void k_fill_int(int *__restrict__ X, long count, int v)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
        X[i] = v;
    }
}

// X[r][p] *= 2^(colexp[p] - blkexp[s][p]) for the rows r of step s (rows [srow[2s], srow[2s+1])):
// one thread per column p of a step (consecutive columns, so each row is read and written as
// contiguous runs), skipping the (step, column) pairs whose exponent did not change after the
// step.  flag[0] == 0: no column ever rescaled.
__global__ __launch_bounds__(256) void k_fixup_steps(double *__restrict__ X, long ldx, int m,
                                                     const int *__restrict__ srow, const int *__restrict__ blkexp,
                                                     long ldb, const int *__restrict__ colexp,
                                                     const int *__restrict__ flag)
{
    if (*flag == 0) {
        return;
    }
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (p >= m) {
        return;
    }
    const int e = colexp[p] - blkexp[(long)s * ldb + p];
    if (e == 0) {
        return;
    }
    const int r0 = srow[2 * s], r1 = srow[2 * s + 1];
    double *x = X + (long)r0 * ldx + p;
    int r = r0;
    for (; r + 4 <= r1; r += 4, x += 4 * ldx) {
        double v[4];
#pragma clang loop unroll(full)
        for (int u = 0; u < 4; ++u) {
            v[u] = x[u * ldx];
        }
#pragma clang loop unroll(full)
        for (int u = 0; u < 4; ++u) {
            x[u * ldx] = ldexp(v[u], e);
        }
    }
    for (; r < r1; ++r, x += ldx) {
        *x = ldexp(*x, e);
    }
}

// X[r][p] *= 2^(colexp[p] - blkexp[step(r)][p]) for the solved rows r <= top(p).
__global__ void k_fixup_scale(double *__restrict__ X, long ldx, int n, int m, const int *__restrict__ rowstep,
                              const int *__restrict__ blkexp, long ldb, const int *__restrict__ colexp,
                              const int *__restrict__ flag)
{
    if (*flag == 0) {
        return;
    }
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (p >= m || r >= n) {
        return;
    }
    int s = rowstep[r];
    int e = colexp[p] - blkexp[(long)s * ldb + p];
    if (e != 0) {
        double *x = X + (long)r * ldx + p;
        *x = ldexp(*x, e);
    }
}

// V[i + outcol(p)*ldv] = X[row(i)][p] with row(i) = rev ? n-1-i : i, through a 32x32 LDS tile.
__global__ void k_copy_out(const double *__restrict__ X, long ldx, int n, int m, double *__restrict__ V, long ldv,
                           const int *__restrict__ outcol, int rev)
{
    __shared__ double tile[32][33];
    int p0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y; // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        int i = i0 + k, p = p0 + tx;
        double v = 0.0;
        if (i < n && p < m) {
            int row = rev ? n - 1 - i : i;
            v = X[(long)row * ldx + p];
        }
        tile[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int p = p0 + k, i = i0 + tx;
        if (i < n && p < m) {
            V[i + (long)outcol[p] * ldv] = tile[tx][k];
        }
    }
}

// One workgroup per vector: largest |v| (pairs: |re| + |im|) over n rows, then scale by its
// reciprocal (LAPACK xTGEVC convention).  vcol[q] = first output column, vw[q] = 1 or 2.
__global__ __launch_bounds__(256) void k_normalize(double *__restrict__ V, long ldv, int n, const int *__restrict__ vcol,
                                                   const int *__restrict__ vw)
{
    __shared__ double red[256 / WAVE];
    const int q = blockIdx.x;
    double *c0 = V + (long)vcol[q] * ldv;
    double *c1 = c0 + ldv;
    const bool two = vw[q] == 2;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t = fabs(c0[i]);
        if (two) {
            t += fabs(c1[i]);
        }
        mx = fmax(mx, t);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
    }
    if (threadIdx.x % WAVE == 0) {
        red[threadIdx.x / WAVE] = mx;
    }
    __syncthreads();
    mx = 0.0;
    for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) {
        mx = fmax(mx, red[w]);
    }
    if (!(mx > kSafeMin)) {
        return;
    }
    double sc = 1.0 / mx;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        c0[i] *= sc;
        if (two) {
            c1[i] *= sc;
        }
    }
}

// out[r + c*ldo] = (float) V[r + c*ldv] for r < rows, c < cols (one workgroup per column):
// FP32 copies of finished vector columns for FP32 host arrays.
__global__ __launch_bounds__(256) void k_round_f32(const double *__restrict__ V, long ldv, long rows,
                                                   float *__restrict__ out, long ldo)
{
    const double *src = V + (long)blockIdx.x * ldv;
    float *dst = out + (long)blockIdx.x * ldo;
    for (long r = threadIdx.x; r < rows; r += blockDim.x) {
        dst[r] = (float)src[r];
    }
}

// D[i + j*ldd] = A[(n-1-j) + (n-1-i)*lda]  (D = J A^T J), 32x32 LDS tiles.
__global__ void k_reverse_transpose(const double *__restrict__ A, long lda, int n, double *__restrict__ D, long ldd)
{
    __shared__ double tile[32][33];
    int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    int tx = threadIdx.x, ty = threadIdx.y;
    // read A rows r = n-1-(j0..j0+31), cols c = n-1-(i0..i0+31); coalesce along r
    for (int k = ty; k < 32; k += 8) {
        int j = j0 + tx, i = i0 + k;
        double v = 0.0;
        if (i < n && j < n) {
            v = A[(long)(n - 1 - j) + (long)(n - 1 - i) * lda];
        }
        tile[k][tx] = v; // tile[i-i0][j-j0]
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int i = i0 + tx, j = j0 + k;
        if (i < n && j < n) {
            D[i + (long)j * ldd] = tile[tx][k];
        }
    }
}

// Column sums of |S| and |P| over rows 0..min(n-1, c+hess) (hess = 1: upper Hessenberg
// part, dtgevc3.cpp:1252-1269; hess = 0: triangle, ztgevc3.cpp:528-542).
__global__ void k_colnorms(const double *__restrict__ S, long lds, const double *__restrict__ P, long ldp, int n,
                           int hess, double *__restrict__ out)
{
    int c = blockIdx.x;
    int lim = min(n - 1, c + hess);
    double ss = 0.0, sp = 0.0;
    for (int r = threadIdx.x; r <= lim; r += blockDim.x) {
        ss += fabs(S[r + (long)c * lds]);
        sp += fabs(P[r + (long)c * ldp]);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        ss += __shfl_xor(ss, off);
        sp += __shfl_xor(sp, off);
    }
    __shared__ double red[2][4];
    if (threadIdx.x % WAVE == 0) {
        red[0][threadIdx.x / WAVE] = ss;
        red[1][threadIdx.x / WAVE] = sp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) {
            a += red[0][w];
            b += red[1][w];
        }
        out[c] = a;
        out[n + c] = b;
    }
}

// out[i] = S[i, i-1] (out[0] = 0): the subdiagonal, gathered so one contiguous copy brings it
// to the host for the block partition.
__global__ void k_subdiag(const double *__restrict__ S, long lds, int n, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = i > 0 ? S[i + (long)(i - 1) * lds] : 0.0;
    }
}

// Fill a row-major block with zeros (a kernel rather than hipMemsetAsync, so that the sweep's
// launch sequence stays replayable from a HIP graph).
__global__ void k_zero(double *__restrict__ X, long count)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        X[i] = 0.0;
    }
}

__global__ void k_zero_int(int *__restrict__ X, long count)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        X[i] = 0;
    }
}

} // namespace xtg

namespace xtg {
template __global__ void k_solve_real<65>(const double2 *, const double2 *, int, int, int, int, unsigned long long,
                                          double *, long, double *, long, const UnitDesc *, int, int,
                                          int, int, double *, int *, int *, int *, int, int, double *, int, GroupInfo,
                                          MixedExp, SplitOut);
} // namespace xtg

// =======================================================================================
// Backward (adjoint) kernels of the real path.  For a selected real eigenvalue k with
// normalised forward vector v (largest |v_p| = 1 at p) and cotangent g = dL/dv:
//   g'' = g - (g . v) sign(v_p) e_p           (pull-back through the normalisation)
//   lambda = (a S - b P)[0:k,0:k]^{-T} g''    (transposed shifted solve, same pencil)
//   dL/dS = -a lambda v^T, dL/dP = b lambda v^T (rows < k, cols <= k)
//   dL/dalpha = s lambda^T (P v), dL/dbeta = -s (b/a) lambda^T (P v), alpha = S_kk,
//   beta = P_kk, (a, b) = s (beta, alpha).
// The solve runs on the reversed transpose J S^T J, so lambda_i sits at working row n-1-i.
// =======================================================================================

namespace xtg {

// RHS of the adjoint solve in working rows: X[(n-1-i)*ldx + q] = g_i - [i == i*] (g.v) sign(v_i*)
// for i < kcol[c] (zero elsewhere), where working column q = (m-1-c)*nbatch + b holds batch b of
// forward column c and i* is the first index of max |v_i| (correct == 0: G is already the
// right-hand side, copied without the correction term).  One workgroup per TQ working
// columns and one pass over G: TQ x 512 tiles of g go through LDS to rows of X (TQ contiguous
// doubles per row) while each thread accumulates g.v over its rows (rows i with the same
// i mod 256, in order), then the partial sums are added in a fixed tree (so a column's
// result depends on neither nbatch nor TQ) and the row i* of each column is rewritten with
// its correction.  TQ = 4 keeps one cotangent batch on enough workgroups; with many batches
// TQ = 16 writes whole 128-byte lines of X.
template <int TQ>
__global__ __launch_bounds__(256) void k_bwd_rhs(const double *__restrict__ V, long ldv, const double *__restrict__ G,
                                                 long ldg, int n, int m, int nbatch, const int *__restrict__ kcol,
                                                 double *__restrict__ X, long ldx, int correct)
{
    constexpr int TI = 512;
    static_assert((TQ * TI) % 256 == 0 && TI == 2 * 256, "a thread takes rows ii and ii + 256 of each column");
    __shared__ double tile[TQ][TI + 1];
    __shared__ double red[256 / WAVE][TQ];
    __shared__ int s_k[TQ];
    const int M = m * nbatch, q0 = blockIdx.x * TQ;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    if (threadIdx.x < TQ) {
        const int q = q0 + threadIdx.x;
        s_k[threadIdx.x] = q < M ? kcol[m - 1 - q / nbatch] : 0;
    }
    __syncthreads();
    double dot[TQ];
#pragma clang loop unroll(full)
    for (int j = 0; j < TQ; ++j) {
        dot[j] = 0.0;
    }
    for (int i0 = 0; i0 < n; i0 += TI) {
#pragma clang loop unroll(full)
        for (int u = 0; u < TQ * TI / 256; ++u) {
            const int t = (int)threadIdx.x + 256 * u;
            const int j = t / TI, ii = t % TI, i = i0 + ii, q = q0 + j;
            double val = 0.0;
            if (q < M && i < n) {
                const int c = m - 1 - q / nbatch, b = q % nbatch;
                const double g = G[(long)(b * m + c) * ldg + i];
                dot[j] = fma(g, V[(long)c * ldv + i], dot[j]);
                val = i < s_k[j] ? g : 0.0;
            }
            tile[j][ii] = val;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < TQ * TI; t += 256) {
            const int ii = t / TQ, j = t % TQ, i = i0 + ii, q = q0 + j;
            if (q < M && i < n) {
                X[(long)(n - 1 - i) * ldx + q] = tile[j][ii];
            }
        }
        __syncthreads();
    }
    if (!correct) {
        return;
    }
#pragma clang loop unroll(full)
    for (int j = 0; j < TQ; ++j) {
        double d = dot[j];
#pragma clang loop unroll(full)
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            d += __shfl_xor(d, off);
        }
        if (lane == 0) {
            red[wave][j] = d;
        }
    }
    __syncthreads();
    // row i* of each column: the first index of max |v_i| (one column per wave at a time)
    for (int j = wave; j < TQ; j += 256 / WAVE) {
        const int q = q0 + j;
        if (q >= M) {
            continue;
        }
        const int c = m - 1 - q / nbatch, b = q % nbatch;
        const double *v = V + (long)c * ldv;
        double best = -1.0;
        int bi = 0;
        for (int i = lane; i < n; i += WAVE) {
            const double a = fabs(v[i]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oi = __shfl_xor(bi, off);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0 && bi < s_k[j]) {
            double tot = 0.0;
#pragma clang loop unroll(full)
            for (int w = 0; w < 256 / WAVE; ++w) {
                tot += red[w][j];
            }
            const double corr = v[bi] < 0.0 ? -tot : tot;
            X[(long)(n - 1 - bi) * ldx + q] = G[(long)(b * m + c) * ldg + bi] - corr;
        }
    }
}

template __global__ void k_bwd_rhs<4>(const double *, long, const double *, long, int, int, int, const int *, double *,
                                      long, int);
template __global__ void k_bwd_rhs<16>(const double *, long, const double *, long, int, int, int, const int *, double *,
                                       long, int);

// Lam[(b*m + c) * ldl + i] = lambda_i = 2^-e X[n-1-i][q] of working column q (batch b, forward
// column c; e = blkexp of the column at the step that solved working row n-1-i, whose scale the
// row keeps), for i < ldl (zero for i >= n: ldl pads the rows), in 32x32 LDS tiles over (q, i).  A
// non-finite entry is written as zero, ovf[q] = 1 marks the column, and nfk[b * ldl + i] keeps
// the smallest reversed column index m - 1 - c of such an entry in row i (the K index of the
// outer products, which take the columns in reverse order; nfk starts at INT_MAX).
__global__ void k_bwd_lambda_t(const double *__restrict__ X, long ldx, int n, int m, int nbatch,
                               const int *__restrict__ kcol, const int *__restrict__ rowstep,
                               const int *__restrict__ blkexp, double *__restrict__ Lam, long ldl,
                               int *__restrict__ ovf, int *__restrict__ nfk)
{
    __shared__ double tile[32][33];
    const int q0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int M = m * nbatch;
    for (int kk = ty; kk < 32; kk += 8) {
        int i = i0 + kk, q = q0 + tx;
        double v = 0.0;
        if (i < n && q < M) {
            int c = m - 1 - q / nbatch;
            if (i < kcol[c]) {
                const long r = n - 1 - i;
                v = ldexp(X[r * ldx + q], -blkexp[(long)rowstep[r] * ldx + q]);
            }
        }
        tile[kk][tx] = v;
    }
    __syncthreads();
    for (int kk = ty; kk < 32; kk += 8) {
        int q = q0 + kk, i = i0 + tx;
        if (i < ldl && q < M) {
            int c = m - 1 - q / nbatch, b = q % nbatch;
            double v = tile[tx][kk];
            if (!isfinite(v)) {
                ovf[q] = 1;
                atomicMin(&nfk[(long)b * ldl + i], m - 1 - c);
                v = 0.0;
            }
            Lam[(long)(b * m + c) * ldl + i] = v; // lambda itself (the outer products apply -a_c, b_c)
        }
    }
}

// out(i, b m + c) = A(ri, b m + cc) for i < n, c < m, b < nbatch, with ri = n - 1 - i when
// rev_rows (else i) and cc = m - 1 - c when rev_cols (else c): reversal of the rows and / or of
// the columns within each batch of m (the working coordinates of left eigenvectors).
__global__ void k_bwd_permute(const double *__restrict__ A, long lda, int n, int m, int nbatch, int rev_rows,
                              int rev_cols, double *__restrict__ out, long ldo)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n || q >= m * nbatch) {
        return;
    }
    const int b = q / m, c = q % m;
    const int src = b * m + (rev_cols ? m - 1 - c : c);
    out[i + (long)q * ldo] = A[(rev_rows ? n - 1 - i : i) + (long)src * lda];
}

// Cotangent of the unnormalised vector: for column q = b m + c of G (batch b, column c) and the
// output vector v_c (V column c), out(:, q) = g - (g . v) sign(v_p) e_p with p the first index of
// max |v_i| (the LAPACK normalisation divides by |v_p|).  One wave per column, fixed order.
__global__ __launch_bounds__(64) void k_bwd_gcorr(const double *__restrict__ V, long ldv,
                                                  const double *__restrict__ G, long ldg, int n, int m,
                                                  double *__restrict__ out, long ldo)
{
    const int q = blockIdx.x, c = q % m, lane = threadIdx.x;
    const double *v = V + (long)c * ldv;
    const double *g = G + (long)q * ldg;
    double best = -1.0, dot = 0.0;
    int bi = 0;
    for (int i = lane; i < n; i += WAVE) {
        const double a = fabs(v[i]);
        if (a > best) {
            best = a;
            bi = i;
        }
        dot = fma(g[i], v[i], dot);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
        dot += __shfl_xor(dot, off);
    }
    const double corr = v[bi] < 0.0 ? -dot : dot;
    for (int i = lane; i < n; i += WAVE) {
        out[i + (long)q * ldo] = g[i] - (i == bi ? corr : 0.0);
    }
}

// W(j, c) = 0 for rows j > kcol[c]: the Schur-basis vector Z^T v of a real eigenvalue k has no
// components below k (the product leaves rounding noise there).
__global__ void k_bwd_trunc(double *__restrict__ W, long ldw, int n, int m, const int *__restrict__ kcol)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j < n && c < m && j > kcol[c]) {
        W[j + (long)c * ldw] = 0.0;
    }
}

// Vt[k * ldt + c] = V[k + c * ldv] (column-major -> row-major), 32x32 tiles.
__global__ void k_to_rowmajor(const double *__restrict__ V, long ldv, int n, int m, double *__restrict__ Vt, long ldt)
{
    __shared__ double tile[32][33];
    const int c0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int kk = ty; kk < 32; kk += 8) {
        int c = c0 + kk, k = k0 + tx;
        tile[kk][tx] = (c < m && k < n) ? V[k + (long)c * ldv] : 0.0;
    }
    __syncthreads();
    for (int kk = ty; kk < 32; kk += 8) {
        int k = k0 + kk, c = c0 + tx;
        if (c < m && k < n) {
            Vt[(long)k * ldt + c] = tile[tx][kk];
        }
    }
}

// dL/dalpha and dL/dbeta of working column q, the terms t_S = s qv and t_P = -s (b/a) qv with
// qv = lambda^T (P v), Y = P V (row-major, ldy), lambda_i = 2^-e X[n-1-i][q] (e = blkexp of the
// column at the step that solved working row n-1-i: the sweep's rows stay at their steps'
// scales), written to out[2q], out[2q + 1] (k_bwd_diag_add folds them into dS_kk, dP_kk):
// 16 working columns per workgroup, thread t = column t % 16 and row phase t / 16 (a row of X is
// read as 16 contiguous values), the rows i < k split over the 16 phases (i = phase + 16 r, four
// running sums per thread), the partial sums added in fixed order.
constexpr int BD_QC = 16;
__global__ __launch_bounds__(256) void k_bwd_diag(const double *__restrict__ X, long ldx, const double *__restrict__ Y,
                                                  long ldy, int n, int m, int nbatch, const int *__restrict__ kcol,
                                                  const double *__restrict__ coef, const int *__restrict__ rowstep,
                                                  const int *__restrict__ blkexp, double *__restrict__ out)
{
    constexpr int QC = BD_QC, PH = 256 / BD_QC;
    __shared__ double part[PH][QC];
    const int col = threadIdx.x % QC, ph = threadIdx.x / QC;
    const int M = m * nbatch;
    const int q = blockIdx.x * QC + col;
    const bool live = q < M;
    const int c = live ? m - 1 - q / nbatch : 0;
    const int k = live ? kcol[c] : 0;
    int kw = k; // the largest row count of the workgroup's columns (each wave holds all of them)
#pragma clang loop unroll(full)
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        kw = max(kw, __shfl_xor(kw, off));
    }
    const double *xq = X + (long)(n - 1) * ldx + (live ? q : 0);
    const double *yc = Y + c;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = ph; i0 < kw; i0 += 4 * PH) {
#pragma clang loop unroll(full)
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + PH * u;
            if (i < k) {
                const double lam = ldexp(xq[-(long)i * ldx], -blkexp[(long)rowstep[n - 1 - i] * ldx + q]);
                acc[u] = fma(lam, yc[(long)i * ldy], acc[u]);
            }
        }
    }
    part[ph][col] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (ph == 0 && live) {
        double tot = 0.0;
#pragma clang loop unroll(full)
        for (int r = 0; r < PH; ++r) {
            tot += part[r][col];
        }
        const double qv = tot;
        const double s = coef[4 * c + 2], ratio = coef[4 * c + 3];
        out[2 * (long)q] = s * qv;
        out[2 * (long)q + 1] = -s * ratio * qv;
    }
}

// dS_kk += t_S, dP_kk += t_P of each working column q (batch b = q % nbatch, forward column
// c = m - 1 - q / nbatch, k = kcol[c]; one column per thread, every (b, k) once).
__global__ void k_bwd_diag_add(const double *__restrict__ t, int n, int m, int nbatch, const int *__restrict__ kcol,
                               double *__restrict__ dS, long ldds, double *__restrict__ dP, long lddp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m * nbatch) {
        return;
    }
    const int c = m - 1 - q / nbatch, b = q % nbatch, k = kcol[c];
    dS[(long)b * n * ldds + k + (long)k * ldds] += t[2 * (long)q];
    dP[(long)b * n * lddp + k + (long)k * lddp] += t[2 * (long)q + 1];
}

// Tile configuration of a product with M x N outputs and depth K: the one with the smallest
// estimated time, counting workgroup rounds (concurrent workgroups per CU from LDS and
// registers; a last partial round costs about half a round plus its share) times the cost of
// one full round measured on MI355X (K = 512: 128x64 and 64x128 tiles run three workgroups per
// CU and beat 128x128 per unit area; for deep products 128x128 wins).  batch products of the
// same shape share the grid.
int gemm_config(int M, int N, int K, int batch)
{
    struct Cfg {
        int id, ga, gb, per_cu;
        double cost, cost_deep; // relative time of one full round (K <= 1024, K > 1024)
    };
    static const Cfg cfgs[] = {
        {GEMM_128x128, 128, 128, 2, 1.00, 1.00},
        {GEMM_128x64, 128, 64, 3, 0.667, 0.80},
        {GEMM_64x128, 64, 128, 3, 0.68, 0.82},
        {GEMM_64x64, 64, 64, 2, 0.26, 0.30},
        {GEMM_32x64, 32, 64, 3, 0.15, 0.17},
    };
    int best = GEMM_128x128;
    double tbest = 1e300;
    for (const Cfg &c : cfgs) {
        if (c.id == GEMM_32x64 && (M > 256 || K > 256)) {
            continue; // measured only on the short near-update products
        }
        const long tiles = (long)((M + c.ga - 1) / c.ga) * ((N + c.gb - 1) / c.gb) * batch;
        const long conc = 256L * c.per_cu;
        const long full = tiles / conc, rem = tiles % conc;
        const double rounds = (double)full + (rem ? 0.5 + 0.5 * (double)rem / conc : 0.0);
        const double t = rounds * (K > 1024 ? c.cost_deep : c.cost);
        if (t < tbest) {
            tbest = t;
            best = c.id;
        }
    }
    return best;
}

// Host launcher: picks the 16-byte-load form when both output extents are >= 2 and the tile
// configuration (returned, GEMM_*) from gemm_config.  tag only separates the uses in profiles:
// 0 = RHS update, 1 = back-transform, 2 = backward.  prec 1: FP32-product form (tags 0, 1)
// with operand exponents sa (per a) and sb (per b).
int launch_gemm_f64(int tag, hipStream_t st, int M, int N, int K1, int K2, const double *L, long ldl, const double *L2,
                     long ldl2, const double *R, long ldr, double *C, long ldc, const int *outcol, int mode,
                     const double *sig, const int *kmax, const double *R2, int prec, const int *sa,
                     const int *sb, int batch, long bsr, long bsc, const int *afl, int aoff)
{
    const bool wide = M >= 2 && N >= 2;
    const int cfg = gemm_config(M, N, K1 + K2, batch);
    const int gta = cfg == GEMM_32x64 ? 32 : (cfg == GEMM_64x64 || cfg == GEMM_64x128 ? 64 : 128);
    const int gtb = cfg == GEMM_64x64 || cfg == GEMM_128x64 || cfg == GEMM_32x64 ? 64 : 128;
    const dim3 grid(((N + gtb - 1) / gtb) * ((M + gta - 1) / gta), batch);
#define XTG_GEMM_LAUNCH(T, WD, GA, GB, K)                                                                            \
    if (prec == 1 && T != 2) {                                                                                       \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_f64<T, WD, GA, GB, K, true>), grid, dim3(256), 0, st, M, N, K1, K2, \
                           L, ldl, L2, ldl2, R, ldr, C, ldc, outcol, mode, sig, kmax, R2, sa, sb, bsr, bsc, afl,    \
                           aoff);                                                                                    \
    }                                                                                                                \
    else {                                                                                                           \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_f64<T, WD, GA, GB, K>), grid, dim3(256), 0, st, M, N, K1, K2, L,  \
                           ldl, L2, ldl2, R, ldr, C, ldc, outcol, mode, sig, kmax, R2, (const int *)nullptr,         \
                           (const int *)nullptr, bsr, bsc, afl, aoff);                                               \
    }
#define XTG_GEMM_WIDE(T, GA, GB, K)                                                                                  \
    if (wide) {                                                                                                      \
        XTG_GEMM_LAUNCH(T, true, GA, GB, K)                                                                          \
    }                                                                                                                \
    else {                                                                                                           \
        XTG_GEMM_LAUNCH(T, false, GA, GB, K)                                                                         \
    }
#define XTG_GEMM_CASE(T)                                                                                             \
    if (tag == T) {                                                                                                  \
        switch (cfg) {                                                                                               \
        case GEMM_64x64:                                                                                             \
            XTG_GEMM_WIDE(T, 64, 64, 32)                                                                             \
            break;                                                                                                   \
        case GEMM_32x64:                                                                                             \
            XTG_GEMM_WIDE(T, 32, 64, 32)                                                                             \
            break;                                                                                                   \
        case GEMM_64x128:                                                                                            \
            XTG_GEMM_WIDE(T, 64, 128, 16)                                                                            \
            break;                                                                                                   \
        case GEMM_128x64:                                                                                            \
            XTG_GEMM_WIDE(T, 128, 64, 16)                                                                            \
            break;                                                                                                   \
        default:                                                                                                     \
            XTG_GEMM_WIDE(T, 128, 128, 16)                                                                           \
        }                                                                                                            \
        return cfg;                                                                                                  \
    }
    XTG_GEMM_CASE(0)
    XTG_GEMM_CASE(1)
    XTG_GEMM_CASE(2)
#undef XTG_GEMM_CASE
#undef XTG_GEMM_WIDE
#undef XTG_GEMM_LAUNCH
    return cfg;
}
} // namespace xtg

namespace xtg {

// ---------------------------------------------------------------------------------------
// Eigenvalue condition numbers (LAPACK dtgsna, JOB = 'E').  U, V, YA, YB row-major n x m
// (element (i, c) at i * ld + c): U the left vectors, V the right vectors, YA = S V and
// YB = P V.  Unit q covers column ucol[q] (uw[q] = 1) or the pair ucol[q], ucol[q] + 1
// (real and imaginary part, uw[q] = 2).  k_tgsna_part: thread (q, chunk) sums rows
// [chunk * rows, +rows) into part[(chunk * nu + q) * TGSNA_SUMS + t]:
//   t = 0..3: u0.a0, u1.a0, u0.a1, u1.a1 (a = YA columns), 4..7 the same with YB,
//   8, 9: |u0|^2, |u1|^2, 10, 11: |v0|^2, |v1|^2.
// k_tgsna_fin sums the chunks in order and forms s (dtgsna's real and complex-pair formulas).
// ---------------------------------------------------------------------------------------
constexpr int TGSNA_SUMS = 12;

__global__ __launch_bounds__(256) void k_tgsna_part(const double *__restrict__ U, const double *__restrict__ V,
                                                    const double *__restrict__ YA, const double *__restrict__ YB,
                                                    long ld, int n, int nu, const int *__restrict__ ucol,
                                                    const int *__restrict__ uw, int rows, double *__restrict__ part)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (q >= nu) {
        return;
    }
    const int c = ucol[q];
    const bool pair = uw[q] == 2;
    double acc[TGSNA_SUMS];
#pragma clang loop unroll(full)
    for (int t = 0; t < TGSNA_SUMS; ++t) {
        acc[t] = 0.0;
    }
    const int i1 = min(n, (chunk + 1) * rows);
    for (int i = chunk * rows; i < i1; ++i) {
        const long o = (long)i * ld + c;
        const double u0 = U[o], v0 = V[o], a0 = YA[o], b0 = YB[o];
        acc[0] = fma(u0, a0, acc[0]);
        acc[4] = fma(u0, b0, acc[4]);
        acc[8] = fma(u0, u0, acc[8]);
        acc[10] = fma(v0, v0, acc[10]);
        if (pair) {
            const double u1 = U[o + 1], v1 = V[o + 1], a1 = YA[o + 1], b1 = YB[o + 1];
            acc[1] = fma(u1, a0, acc[1]);
            acc[2] = fma(u0, a1, acc[2]);
            acc[3] = fma(u1, a1, acc[3]);
            acc[5] = fma(u1, b0, acc[5]);
            acc[6] = fma(u0, b1, acc[6]);
            acc[7] = fma(u1, b1, acc[7]);
            acc[9] = fma(u1, u1, acc[9]);
            acc[11] = fma(v1, v1, acc[11]);
        }
    }
    double *p = part + ((long)chunk * nu + q) * TGSNA_SUMS;
#pragma clang loop unroll(full)
    for (int t = 0; t < TGSNA_SUMS; ++t) {
        p[t] = acc[t];
    }
}

__global__ void k_tgsna_fin(const double *__restrict__ part, int nchunk, int nu, const int *__restrict__ ucol,
                            const int *__restrict__ uw, double *__restrict__ s)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nu) {
        return;
    }
    double t[TGSNA_SUMS];
#pragma clang loop unroll(full)
    for (int k = 0; k < TGSNA_SUMS; ++k) {
        t[k] = 0.0;
    }
    for (int ch = 0; ch < nchunk; ++ch) {
        const double *p = part + ((long)ch * nu + q) * TGSNA_SUMS;
#pragma clang loop unroll(full)
        for (int k = 0; k < TGSNA_SUMS; ++k) {
            t[k] += p[k];
        }
    }
    const int c = ucol[q];
    if (uw[q] == 1) {
        // real eigenvalue: |(u'Sv, u'Pv)| / (|u| |v|)
        s[c] = hypot(t[0], t[4]) / (sqrt(t[8]) * sqrt(t[10]));
    }
    else {
        // complex pair u = u0 + i u1, v = v0 + i v1: |u^H S v| from the four real products
        const double av = t[0] + t[3], avi = t[2] - t[1];
        const double bv = t[4] + t[7], bvi = t[6] - t[5];
        const double cond = hypot(hypot(av, avi), hypot(bv, bvi));
        const double rn = hypot(sqrt(t[10]), sqrt(t[11])), ln = hypot(sqrt(t[8]), sqrt(t[9]));
        s[c] = s[c + 1] = cond / (rn * ln);
    }
}

} // namespace xtg
