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
#include <cmath>
#include <type_traits>

#include "common.h"

namespace xtg {

struct cpx {
    double r, i;
};

__device__ __forceinline__ cpx cmul(cpx a, cpx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ double abs1(cpx a) { return fabs(a.r) + fabs(a.i); }

// Smith's complex division (no overflow of |c|^2).  Real lanes have c.i == 0 and get an
// exact real division.
__device__ __forceinline__ cpx cdiv(cpx x, cpx c)
{
    if (fabs(c.r) >= fabs(c.i)) {
        double t = c.i / c.r;
        double den = c.r + c.i * t;
        return {(x.r + x.i * t) / den, (x.i - x.r * t) / den};
    }
    double t = c.r / c.i;
    double den = c.i + c.r * t;
    return {(x.r * t + x.i) / den, (x.i * t - x.r) / den};
}

// Exponent e >= 0 such that 2^-e * num <= lim * den (lim = 2^k), computed without
// forming num / den (which may overflow).
__device__ __forceinline__ int scale_exponent(double num, double den, int log2lim)
{
    int e = ilogb(num) - ilogb(den) - log2lim + 2;
    return e > 0 ? e : 0;
}

constexpr double kEps = 2.220446049250313e-16;     // DBL_EPSILON
constexpr double kSafeMin = 2.2250738585072014e-308; // DBL_MIN
constexpr int kLog2Big = 960;                        // |x| entries are kept <= 2^960

struct Bump2 {
    cpx x1, x2;
    int e;
};

// 2x2 (complex) block solve with partial pivoting on the first column and overflow guard.
// Restates the effect of the Hessenberg GEPP of dlauhs/dlau2s (dtgevc3.cpp:132-344) on a
// 2x2 diagonal bump of S: [c11 c12; c21 c22] [x1; x2] = [r1; r2].
__device__ __noinline__ Bump2 bump_solve(cpx r1, cpx r2, cpx c11, cpx c12, cpx c21, cpx c22, double dmin1,
                                         double dmin2)
{
    bool sw = abs1(c21) > abs1(c11);
    cpx p = sw ? c21 : c11, q = sw ? c22 : c12;
    cpx s = sw ? c11 : c21, t = sw ? c12 : c22;
    cpx rp = sw ? r2 : r1, rs = sw ? r1 : r2;
    if (abs1(p) < dmin1) {
        p = {dmin1, 0.0};
    }
    cpx mlt = cdiv(s, p);
    cpx u = csub(t, cmul(mlt, q));
    cpx y = csub(rs, cmul(mlt, rp));
    if (abs1(u) < dmin2) {
        u = {dmin2, 0.0};
    }
    Bump2 out;
    out.e = 0;
    double yn = abs1(y);
    if (yn > 0.0 && ilogb(yn) - ilogb(abs1(u)) > kLog2Big - 2) {
        int e = scale_exponent(yn, abs1(u), kLog2Big);
        double f = ldexp(1.0, -e);
        y = {y.r * f, y.i * f};
        rp = {rp.r * f, rp.i * f};
        out.e += e;
    }
    cpx x2 = cdiv(y, u);
    cpx w = csub(rp, cmul(q, x2));
    double wn = abs1(w);
    if (wn > 0.0 && ilogb(wn) - ilogb(abs1(p)) > kLog2Big - 2) {
        int e = scale_exponent(wn, abs1(p), kLog2Big);
        double f = ldexp(1.0, -e);
        w = {w.r * f, w.i * f};
        x2 = {x2.r * f, x2.i * f};
        out.e += e;
    }
    out.x1 = cdiv(w, p);
    out.x2 = x2;
    return out;
}

// ---------------------------------------------------------------------------------------
// k_solve_real: one step (row block [i0, i0+nb)) for all active columns.
//
// X is the working matrix in row-major layout X[row * ldx + col] (rows = vector components,
// columns = packed eigenvector columns).  Rows above the current block hold the running
// right-hand sides, rows of the block hold RHS on entry and the solution on exit.
// W receives the update operands of the block's group, row-major W[k * ldw + col]:
// rows w_s + r hold -a * x_r, rows w_p + r hold (b x)_r (own part), so that
// X_above := sig * X_above + [S_w | P_w] * W.
//
// Four lanes (a quad) serve one column: lane j of the quad keeps rows r = 4 i + j of the
// block in registers.  Rows are solved bottom-up; the owner of row r broadcasts its value
// to the quad (DPP quad_perm), every lane of the quad computes x_r, and each lane applies
// x_r to its own rows above r.  {S,P} of the diagonal block is staged in LDS.  A column
// whose solution would exceed 2^960 is scaled by a power of two (recorded as exponent).
// ---------------------------------------------------------------------------------------
constexpr int QL = 4;            // lanes per column

template <int SRC> __device__ __forceinline__ double quad_bcast_c(double v)
{
    constexpr int ctrl = SRC * 0x55; // quad_perm [SRC, SRC, SRC, SRC]
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double quad_bcast(double v, int src)
{
    switch (src) {
    case 0:
        return quad_bcast_c<0>(v);
    case 1:
        return quad_bcast_c<1>(v);
    case 2:
        return quad_bcast_c<2>(v);
    default:
        return quad_bcast_c<3>(v);
    }
}

template <int MAXB>
__global__ __launch_bounds__(256) void k_solve_real(
    const double *__restrict__ S, long lds, const double *__restrict__ P, long ldp, int i0, int nb,
    double *__restrict__ X, long ldx, double *__restrict__ W, long ldw, const ColDesc *__restrict__ desc,
    const int2 *__restrict__ tiles, int t_begin, int t_end, int c_begin, int c_end, double *__restrict__ sig,
    int *__restrict__ colexp, int *__restrict__ blkexp, int *__restrict__ flag, int *__restrict__ gflag, int w_s, int w_p,
    double *__restrict__ sig_grp, int first_in_group)
{
    constexpr int QR = (MAXB + QL - 1) / QL; // rows per lane
    __shared__ double2 sp[MAXB * MAXB];
    // stage {S,P} of the diagonal block (zero outside nb x nb); 8 loads in flight per thread
    for (int q0 = threadIdx.x; q0 < MAXB * MAXB; q0 += 8 * blockDim.x) {
        double2 t[8];
#pragma clang loop unroll(full)
        for (int jj = 0; jj < 8; ++jj) {
            int q = q0 + jj * blockDim.x;
            int l = q % MAXB, r = q / MAXB;
            t[jj] = make_double2(0.0, 0.0);
            if (q < MAXB * MAXB && l < nb && r < nb) {
                t[jj] = make_double2(S[(i0 + l) + (long)(i0 + r) * lds], P[(i0 + l) + (long)(i0 + r) * ldp]);
            }
        }
#pragma clang loop unroll(full)
        for (int jj = 0; jj < 8; ++jj) {
            int q = q0 + jj * blockDim.x;
            if (q < MAXB * MAXB) {
                sp[q] = t[jj];
            }
        }
    }
    __syncthreads();

    const int wave = threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int j = lane & (QL - 1);
    const int cq = lane / QL;
    const int t = t_begin + blockIdx.x * (blockDim.x / WAVE) + wave;
    if (t >= t_end) {
        return;
    }
    const int2 tl = tiles[t];
    const int p = tl.x + cq;
    const bool act = cq < tl.y && p >= c_begin && p < c_end;
    const ColDesc cd = desc[act ? p : tl.x];
    const bool is_im = cd.flags & 1;
    const bool is_pair = act && (cd.flags & 2) != 0;
    const int init = (cd.flags >> 2) & 3;
    const int plane = act ? lane + QL * (cd.partner - p) : lane;
    const bool wave_pairs = __any(is_pair);
    const double a = cd.a, bR = cd.bR, bI = cd.bI;
    const double bIs = is_im ? bI : -bI;
    const double bn1 = fabs(bR) + fabs(bI);
    const int dl = act ? cd.d - i0 : -1 - MAXB; // inactive columns: every row is "fixed zero"
    const int lol = cd.lo - i0;

    // initial values of the column's own diagonal block rows (complex)
    cpx x0 = {0.0, 0.0}, x1v = {0.0, 0.0};
    if (act && dl >= 0 && dl < nb) {
        if (init == INIT_UNIT) {
            x0 = {1.0, 0.0};
        }
        else if (init == INIT_PAIR_ROW1 && dl + 1 < nb) {
            double2 e11 = sp[dl + dl * MAXB], e12 = sp[dl + (dl + 1) * MAXB];
            x0 = {-(a * e12.x - bR * e12.y), bI * e12.y};
            x1v = {a * e11.x - bR * e11.y, -bI * e11.y};
        }
        else if (init == INIT_PAIR_ROW2 && dl + 1 < nb) {
            double2 e22 = sp[(dl + 1) + (dl + 1) * MAXB], e21 = sp[(dl + 1) + dl * MAXB];
            x0 = {a * e22.x - bR * e22.y, -bI * e22.y};
            x1v = {-a * e21.x, 0.0};
        }
    }

    double vr[QR];
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = QL * i + j;
        vr[i] = (r < nb && r < dl && r >= lol) ? X[(long)(i0 + r) * ldx + p] : 0.0;
    }

    int my_exp = 0;
    bool pre = false;    // row r was solved together with row r+1 (2x2 bump)
    double pre_xp = 0.0; // partner part of that row's solution
#pragma clang loop unroll(full)
    for (int r = MAXB - 1; r >= 0; --r) {
        if (r < nb) {
            const int jr = r & (QL - 1), ir = r / QL;
            double xs = 0.0, xp = 0.0;
            bool bump_now = false;
            int e = 0;
            Bump2 bs;
            cpx xsol = {0.0, 0.0};
            const bool from_pre = pre;
            if (pre) {
                xs = quad_bcast(vr[ir], jr);
                xp = pre_xp;
                pre = false;
            }
            else if (r >= 1 && sp[r + (r - 1) * MAXB].x != 0.0) {
                // 2x2 bump at rows (r-1, r)
                bump_now = true;
                const int jr1 = (r - 1) & (QL - 1), ir1 = (r - 1) / QL;
                double o1 = quad_bcast(vr[ir1], jr1), o2 = quad_bcast(vr[ir], jr);
                double q1 = wave_pairs ? __shfl(o1, plane) : 0.0;
                double q2 = wave_pairs ? __shfl(o2, plane) : 0.0;
                cpx r1 = is_im ? cpx{q1, o1} : cpx{o1, is_pair ? q1 : 0.0};
                cpx r2 = is_im ? cpx{q2, o2} : cpx{o2, is_pair ? q2 : 0.0};
                double2 e11 = sp[(r - 1) + (r - 1) * MAXB], e12 = sp[(r - 1) + r * MAXB];
                double2 e21 = sp[r + (r - 1) * MAXB], e22 = sp[r + r * MAXB];
                cpx c11 = {a * e11.x - bR * e11.y, -bI * e11.y};
                cpx c12 = {a * e12.x - bR * e12.y, -bI * e12.y};
                cpx c21 = {a * e21.x - bR * e21.y, -bI * e21.y};
                cpx c22 = {a * e22.x - bR * e22.y, -bI * e22.y};
                double dm1 = fmax(kEps * (fabs(a) * (fabs(e11.x) + fabs(e21.x)) + bn1 * fabs(e11.y)), kSafeMin);
                double dm2 = fmax(kEps * (fabs(a) * (fabs(e12.x) + fabs(e22.x)) + bn1 * fabs(e22.y)), kSafeMin);
                bs = bump_solve(r1, r2, c11, c12, c21, c22, dm1, dm2);
                const bool solve1 = (r - 1) < dl && (r - 1) >= lol;
                const bool solve2 = r < dl && r >= lol;
                e = (solve1 || solve2) ? bs.e : 0;
            }
            else {
                double o = quad_bcast(vr[ir], jr);
                double qv = wave_pairs ? __shfl(o, plane) : 0.0;
                cpx rhs = is_im ? cpx{qv, o} : cpx{o, is_pair ? qv : 0.0};
                double2 ee = sp[r + r * MAXB];
                cpx c = {a * ee.x - bR * ee.y, -bI * ee.y};
                double dmin = fmax(kEps * (fabs(a) * fabs(ee.x) + bn1 * fabs(ee.y)), kSafeMin);
                if (abs1(c) < dmin) {
                    c = {dmin, 0.0};
                }
                const bool solve = r < dl && r >= lol;
                // |x| <= |rhs|_1 / |c|_1 * 2 must stay below 2^960
                if (solve && abs1(rhs) > 0x1p958 * abs1(c)) {
                    e = scale_exponent(abs1(rhs), abs1(c), kLog2Big);
                }
                if (e != 0) {
                    double f = ldexp(1.0, -e);
                    rhs = {rhs.r * f, rhs.i * f};
                }
                xsol = solve ? cdiv(rhs, c) : (r == dl ? x0 : (r == dl + 1 ? x1v : cpx{0.0, 0.0}));
            }
            if (__any(e != 0)) {
                // rescale the whole column by 2^-e (rare)
                const double f = ldexp(1.0, -e);
#pragma clang loop unroll(full)
                for (int i = 0; i < QR; ++i) {
                    vr[i] *= f;
                }
                my_exp -= e;
            }
            if (bump_now) {
                const bool solve1 = (r - 1) < dl && (r - 1) >= lol;
                const bool solve2 = r < dl && r >= lol;
                cpx s1 = solve1 ? bs.x1 : ((r - 1) == dl ? x0 : ((r - 1) == dl + 1 ? x1v : cpx{0.0, 0.0}));
                cpx s2 = solve2 ? bs.x2 : (r == dl ? x0 : (r == dl + 1 ? x1v : cpx{0.0, 0.0}));
                const int jr1 = (r - 1) & (QL - 1), ir1 = (r - 1) / QL;
                if (j == jr1) {
                    vr[ir1] = is_im ? s1.i : s1.r;
                }
                pre_xp = is_im ? s1.r : s1.i;
                xs = is_im ? s2.i : s2.r;
                xp = is_im ? s2.r : s2.i;
                pre = true;
            }
            else if (!from_pre) {
                xs = is_im ? xsol.i : xsol.r;
                xp = is_im ? xsol.r : xsol.i;
            }
            if (j == jr) {
                vr[ir] = xs;
            }
            const double ax = a * xs;
            const double bx = bR * xs + bIs * xp;
            // this lane's rows l = 4 i + j above r (row r-1 excluded after a bump)
#pragma clang loop unroll(full)
            for (int i = 0; i < QR; ++i) {
                if (QL * i < r) {
                    const int l = QL * i + j;
                    const bool upd = l < r && !(bump_now && l == r - 1);
                    double2 el = sp[l + r * MAXB];
                    double nv = fma(-el.x, ax, vr[i]);
                    nv = fma(el.y, bx, nv);
                    vr[i] = upd ? nv : vr[i];
                }
            }
        }
    }

    // W of this block at rows w_s + r (S part) and w_p + r (P part) of the group's operand
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = QL * i + j;
        const double o = vr[i];
        const double q = wave_pairs ? __shfl(o, plane) : 0.0;
        if (act && r < nb) {
            X[(long)(i0 + r) * ldx + p] = o;
            W[(long)(w_s + r) * ldw + p] = -a * o;
            W[(long)(w_p + r) * ldw + p] = bR * o + bIs * (is_pair ? q : 0.0);
        }
    }
    if (act && j == 0) {
        const double sg = ldexp(1.0, my_exp);
        sig[p] = sg;
        // cumulative scale over the group's steps (the group's far update applies it);
        // a column starting in this block has no earlier steps in the group
        const bool starts_here = dl >= 0 && dl < nb;
        sig_grp[p] = (first_in_group || starts_here) ? sg : sig_grp[p] * sg;
        int ce = colexp[p] + my_exp;
        colexp[p] = ce;
        blkexp[p] = ce;
        if (my_exp != 0) {
            // the operands of the group's earlier (lower) blocks carry the old scale: the
            // group's rescale pass (k_w_rescale) brings them to the final one
            *flag = 1;
            *gflag = 1;
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_w_rescale: after the last step of a group, bring the group's update operands W (rows
// [0, wr) S part, [wr, 2 wr) P part; local row k belongs to the block holding row g0 + k)
// to the columns' final scale: W[k][p] *= 2^(blkexp[last][p] - blkexp[step(k)][p]).
// Blocks are listed top-down: block j covers local rows [off_j, off_{j+1}) and was solved
// in step steps_j; steps.x is the group's last step.  Skipped when no column of the group
// was rescaled (*gflag == 0).
// ---------------------------------------------------------------------------------------
__global__ void k_w_rescale(double *__restrict__ W, long ldw, int wr, int c0, int ncols, const int *__restrict__ blkexp,
                            long ldb, int4 offs, int4 steps, const int *__restrict__ gflag)
{
    if (*gflag == 0) {
        return;
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y; // 0 .. 2 wr - 1
    if (x >= ncols) {
        return;
    }
    const int kl = k < wr ? k : k - wr;
    const int j = (kl >= offs.y) + (kl >= offs.z) + (kl >= offs.w);
    const int st = j == 0 ? steps.x : (j == 1 ? steps.y : (j == 2 ? steps.z : steps.w));
    const int p = c0 + x;
    const int e = blkexp[(long)steps.x * ldb + p] - blkexp[(long)st * ldb + p];
    if (e != 0) {
        double *w = W + (long)k * ldw + p;
        *w = ldexp(*w, e);
    }
}

// W[k][c0 + x] = 0 for k < rows, x < ncols
__global__ void k_zero_cols(double *__restrict__ W, long ldw, int rows, int c0, int ncols)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < ncols && (int)blockIdx.y < rows) {
        W[(long)blockIdx.y * ldw + c0 + x] = 0.0;
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
// Workgroup: 256 threads (2x2 waves of 64x64), tile 128 (a) x 128 (b), K chunks of 16 rows
// double-buffered in LDS; the next chunk is fetched into registers while the current one
// feeds the MFMAs.  Each segment is padded to whole chunks with zero rows (masked at the
// LDS store, so no wait on the loads precedes the MFMAs).  Columns a >= M, b >= N are
// clamped to the last one: they only feed outputs that are never stored.
// WIDE: 16-byte loads of column pairs (M, N >= 2); otherwise one double per load.
// ---------------------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int GT = 128; // tile edge
constexpr int GK = 16;  // K chunk

template <int TAG, bool WIDE>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K1, int K2, const double *__restrict__ L, long ldl,
                                                  const double *__restrict__ L2, long ldl2,
                                                  const double *__restrict__ R, long ldr, double *__restrict__ C,
                                                  long ldc, const int *__restrict__ outcol, int mode,
                                                  const double *__restrict__ sig, const int *__restrict__ kmax,
                                                  const double *__restrict__ R2)
{
    if (R2 == nullptr) {
        R2 = R + (long)K1 * ldr;
    }
    __shared__ double Ls[2][GK][GT];
    __shared__ double Rs[2][GK][GT];

    // with a per-row K bound (triangular operand) the heaviest tiles are dispatched first
    const int ya = kmax != nullptr ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
    const int a0 = ya * GT;
    const int b0 = blockIdx.x * GT;
    const int tid = threadIdx.x;
    const int wave = tid / WAVE;
    const int lane = tid % WAVE;
    const int wa = (wave >> 1) * 64; // wave offset in a
    const int wb = (wave & 1) * 64;  // wave offset in b

    int K = K1 + K2;
    if (kmax != nullptr) {
        __shared__ int kred;
        if (tid == 0) {
            kred = 0;
        }
        __syncthreads();
        if (tid < GT && a0 + tid < M) {
            atomicMax(&kred, kmax[a0 + tid]);
        }
        __syncthreads();
        K = min(K, kred);
    }
    const int rows1 = min(K1, K);
    const int rows2 = max(0, min(K2, K - K1));
    const int nch1 = (rows1 + GK - 1) / GK;
    const int nchunks = nch1 + (rows2 + GK - 1) / GK;

    double4_t acc[4][4];
#pragma clang loop unroll(full)
    for (int i = 0; i < 4; ++i) {
#pragma clang loop unroll(full)
        for (int j = 0; j < 4; ++j) {
            acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
        }
    }

    // register staging of one chunk: NS rows per thread (WIDE: column pair c, rows lk + 4 s;
    // otherwise column c, rows lk + 2 s)
    constexpr int VW = WIDE ? 2 : 1;
    constexpr int RSTEP = 256 / (GT / VW);
    constexpr int NS = GK / RSTEP;
    const int c = VW * (tid % (GT / VW));
    const int lk = tid / (GT / VW);
    const int aa = min(a0 + c, M - VW);
    const int bb = min(b0 + c, N - VW);
    const bool odd_a = WIDE && a0 + c == M - 1; // pair shifted left at the last column
    const bool odd_b = WIDE && b0 + c == N - 1;
    typedef typename std::conditional<WIDE, double2, double>::type vec_t;
    vec_t lreg[NS], rreg[NS];
    int valid = 0; // rows of the staged chunk inside their segment
    auto load_chunk = [&](int ch) {
        const bool s1 = ch < nch1;
        const int k0 = (s1 ? ch : ch - nch1) * GK;
        const int rows = s1 ? rows1 : rows2;
        const long ll = s1 ? ldl : ldl2;
        const double *lb = (s1 ? L : L2) + aa;
        const double *rb = (s1 ? R : R2) + bb;
        valid = rows - k0;
        if (valid >= GK) {
#pragma clang loop unroll(full)
            for (int s = 0; s < NS; ++s) {
                const int k = k0 + lk + RSTEP * s;
                lreg[s] = *(const vec_t *)(lb + (long)k * ll);
                rreg[s] = *(const vec_t *)(rb + (long)k * ldr);
            }
        }
        else {
#pragma clang loop unroll(full)
            for (int s = 0; s < NS; ++s) {
                const int k = min(k0 + lk + RSTEP * s, rows - 1);
                lreg[s] = *(const vec_t *)(lb + (long)k * ll);
                rreg[s] = *(const vec_t *)(rb + (long)k * ldr);
            }
        }
    };
    auto store_chunk = [&](int buf) {
#pragma clang loop unroll(full)
        for (int s = 0; s < NS; ++s) {
            const int r = lk + RSTEP * s;
            const bool in = r < valid;
            if constexpr (WIDE) {
                double2 lv = lreg[s], rv = rreg[s];
                lv.x = odd_a ? lv.y : lv.x;
                rv.x = odd_b ? rv.y : rv.x;
                if (!in) {
                    lv = make_double2(0.0, 0.0);
                    rv = make_double2(0.0, 0.0);
                }
                *(double2 *)&Ls[buf][r][c] = lv;
                *(double2 *)&Rs[buf][r][c] = rv;
            }
            else {
                Ls[buf][r][c] = in ? lreg[s] : 0.0;
                Rs[buf][r][c] = in ? rreg[s] : 0.0;
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
            double af[4], bf[4];
#pragma clang loop unroll(full)
            for (int i = 0; i < 4; ++i) {
                af[i] = Ls[buf][ks + fk][wa + i * 16 + fr];
                bf[i] = Rs[buf][ks + fk][wb + i * 16 + fr];
            }
#pragma clang loop unroll(full)
            for (int i = 0; i < 4; ++i) {
#pragma clang loop unroll(full)
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                }
            }
        }
        if (ch + 1 < nchunks) {
            store_chunk(buf ^ 1);
        }
        __syncthreads();
    }

    // epilogue: lane holds D[lane/16 + 4*v][lane%16] of each 16x16 tile (v_mfma_f64_16x16x4 layout)
#pragma clang loop unroll(full)
    for (int i = 0; i < 4; ++i) {
#pragma clang loop unroll(full)
        for (int vv = 0; vv < 4; ++vv) {
            int ar = a0 + wa + i * 16 + fk + 4 * vv;
            if (ar >= M) {
                continue;
            }
            long rowoff = (long)(outcol ? outcol[ar] : ar) * ldc;
#pragma clang loop unroll(full)
            for (int j = 0; j < 4; ++j) {
                int bc = b0 + wb + j * 16 + fr;
                if (bc < N) {
                    double d = acc[i][j][vv];
                    double *cp = C + rowoff + bc;
                    if (mode == 1) {
                        d = fma(sig[bc], *cp, d);
                    }
                    *cp = d;
                }
            }
        }
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

// Fill a row-major block with zeros (used instead of hipMemsetAsync so the sequence can be
// captured in a HIP graph).
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
template __global__ void k_solve_real<65>(const double *, long, const double *, long, int, int, double *, long,
                                          double *, long, const ColDesc *, const int2 *, int, int, int, int,
                                          double *, int *, int *, int *, int *, int, int, double *, int);
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

// One workgroup per (batch b, forward column c): RHS of the adjoint solve in working rows.
__global__ __launch_bounds__(256) void k_bwd_rhs(const double *__restrict__ V, long ldv, const double *__restrict__ G,
                                                 long ldg, int n, int m, int nbatch, const int *__restrict__ kcol,
                                                 double *__restrict__ X, long ldx)
{
    __shared__ double rv[256 / WAVE];
    __shared__ int ri[256 / WAVE];
    __shared__ double rd[256 / WAVE];
    const int c = blockIdx.x % m, b = blockIdx.x / m;
    const double *v = V + (long)c * ldv;
    const double *g = G + (long)(b * m + c) * ldg;
    const int k = kcol[c];
    const int q = (m - 1 - c) * nbatch + b;
    double best = -1.0, dot = 0.0;
    int bi = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double a = fabs(v[i]);
        if (a > best) {
            best = a;
            bi = i;
        }
        dot = fma(g[i], v[i], dot);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off);
        int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
        dot += __shfl_xor(dot, off);
    }
    if (threadIdx.x % WAVE == 0) {
        rv[threadIdx.x / WAVE] = best;
        ri[threadIdx.x / WAVE] = bi;
        rd[threadIdx.x / WAVE] = dot;
    }
    __syncthreads();
    best = rv[0];
    bi = ri[0];
    dot = rd[0];
    for (int w = 1; w < (int)(blockDim.x / WAVE); ++w) {
        if (rv[w] > best || (rv[w] == best && ri[w] < bi)) {
            best = rv[w];
            bi = ri[w];
        }
        dot += rd[w];
    }
    const double sp = v[bi] < 0.0 ? -1.0 : 1.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double val = 0.0;
        if (i < k) {
            val = g[i] - (i == bi ? dot * sp : 0.0);
        }
        X[(long)(n - 1 - i) * ldx + q] = val;
    }
}

// Lam[(b*m + c) * n + i] = coef * 2^-e * lambda_i  for coef = -a_c (S part) and b_c (P part);
// 32x32 LDS tiles over (working column q, row i).
__global__ void k_bwd_lambda_t(const double *__restrict__ X, long ldx, int n, int m, int nbatch,
                               const int *__restrict__ kcol, const double *__restrict__ coef,
                               const int *__restrict__ colexp, double *__restrict__ LamS, double *__restrict__ LamP)
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
                v = ldexp(X[(long)(n - 1 - i) * ldx + q], -colexp[q]);
            }
        }
        tile[kk][tx] = v;
    }
    __syncthreads();
    for (int kk = ty; kk < 32; kk += 8) {
        int q = q0 + kk, i = i0 + tx;
        if (i < n && q < M) {
            int c = m - 1 - q / nbatch, b = q % nbatch;
            double v = tile[tx][kk];
            long o = (long)(b * m + c) * n + i;
            LamS[o] = -coef[4 * c + 0] * v;
            LamP[o] = coef[4 * c + 1] * v;
        }
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

// One wave per (working column q): dL/dalpha and dL/dbeta folded into dS_kk, dP_kk.
//   Y = P V (row-major, ldy), lambda_i = 2^-e X[n-1-i][q].
__global__ __launch_bounds__(64) void k_bwd_diag(const double *__restrict__ X, long ldx, const double *__restrict__ Y,
                                                 long ldy, int n, int m, int nbatch, const int *__restrict__ kcol,
                                                 const double *__restrict__ coef, const int *__restrict__ colexp,
                                                 double *__restrict__ dS, long ldds, double *__restrict__ dP, long lddp)
{
    const int q = blockIdx.x;
    const int c = m - 1 - q / nbatch, b = q % nbatch;
    const int k = kcol[c];
    double acc = 0.0;
    for (int i = threadIdx.x; i < k; i += WAVE) {
        acc = fma(X[(long)(n - 1 - i) * ldx + q], Y[(long)i * ldy + c], acc);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        acc += __shfl_xor(acc, off);
    }
    if (threadIdx.x == 0) {
        double qv = ldexp(acc, -colexp[q]);
        double s = coef[4 * c + 2], ratio = coef[4 * c + 3];
        dS[(long)b * n * ldds + k + (long)k * ldds] += s * qv;
        dP[(long)b * n * lddp + k + (long)k * lddp] += -s * ratio * qv;
    }
}

// Host launcher: picks the 16-byte-load form when both output extents are >= 2.
// tag only separates the uses in profiles: 0 = RHS update, 1 = back-transform, 2 = backward.
void launch_gemm_f64(int tag, hipStream_t st, int M, int N, int K1, int K2, const double *L, long ldl, const double *L2,
                     long ldl2, const double *R, long ldr, double *C, long ldc, const int *outcol, int mode,
                     const double *sig, const int *kmax, const double *R2)
{
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    const bool wide = M >= 2 && N >= 2;
#define XTG_GEMM_CASE(T)                                                                                             \
    if (tag == T) {                                                                                                  \
        if (wide) {                                                                                                  \
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_f64<T, true>), grid, dim3(256), 0, st, M, N, K1, K2, L, ldl, L2, \
                               ldl2, R, ldr, C, ldc, outcol, mode, sig, kmax, R2);                                   \
        }                                                                                                            \
        else {                                                                                                       \
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_f64<T, false>), grid, dim3(256), 0, st, M, N, K1, K2, L, ldl,  \
                               L2, ldl2, R, ldr, C, ldc, outcol, mode, sig, kmax, R2);                               \
        }                                                                                                            \
        return;                                                                                                      \
    }
    XTG_GEMM_CASE(0)
    XTG_GEMM_CASE(1)
    XTG_GEMM_CASE(2)
#undef XTG_GEMM_CASE
}
} // namespace xtg
