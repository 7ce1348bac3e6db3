// gfx950 kernels of the complex128 xTGEVC3 path (ztgevc3 / ctgevc3 interface).
//
//   kz_solve_units -- one block-row step: per-column shifted triangular solve of a diagonal
//                   block (complex, no 2x2 bumps), 16 lanes per eigenvector column, S/P
//                   block in LDS, per-column power-of-two overflow scaling, group operands
//                   of the near/far updates (replaces zlalsr/zlalsl + ztrsm_, reference
//                   src/ztgevc3.cpp:116-358)
//   kz_gemm      -- complex FP64 MFMA GEMM (four v_mfma_f64_16x16x4 per complex product) for
//                   the RHS update (ztgevc3.cpp:664-665, :808-809) and the back-transform
//                   (ztgevc3.cpp:676, :820)
//   kz_*         -- fix-up of the scale history, copy-out, normalisation, conjugate
//                   reversed transpose, column norms.
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.h"

namespace xtg {

struct zc {
    double r, i;
};

__device__ __forceinline__ zc zmul(zc a, zc b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ double zabs1(zc a) { return fabs(a.r) + fabs(a.i); }

constexpr double kzEps = 2.220446049250313e-16;
constexpr double kzSafeMin = 2.2250738585072014e-308;

// ---------------------------------------------------------------------------------------
// kz_solve_units: one block-row step of the complex sweep, 16 lanes per eigenvector column.
// Lane j of a column keeps rows 16 i + j of the block.  Before the sweep every lane
// computes the reciprocal of the shifted diagonal of its rows (a S_rr - b P_rr, perturbed to
// the rounding level when tiny; rows outside the column's system get 0, its own row 1).  The
// sweep runs bottom-up: the owner lane forms x_r = rhs_r / c_r, broadcasts it in the 16-lane
// row (DPP row_newbcast), and every lane updates its rows above with the {S,P} entries of
// column r, fetched from LDS one row ahead.  A solution entry above 2^958 rescales the column
// by a power of two down to 2^446 (exponent bookkeeping as in the real sweep).
// Operands of the group (rows w_s + r: -a x, rows w_p + r: b x); a column starting in this
// block (not the group's first step) zeros its operand rows of the group's lower blocks; the
// group's last step brings the lower blocks' operand rows to the column's final scale.
// X, W: complex (double2) row-major, X[row * ldx + col].  sq (split-int8 form, bexp ==
// nullptr: off): block exponents of the operand and of X per step, the group operand's at the
// group's last step, as in the real solve kernel.
// ---------------------------------------------------------------------------------------
constexpr int ZQL = 16;
constexpr int ZSP = 33; // LDS column pitch of the staged block (rows)

template <int SRC> __device__ __forceinline__ double zrow_bcast_c(double v)
{
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + SRC, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + SRC, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int SRC> __device__ __forceinline__ int zrow_bcast_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + SRC, 0xF, 0xF, false);
}
#define XTG_ZRB_CASES(F)                                                                                         \
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
__device__ __forceinline__ double zbd(double v, int src) { XTG_ZRB_CASES(zrow_bcast_c) }
__device__ __forceinline__ int zbi(int v, int src) { XTG_ZRB_CASES(zrow_bcast_i) }
#undef XTG_ZRB_CASES
__device__ __forceinline__ zc zbc(zc v, int src) { return {zbd(v.r, src), zbd(v.i, src)}; }
__device__ __forceinline__ zc zsel(bool c, zc x, zc y) { return {c ? x.r : y.r, c ? x.i : y.i}; }
__device__ __forceinline__ zc zrecip(zc c)
{
    const bool sw = fabs(c.i) > fabs(c.r);
    const double big = sw ? c.i : c.r, sml = sw ? c.r : c.i;
    const double t = sml / big;
    const double rd = 1.0 / (big + sml * t);
    const double tr = t * rd;
    return sw ? zc{tr, -rd} : zc{rd, -tr};
}

template <int MAXB>
__global__ __launch_bounds__(256) void kz_solve_units(
    const double2 *__restrict__ S, long lds, const double2 *__restrict__ P, long ldp, int i0, int nb,
    double2 *__restrict__ X, long ldx, double2 *__restrict__ W, long ldw, const ZColDesc *__restrict__ desc,
    int c_begin, int m, double *__restrict__ sig, int *__restrict__ colexp, int *__restrict__ blkexp,
    int *__restrict__ flag, int w_s, int w_p, double *__restrict__ sig_grp, int first_in_group, GroupInfo grp,
    SplitOut sq)
{
    static_assert(MAXB <= 32 && ZSP > MAXB, "block layout");
    constexpr int QR = (MAXB + ZQL - 1) / ZQL;
    __shared__ double2 ss[ZSP * MAXB], pp[ZSP * MAXB]; // ss[l + r * ZSP] = S_lr
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int j = lane & (ZQL - 1);
    const int p = c_begin + (int)(blockIdx.x * (blockDim.x / WAVE) + wave) * (WAVE / ZQL) + lane / ZQL;
    const bool act = p < m;
    const ZColDesc cd = desc[act ? p : c_begin];
    const zc a = {cd.aR, cd.aI}, b = {cd.bR, cd.bI};
    const double an1 = zabs1(a), bn1 = zabs1(b);
    const int dl = act ? cd.d - i0 : -1 - MAXB;
    const int lol = cd.lo - i0;
    const int init = (cd.flags >> 2) & 3;

    // right-hand side of the column's rows (issued before the staging)
    zc vr[QR];
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = ZQL * i + j;
        vr[i] = {0.0, 0.0};
        if (r < nb && r < dl && r >= lol) {
            const double2 t = X[(long)(i0 + r) * ldx + p];
            vr[i] = {t.x, t.y};
        }
    }
    for (int q = threadIdx.x; q < nb * nb; q += blockDim.x) {
        const int l = q % nb, r = q / nb;
        ss[l + r * ZSP] = S[(i0 + l) + (long)(i0 + r) * lds];
        pp[l + r * ZSP] = P[(i0 + l) + (long)(i0 + r) * ldp];
    }
    __syncthreads();

    // reciprocals of the shifted diagonal of this lane's rows
    zc fa[QR];
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = ZQL * i + j;
        fa[i] = {0.0, 0.0};
        if (r < nb) {
            const double2 es = ss[r + r * ZSP], ep = pp[r + r * ZSP];
            zc c = {a.r * es.x - a.i * es.y - (b.r * ep.x - b.i * ep.y), a.r * es.y + a.i * es.x - (b.r * ep.y + b.i * ep.x)};
            const double dmin =
                fmax(kzEps * (an1 * (fabs(es.x) + fabs(es.y)) + bn1 * (fabs(ep.x) + fabs(ep.y))), kzSafeMin);
            if (zabs1(c) < dmin) {
                c = {dmin, 0.0};
            }
            const bool solve = r < dl && r >= lol;
            const bool own = r == dl && init == INIT_UNIT;
            fa[i] = zsel(solve, zrecip(c), zc{own ? 1.0 : 0.0, 0.0});
            if (own) {
                vr[i] = {1.0, 0.0};
            }
        }
    }

    int my_exp = 0;
    double2 es_c[QR], ep_c[QR], es_n[QR], ep_n[QR]; // column r (c) and r-1 (n) of S, P at this lane's rows
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        es_c[i] = es_n[i] = ep_c[i] = ep_n[i] = make_double2(0.0, 0.0);
    }
#pragma clang loop unroll(full)
    for (int r = MAXB - 1; r >= 0; --r) {
        if (r < nb) {
            if (r == nb - 1) {
#pragma clang loop unroll(full)
                for (int i = 0; i < QR; ++i) {
                    if (ZQL * i < r) {
                        es_c[i] = ss[ZQL * i + j + r * ZSP];
                        ep_c[i] = pp[ZQL * i + j + r * ZSP];
                    }
                }
            }
            if (r >= 1) {
#pragma clang loop unroll(full)
                for (int i = 0; i < QR; ++i) {
                    if (ZQL * i < r - 1) {
                        es_n[i] = ss[ZQL * i + j + (r - 1) * ZSP];
                        ep_n[i] = pp[ZQL * i + j + (r - 1) * ZSP];
                    }
                }
            }
            const int jr = r & (ZQL - 1), ir = r / ZQL;
            zc x = zbc(zmul(vr[ir], fa[ir]), jr);
            const bool bad = !(zabs1(x) <= 0x1p958);
            if (__any(bad)) {
                int eo = 0;
                const double rn = zabs1(vr[ir]), fn = zabs1(fa[ir]);
                if (bad && rn > 0.0 && fn > 0.0) {
                    eo = max(0, ilogb(rn) + ilogb(fn) + 2 - 446);
                }
                const int e = zbi(eo, jr);
                const double f = ldexp(1.0, -e);
#pragma clang loop unroll(full)
                for (int i = 0; i < QR; ++i) {
                    vr[i] = {vr[i].r * f, vr[i].i * f};
                }
                my_exp -= e;
                x = zbc(zmul(vr[ir], fa[ir]), jr);
            }
            vr[ir] = zsel(j == jr, x, vr[ir]);
            const zc ax = zmul(a, x), bx = zmul(b, x);
#pragma clang loop unroll(full)
            for (int i = 0; i < QR; ++i) {
                if (ZQL * i < r) {
                    const double2 s2 = es_c[i], p2 = ep_c[i];
                    // v_l -= S_lr * ax - P_lr * bx
                    double nr = fma(-s2.x, ax.r, vr[i].r);
                    nr = fma(s2.y, ax.i, nr);
                    nr = fma(p2.x, bx.r, nr);
                    nr = fma(-p2.y, bx.i, nr);
                    double ni = fma(-s2.x, ax.i, vr[i].i);
                    ni = fma(-s2.y, ax.r, ni);
                    ni = fma(p2.x, bx.i, ni);
                    ni = fma(p2.y, bx.r, ni);
                    const bool upd = ZQL * i + ZQL - 1 < r || ZQL * i + j < r;
                    vr[i] = zsel(upd, zc{nr, ni}, vr[i]);
                }
            }
#pragma clang loop unroll(full)
            for (int i = 0; i < QR; ++i) {
                es_c[i] = es_n[i];
                ep_c[i] = ep_n[i];
            }
        }
    }

    // split-int8 form: exponents of the block's operand and solution (exact maxima of |Re|,
    // |Im| over the column's 16 lanes)
    double mw = 0.0, mxv = 0.0;
    if (sq.bexp != nullptr) {
#pragma clang loop unroll(full)
        for (int i = 0; i < QR; ++i) {
            if (ZQL * i + j < nb) {
                const zc x = vr[i];
                const zc ax = zmul(a, x), bx = zmul(b, x);
                mxv = fmax(mxv, fmax(fabs(x.r), fabs(x.i)));
                mw = fmax(mw, fmax(fmax(fabs(ax.r), fabs(ax.i)), fmax(fabs(bx.r), fabs(bx.i))));
            }
        }
#pragma clang loop unroll(full)
        for (int off = 1; off < ZQL; off <<= 1) {
            mw = fmax(mw, __shfl_xor(mw, off));
            mxv = fmax(mxv, __shfl_xor(mxv, off));
        }
    }
    if (!act) {
        return;
    }
#pragma clang loop unroll(full)
    for (int i = 0; i < QR; ++i) {
        const int r = ZQL * i + j;
        if (r < nb) {
            const zc x = vr[i];
            X[(long)(i0 + r) * ldx + p] = make_double2(x.r, x.i);
            const zc ax = zmul(a, x), bx = zmul(b, x);
            W[(long)(w_s + r) * ldw + p] = make_double2(-ax.r, -ax.i);
            W[(long)(w_p + r) * ldw + p] = make_double2(bx.r, bx.i);
        }
    }
    const bool starts_here = dl >= 0 && dl < nb;
    const int ce = colexp[p] + my_exp;
    if (starts_here && !first_in_group) {
        for (int k = w_s + nb + j; k < grp.wr; k += ZQL) {
            W[(long)k * ldw + p] = make_double2(0.0, 0.0);
            W[(long)(grp.wr + k) * ldw + p] = make_double2(0.0, 0.0);
        }
    }
    if (grp.last && !starts_here) {
        for (int jb = 1; jb < 8; ++jb) {
            const int k0 = grp.off[jb], k1 = grp.off[jb + 1];
            if (k0 < k1) {
                const int e = ce - blkexp[(long)(grp.step[jb] - grp.step[0]) * ldx + p];
                if (e != 0) {
                    for (int k = k0 + j; k < k1; k += ZQL) {
                        double2 *w0 = W + (long)k * ldw + p;
                        double2 *w1 = W + (long)(grp.wr + k) * ldw + p;
                        const double2 v0 = *w0, v1 = *w1;
                        *w0 = make_double2(ldexp(v0.x, e), ldexp(v0.y, e));
                        *w1 = make_double2(ldexp(v1.x, e), ldexp(v1.y, e));
                    }
                }
            }
        }
    }
    if (j == 0 && sq.bexp != nullptr) {
        const int ew = mw > 0.0 ? min(max(ilogb(mw) + 1, -1000), 1000) : -1000;
        sq.bexp[p] = ew;
        sq.xbexp[p] = mxv > 0.0 ? min(max(ilogb(mxv) + 1, -1000), 1000) : -1000;
        if (grp.last) {
            // group operand exponent: this block and the lower blocks the column reaches, each
            // at the final scale (lower blocks were brought there by 2^(ce - blkexp) above)
            int eg = ew;
            for (int jb = 1; jb < GROUP_MAX_BLOCKS; ++jb) {
                if (grp.off[jb] < grp.off[jb + 1] && i0 + grp.off[jb] <= cd.d) {
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
    if (j == 0) {
        const double sg = ldexp(1.0, my_exp);
        sig[p] = sg;
        sig_grp[p] = (first_in_group || starts_here) ? sg : sig_grp[p] * sg;
        colexp[p] = ce;
        blkexp[p] = ce;
        if (my_exp != 0) {
            *flag = 1;
        }
    }
}

template __global__ void kz_solve_units<32>(const double2 *, long, const double2 *, long, int, int, double2 *, long,
                                            double2 *, long, const ZColDesc *, int, int, double *, int *, int *,
                                            int *, int, int, double *, int, GroupInfo, SplitOut);

// ---------------------------------------------------------------------------------------
// kz_gemm: D[a][b] = sum_k L[k][a] * R[k][b], complex, FP64 MFMA; rows k >= K1 of L and R
// come from the second segments L2 and R2 (row k - K1).
// Tile TA (a) x 64 (b) per 256-thread workgroup (waves 2x2, each TA/2 x 32), K chunks of 8.
// Same operand / output conventions as k_gemm_f64 (element = complex double2).
// ---------------------------------------------------------------------------------------
typedef double zdouble4_t __attribute__((ext_vector_type(4)));
constexpr int ZTB = 64, ZGK = 8;

// TA: tile rows (128, or 32 for short products over few rows)
template <int TAG, int TA>
__global__ __launch_bounds__(256) void kz_gemm(int M, int N, int K1, int K2, const double2 *__restrict__ L, long ldl,
                                               const double2 *__restrict__ L2, long ldl2, const double2 *__restrict__ R,
                                               long ldr, double2 *__restrict__ C, long ldc,
                                               const int *__restrict__ outcol, int mode, const double *__restrict__ sig,
                                               const int *__restrict__ kmax, const double2 *__restrict__ R2)
{
    constexpr int AI = TA / 32;       // 16-row MFMA blocks per wave (waves 2 x 2)
    constexpr int LS = 256 / TA;      // k rows of an L chunk loaded per pass
    __shared__ double Lr[2][ZGK][TA], Li[2][ZGK][TA];
    __shared__ double Rr[2][ZGK][ZTB], Ri[2][ZGK][ZTB];
    int tm, tn;
    xcd_tile(blockIdx.x, (M + TA - 1) / TA, (N + ZTB - 1) / ZTB, tm, tn);
    const int a0 = tm * TA, b0 = tn * ZTB;
    const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
    const int wa = (wave >> 1) * (TA / 2), wb = (wave & 1) * 32;
    int K = K1 + K2;
    if (kmax != nullptr) {
        int kk = 0;
        int amax = min(a0 + TA, M) - 1;
        for (int q = a0; q <= amax; ++q) {
            kk = max(kk, kmax[q]);
        }
        K = min(K, kk);
    }
    zdouble4_t accr[AI][2], acci[AI][2];
#pragma clang loop unroll(full)
    for (int i = 0; i < AI; ++i) {
#pragma clang loop unroll(full)
        for (int j = 0; j < 2; ++j) {
            accr[i][j] = (zdouble4_t){0, 0, 0, 0};
            acci[i][j] = (zdouble4_t){0, 0, 0, 0};
        }
    }
    // staging: L chunk 8 x TA complex (TA/32 per thread), R chunk 8 x 64 complex (2 per thread)
    const int lc = tid % TA, lk = tid / TA;
    const int rc = tid % ZTB, rk = tid / ZTB;  // rk in {0..3}
    double2 lreg[AI], rreg[2];
    auto load_chunk = [&](int k0) {
#pragma clang loop unroll(full)
        for (int s = 0; s < AI; ++s) {
            int k = k0 + lk + LS * s;
            int aa = a0 + lc;
            double2 v = make_double2(0.0, 0.0);
            if (k < K && aa < M) {
                v = (k < K1) ? L[aa + (long)k * ldl] : L2[aa + (long)(k - K1) * ldl2];
            }
            lreg[s] = v;
        }
#pragma clang loop unroll(full)
        for (int s = 0; s < 2; ++s) {
            int k = k0 + rk + 4 * s;
            int bb = b0 + rc;
            double2 v = make_double2(0.0, 0.0);
            if (k < K && bb < N) {
                v = (k < K1) ? R[(long)k * ldr + bb] : R2[(long)(k - K1) * ldr + bb];
            }
            rreg[s] = v;
        }
    };
    auto store_chunk = [&](int buf) {
#pragma clang loop unroll(full)
        for (int s = 0; s < AI; ++s) {
            Lr[buf][lk + LS * s][lc] = lreg[s].x;
            Li[buf][lk + LS * s][lc] = lreg[s].y;
        }
#pragma clang loop unroll(full)
        for (int s = 0; s < 2; ++s) {
            Rr[buf][rk + 4 * s][rc] = rreg[s].x;
            Ri[buf][rk + 4 * s][rc] = rreg[s].y;
        }
    };
    const int nch = (K + ZGK - 1) / ZGK;
    if (nch > 0) {
        load_chunk(0);
        store_chunk(0);
    }
    __syncthreads();
    const int fr = lane % 16, fk = lane / 16;
    for (int ch = 0; ch < nch; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nch) {
            load_chunk((ch + 1) * ZGK);
        }
#pragma clang loop unroll(full)
        for (int ks = 0; ks < ZGK; ks += 4) {
            double ar_[AI], ai_[AI], br_[2], bi_[2];
#pragma clang loop unroll(full)
            for (int i = 0; i < AI; ++i) {
                ar_[i] = Lr[buf][ks + fk][wa + i * 16 + fr];
                ai_[i] = Li[buf][ks + fk][wa + i * 16 + fr];
            }
#pragma clang loop unroll(full)
            for (int j = 0; j < 2; ++j) {
                br_[j] = Rr[buf][ks + fk][wb + j * 16 + fr];
                bi_[j] = Ri[buf][ks + fk][wb + j * 16 + fr];
            }
#pragma clang loop unroll(full)
            for (int i = 0; i < AI; ++i) {
                double nai = -ai_[i];
#pragma clang loop unroll(full)
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar_[i], br_[j], accr[i][j], 0, 0, 0);
                    accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi_[j], accr[i][j], 0, 0, 0);
                    acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar_[i], bi_[j], acci[i][j], 0, 0, 0);
                    acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai_[i], br_[j], acci[i][j], 0, 0, 0);
                }
            }
        }
        if (ch + 1 < nch) {
            store_chunk(buf ^ 1);
        }
        __syncthreads();
    }
#pragma clang loop unroll(full)
    for (int i = 0; i < AI; ++i) {
#pragma clang loop unroll(full)
        for (int vv = 0; vv < 4; ++vv) {
            int aa = a0 + wa + i * 16 + fk + 4 * vv;
            if (aa >= M) {
                continue;
            }
            long rowoff = (long)(outcol ? outcol[aa] : aa) * ldc;
#pragma clang loop unroll(full)
            for (int j = 0; j < 2; ++j) {
                int bb = b0 + wb + j * 16 + fr;
                if (bb < N) {
                    double dr = accr[i][j][vv], di = acci[i][j][vv];
                    double2 *cp = C + rowoff + bb;
                    if (mode == 1) {
                        double2 o = *cp;
                        double s = sig[bb];
                        dr = fma(s, o.x, dr);
                        di = fma(s, o.y, di);
                    }
                    *cp = make_double2(dr, di);
                }
            }
        }
    }
}

#define XTG_KZ_INST(T, TA)                                                                                       \
    template __global__ void kz_gemm<T, TA>(int, int, int, int, const double2 *, long, const double2 *, long,          \
                                            const double2 *, long, double2 *, long, const int *, int, const double *,  \
                                            const int *, const double2 *);
XTG_KZ_INST(0, 128)
XTG_KZ_INST(0, 32)
XTG_KZ_INST(1, 128)
#undef XTG_KZ_INST

__global__ void kz_fixup_scale(double2 *__restrict__ X, long ldx, int n, int m, const int *__restrict__ rowstep,
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
    int e = colexp[p] - blkexp[(long)rowstep[r] * ldb + p];
    if (e != 0) {
        double2 *x = X + (long)r * ldx + p;
        double2 v = *x;
        *x = make_double2(ldexp(v.x, e), ldexp(v.y, e));
    }
}

__global__ void kz_copy_out(const double2 *__restrict__ X, long ldx, int n, int m, double2 *__restrict__ V, long ldv,
                            const int *__restrict__ outcol, int rev)
{
    __shared__ double2 tile[32][33];
    int p0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;
    for (int k = ty; k < 32; k += 8) {
        int i = i0 + k, p = p0 + tx;
        double2 v = make_double2(0.0, 0.0);
        if (i < n && p < m) {
            v = X[(long)(rev ? n - 1 - i : i) * ldx + p];
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

__global__ __launch_bounds__(256) void kz_normalize(double2 *__restrict__ V, long ldv, int n)
{
    __shared__ double red[256 / WAVE];
    double2 *c0 = V + (long)blockIdx.x * ldv;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double2 t = c0[i];
        mx = fmax(mx, fabs(t.x) + fabs(t.y));
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
    if (!(mx > kzSafeMin)) {
        return;
    }
    double sc = 1.0 / mx;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double2 t = c0[i];
        c0[i] = make_double2(t.x * sc, t.y * sc);
    }
}

// D = J A^H J: D[i + j*ldd] = conj(A[(n-1-j) + (n-1-i)*lda])
__global__ void kz_reverse_conjtranspose(const double2 *__restrict__ A, long lda, int n, double2 *__restrict__ D,
                                         long ldd)
{
    __shared__ double2 tile[32][33];
    int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    int tx = threadIdx.x, ty = threadIdx.y;
    for (int k = ty; k < 32; k += 8) {
        int j = j0 + tx, i = i0 + k;
        double2 v = make_double2(0.0, 0.0);
        if (i < n && j < n) {
            v = A[(long)(n - 1 - j) + (long)(n - 1 - i) * lda];
        }
        tile[k][tx] = make_double2(v.x, -v.y);
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int i = i0 + tx, j = j0 + k;
        if (i < n && j < n) {
            D[i + (long)j * ldd] = tile[tx][k];
        }
    }
}

// column sums of |S| and |P| (modulus) over rows 0..c (ztgevc3.cpp:528-542)
__global__ void kz_colnorms(const double2 *__restrict__ S, long lds, const double2 *__restrict__ P, long ldp, int n,
                            double *__restrict__ out)
{
    int c = blockIdx.x;
    double ss = 0.0, sp = 0.0;
    for (int r = threadIdx.x; r <= c; r += blockDim.x) {
        double2 a = S[r + (long)c * lds], b = P[r + (long)c * ldp];
        ss += hypot(a.x, a.y);
        sp += hypot(b.x, b.y);
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

// ---------------------------------------------------------------------------------------
// Eigenvalue condition numbers of the complex pencil (LAPACK ztgsna, JOB = 'E').
// kz_to_rowmajor: Vt[k * ldt + c] = V[k + c * ldv] (32 x 32 tiles).
// kz_tgsna: one wave per column c of the column-major n x m arrays VL, VR, YA = S VR,
// YB = P VR: s[c] = hypot(|VL_c^H YA_c|, |VL_c^H YB_c|) / (|VL_c| |VR_c|).
// ---------------------------------------------------------------------------------------
__global__ void kz_to_rowmajor(const double2 *__restrict__ V, long ldv, int n, int m, double2 *__restrict__ Vt,
                               long ldt)
{
    __shared__ double2 tile[32][33];
    const int c0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int kk = ty; kk < 32; kk += 8) {
        const int c = c0 + kk, k = k0 + tx;
        tile[kk][tx] = (c < m && k < n) ? V[k + (long)c * ldv] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int kk = ty; kk < 32; kk += 8) {
        const int k = k0 + kk, c = c0 + tx;
        if (c < m && k < n) {
            Vt[(long)k * ldt + c] = tile[tx][kk];
        }
    }
}

__global__ __launch_bounds__(64) void kz_tgsna(const double2 *__restrict__ VL, long ldvl,
                                               const double2 *__restrict__ VR, long ldvr,
                                               const double2 *__restrict__ YA, const double2 *__restrict__ YB,
                                               long ldy, int n, double *__restrict__ s)
{
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0, nu = 0.0, nv = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double2 u = VL[i + (long)c * ldvl], v = VR[i + (long)c * ldvr];
        const double2 ya = YA[i + (long)c * ldy], yb = YB[i + (long)c * ldy];
        // conj(u) * y
        ar = fma(u.x, ya.x, fma(u.y, ya.y, ar));
        ai = fma(u.x, ya.y, fma(-u.y, ya.x, ai));
        br = fma(u.x, yb.x, fma(u.y, yb.y, br));
        bi = fma(u.x, yb.y, fma(-u.y, yb.x, bi));
        nu = fma(u.x, u.x, fma(u.y, u.y, nu));
        nv = fma(v.x, v.x, fma(v.y, v.y, nv));
    }
#pragma clang loop unroll(full)
    for (int off = 32; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off);
        ai += __shfl_xor(ai, off);
        br += __shfl_xor(br, off);
        bi += __shfl_xor(bi, off);
        nu += __shfl_xor(nu, off);
        nv += __shfl_xor(nv, off);
    }
    if (lane == 0) {
        s[c] = hypot(hypot(ar, ai), hypot(br, bi)) / (sqrt(nu) * sqrt(nv));
    }
}

} // namespace xtg

