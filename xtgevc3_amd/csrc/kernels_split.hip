// gfx950 kernels of the split-int8 form of the real path (XTGEVC3_PREC_SPLIT_I8).
//
// A value v of a row (or column) with exponent E (2^E > max |row|) is written as five
// balanced base-255 digits d_0..d_4 in [-127, 127] (split_digits):
//     v = 2^(E+1) * sum_s d_s 255^-(s+1) + r,   |r| <= 2^(E+1) 255^-5 / 2  (about 2^-41 2^E).
// A product sum_k L(a,k) R(b,k) then becomes, with exact int32 MFMA accumulation,
//     2^(EL_a + ER_b + 2) * sum_t 255^-(t+2) acc_t,    acc_t = sum_{i+j=t} sum_k dL_i(a,k) dR_j(b,k),
// keeping the levels t = i + j <= 4 (15 digit products).  The dropped levels and the digit
// remainders bound the error by about 2^-39 * K * max|L(a,:)| * max|R(b,:)|; |acc_t| stays
// below 5 * 127^2 * K, exact in int32 for K <= SPLIT_KMAX (longer products run in chunks).
//
//   k_slice_digits -- digits of a matrix (real, or complex expanded along k), k-blocked
//   kz_rowexp      -- row exponents of complex matrices
//   k_gemm_i8      -- C = sig * C + sum_k L(a,k) R(b,k) on v_mfma_i32_32x32x32_i8, digits
//                     staged through LDS, FP64 epilogue
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>

#include "common.h"

namespace xtg {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// Digit layout ("k-blocked"): digit s of element (a, k) at Q[s * qs + (k / 32) * kbp + a * 32 +
// k % 32], kbp = 32 * (rows of the array): a K step of 32 over consecutive rows is one
// contiguous run of 32-byte rows.
//
// Q = digits of A(a, k) * 2^-(E[a] + 1) for k in [0, Kp), A(a, k) = A1[a + (k - o1) * lda1] for
// o1 <= k < o1 + K1, A2[a + (k - o2) * lda2] for o2 <= k < o2 + K2, zero elsewhere.
// Complex operands (cmode != 0; A1, A2 hold double2 and o1, K1, o2, K2 count complex
// entries) are expanded along k: position k holds part k % 2 of complex entry k / 2 as
//     cmode 1: (re, -im),   cmode 2: (im, re),   cmode 3: (re, im),
// so that a complex product becomes two real ones, Re = [Lr, -Li] . [Rr, Ri] and
// Im = [Li, Lr] . [Rr, Ri].
// With Q2 != nullptr (cmode 1 only) the same pass also writes the (im, re) expansion to Q2:
// its digits at positions (2k, 2k+1) are (-d(2k+1), d(2k)) of the (re, -im) digits, exact
// because the digits of -v are those of v negated.
// Block: 64 rows a x 16 k, thread t: row t / 4, k quad 4 (t % 4) (many small blocks: the
// kernel is latency-bound at the sizes of the group operands).
__global__ __launch_bounds__(256) void k_slice_digits(const double *__restrict__ A1, long lda1, int o1, int K1,
                                                      const double *__restrict__ A2, long lda2, int o2, int K2, int M,
                                                      int Kp, const int *__restrict__ E, signed char *__restrict__ Q,
                                                      long qs, long kbp, int cmode, signed char *__restrict__ Q2)
{
    const int a = blockIdx.x * 64 + threadIdx.x / 4;
    const int k0 = blockIdx.y * 16 + (threadIdx.x % 4) * 4;
    if (a >= M || k0 >= Kp) {
        return;
    }
    const int e = E[a] + 1;
    int dg[SPLIT_DIGITS][4];
#pragma clang loop unroll(full)
    for (int q = 0; q < 4; ++q) {
        const int kk = k0 + q;
        double v = 0.0;
        if (cmode == 0) {
            if (kk >= o1 && kk < o1 + K1) {
                v = A1[a + (long)(kk - o1) * lda1];
            }
            else if (kk >= o2 && kk < o2 + K2) {
                v = A2[a + (long)(kk - o2) * lda2];
            }
        }
        else {
            const int k = kk >> 1, part = kk & 1;
            double2 z = make_double2(0.0, 0.0);
            if (k >= o1 && k < o1 + K1) {
                z = reinterpret_cast<const double2 *>(A1)[a + (long)(k - o1) * lda1];
            }
            else if (k >= o2 && k < o2 + K2) {
                z = reinterpret_cast<const double2 *>(A2)[a + (long)(k - o2) * lda2];
            }
            v = cmode == 1 ? (part ? -z.y : z.x) : (cmode == 2 ? (part ? z.x : z.y) : (part ? z.y : z.x));
        }
        int d[SPLIT_DIGITS];
        split_digits(ldexp(v, -e), d);
#pragma clang loop unroll(full)
        for (int s = 0; s < SPLIT_DIGITS; ++s) {
            dg[s][q] = d[s];
        }
    }
    const long qo = (long)(k0 / 32) * kbp + (long)a * 32 + (k0 % 32);
#pragma clang loop unroll(full)
    for (int s = 0; s < SPLIT_DIGITS; ++s) {
        *reinterpret_cast<int *>(Q + qo + s * qs) =
            (dg[s][0] & 0xff) | ((dg[s][1] & 0xff) << 8) | ((dg[s][2] & 0xff) << 16) | ((dg[s][3] & 0xff) << 24);
    }
    if (Q2 != nullptr) {
#pragma clang loop unroll(full)
        for (int s = 0; s < SPLIT_DIGITS; ++s) {
            *reinterpret_cast<int *>(Q2 + qo + s * qs) = ((-dg[s][1]) & 0xff) | ((dg[s][0] & 0xff) << 8) |
                                                         (((-dg[s][3]) & 0xff) << 16) | ((dg[s][2] & 0xff) << 24);
        }
    }
}

// out[r] = max(out[r], ilogb(max |Re|, |Im| over the chunk) + 1) for the complex rows r of A
// and B (B may be null), A(r, k) = A[r + k * lda]; chunks of 64 columns per blockIdx.y, and
// only columns k >= cfirst[r] when cfirst is given.
__global__ __launch_bounds__(256) void kz_rowexp(const double2 *__restrict__ A, long lda, const double2 *__restrict__ B,
                                                 long ldb, int n, int ncols, int *__restrict__ out,
                                                 const int *__restrict__ cfirst)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = blockIdx.y * 64;
    if (r >= n) {
        return;
    }
    const int k1 = min(ncols, k0 + 64);
    double mx = 0.0;
    for (int k = cfirst != nullptr ? max(k0, cfirst[r]) : k0; k < k1; ++k) {
        const double2 z = A[r + (long)k * lda];
        mx = fmax(mx, fmax(fabs(z.x), fabs(z.y)));
        if (B != nullptr) {
            const double2 w = B[r + (long)k * ldb];
            mx = fmax(mx, fmax(fabs(w.x), fabs(w.y)));
        }
    }
    if (mx > 0.0) {
        atomicMax(out + r, min(max(ilogb(mx) + 1, -1000), 1000));
    }
}

// out[p] = E of the final column p of X: the maximum over the steps s >= sstart[p] (the
// column's first block) of xbexp[s][p] + colexp[p] - blkexp[s][p] (the block's exponent at
// the column's final scale); -1000 for a zero column.
__global__ void k_xexp_final(const int *__restrict__ xbexp, const int *__restrict__ blkexp, long ldb,
                             const int *__restrict__ colexp, const int *__restrict__ sstart, int nsteps, int m,
                             int *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) {
        return;
    }
    int e = -1000;
    const int ce = colexp[p];
    int s = sstart[p];
    // eight steps' loads in flight at a time
    for (; s + 8 <= nsteps; s += 8) {
        int eb[8], bb[8];
#pragma clang loop unroll(full)
        for (int u = 0; u < 8; ++u) {
            eb[u] = xbexp[(long)(s + u) * ldb + p];
            bb[u] = blkexp[(long)(s + u) * ldb + p];
        }
#pragma clang loop unroll(full)
        for (int u = 0; u < 8; ++u) {
            if (eb[u] > -1000) {
                e = max(e, eb[u] + ce - bb[u]);
            }
        }
    }
    for (; s < nsteps; ++s) {
        const int eb = xbexp[(long)s * ldb + p];
        if (eb > -1000) {
            e = max(e, eb + ce - blkexp[(long)s * ldb + p]);
        }
    }
    out[p] = min(max(e, -1000), 1000);
}

// Digits of the columns of X at their final scale, as k_slice_digits (cmode 0) over
// A(a, k) = X[k * ldx + a] for k < n, each entry first brought from the exponent of its block's
// step to the column's final one: 2^(colexp[a] - blkexp[rowstep[k] * ldb + a]) (the scale
// fix-up of the sweep, fused into this pass over X).
__global__ __launch_bounds__(256) void k_slice_xfix(const double *__restrict__ X, long ldx, int n, int M, int Kp,
                                                    const int *__restrict__ rowstep, const int *__restrict__ blkexp,
                                                    long ldb, const int *__restrict__ colexp,
                                                    const int *__restrict__ E, signed char *__restrict__ Q, long qs,
                                                    long kbp)
{
    const int a = blockIdx.x * 64 + threadIdx.x / 4;
    const int k0 = blockIdx.y * 16 + (threadIdx.x % 4) * 4;
    if (a >= M || k0 >= Kp) {
        return;
    }
    const int e = E[a] + 1;
    const int ce = colexp[a];
    int dg[SPLIT_DIGITS][4];
#pragma clang loop unroll(full)
    for (int q = 0; q < 4; ++q) {
        const int kk = k0 + q;
        double v = 0.0;
        if (kk < n) {
            v = ldexp(X[a + (long)kk * ldx], ce - blkexp[(long)rowstep[kk] * ldb + a]);
        }
        int d[SPLIT_DIGITS];
        split_digits(ldexp(v, -e), d);
#pragma clang loop unroll(full)
        for (int s = 0; s < SPLIT_DIGITS; ++s) {
            dg[s][q] = d[s];
        }
    }
    const long qo = (long)(k0 / 32) * kbp + (long)a * 32 + (k0 % 32);
#pragma clang loop unroll(full)
    for (int s = 0; s < SPLIT_DIGITS; ++s) {
        *reinterpret_cast<int *>(Q + qo + s * qs) =
            (dg[s][0] & 0xff) | ((dg[s][1] & 0xff) << 8) | ((dg[s][2] & 0xff) << 16) | ((dg[s][3] & 0xff) << 24);
    }
}

// ---------------------------------------------------------------------------------------
// k_gemm_i8: C[out(a) * ldc + b * cst] = sig[b] * C + 2^(EL[a] + ER[b] + 2) sum_t 255^-(t+2) acc_t(a, b)
// (mode 0: no sig * C term; sig == nullptr: sig = 1).  L digits (k-blocked, k-block pitch kbl,
// digit pitch lslice): k < K1 from L1, k in [K1, K1 + K2) from L2 at k - K1; R digits (pitches
// kbr, rslice) over [0, K1 + K2).  K1, K2 are multiples of 32; kmax (optional) bounds k + kofs
// per row a (rounded up to 32 per tile).
// Block tile BM x BN, waves WMW x WNW, each wave (BM/WMW) x (BN/WNW) in 32x32 MFMA blocks.
// K step 32.  A stage holds the step's digits of the tile, per digit BM (then BN) rows of
// 32 bytes; NSTAGE stages rotate, filled by direct global-to-LDS loads (16 bytes per lane,
// consecutive lanes to consecutive LDS addresses) two steps ahead of the MFMA work.  Rows
// beyond M (N) load row M-1 (N-1): they only feed outputs that are not written.
// Tiles: the launch is 1-D; workgroup w runs on XCD w % 8, so each XCD takes a contiguous
// range of tiles: a group of 8 column tiles over all row tiles (even work per XCD also for a
// triangular kmax; the XCD's concurrent tiles share L2 panels of L and R).
// MFMA operand bytes: lane l, byte q <-> (row l % 32, k = 16 (l / 32) + q) for A and B alike;
// result register v of lane l <-> (row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).
// ---------------------------------------------------------------------------------------
constexpr int QK = 32;
constexpr int NSTAGE = 3;
constexpr int SPLIT_KMAX = 26624; // 5 * 127^2 * K < 2^31

template <int BM, int BN, int WMW, int WNW>
__global__ __launch_bounds__(WMW *WNW * 64) void k_gemm_i8(
    int M, int N, int K1, int K2, const signed char *__restrict__ L1, const signed char *__restrict__ L2, long lslice,
    long kbl, const signed char *__restrict__ R, long rslice, long kbr, const int *__restrict__ EL,
    const int *__restrict__ ER, double *__restrict__ C, long ldc, int cst, const int *__restrict__ outcol, int mode,
    const double *__restrict__ sig, const int *__restrict__ kmax, int kofs)
{
    constexpr int NW = WMW * WNW;
    constexpr int WTM = BM / WMW, WTN = BN / WNW;
    constexpr int AM = WTM / 32, AN = WTN / 32;
    static_assert(AM * 32 == WTM && AN * 32 == WTN, "wave tile");
    constexpr int CA = SPLIT_DIGITS * BM * 2, CB = SPLIT_DIGITS * BN * 2; // 16-byte chunks per step
    constexpr int CH = CA + CB;
    static_assert(CA % 64 == 0 && CH % 64 == 0, "load instructions cover one operand each");
    constexpr int IPW = (CH + 64 * NW - 1) / (64 * NW); // load instructions per wave and step
    constexpr int STAGE = IPW * NW * 64 * 16;            // bytes per stage (chunks past CH: padding)
    constexpr int LV = SPLIT_LEVELS;
    __shared__ __attribute__((aligned(16))) signed char lds[NSTAGE * STAGE];

    int tm, tn;
    xcd_tile(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int a0 = tm * BM, b0 = tn * BN;

    const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
    const int wa = (wave / WNW) * WTM, wb = (wave % WNW) * WTN;
    int K = K1 + K2;
    if (kmax != nullptr) {
        int kk = 0;
        const int amax = min(a0 + BM, M);
        for (int q = a0; q < amax; ++q) {
            kk = max(kk, kmax[q]);
        }
        K = min(K, max(0, (kk - kofs + QK - 1) / QK * QK));
    }
    // per-lane source offsets of its chunks (fixed over the steps); chunks past CH reload
    // chunk CA (a valid R row) into the stage's padding.  LDS position 2 r + h' of a row pair
    // holds half h = h' ^ ((r >> 3) & 1) of row r, so the 16 lanes of a b128 fragment read
    // (rows r..r+15, one half) cover all 64 banks.
    long off[IPW];
#pragma clang loop unroll(full)
    for (int u = 0; u < IPW; ++u) {
        const int c = (u * NW + wave) * 64 + lane;
        const int cs = c < CH ? c : CA;
        const int cc = cs < CA ? cs : cs - CA;
        const int rows2 = cs < CA ? BM * 2 : BN * 2;
        const int sd = cc / rows2, rr = (cc % rows2) / 2;
        const int h = (cc % 2) ^ ((rr >> 3) & 1);
        if (cs < CA) {
            off[u] = sd * lslice + (long)min(a0 + rr, M - 1) * 32 + h * 16;
        }
        else {
            off[u] = sd * rslice + (long)min(b0 + rr, N - 1) * 32 + h * 16;
        }
    }
    auto issue = [&](int st) {
        const int k0 = st * QK;
        const signed char *lbase = k0 < K1 ? L1 + (long)(k0 / QK) * kbl : L2 + (long)((k0 - K1) / QK) * kbl;
        const signed char *rbase = R + (long)(k0 / QK) * kbr;
        signed char *stage = lds + (st % NSTAGE) * STAGE;
#pragma clang loop unroll(full)
        for (int u = 0; u < IPW; ++u) {
            const int cbase = (u * NW + wave) * 64; // wave-uniform; an instruction loads one operand
            const signed char *g = (cbase < CA ? lbase : rbase) + off[u];
            __builtin_amdgcn_global_load_lds((const void *)g,
                                             (__attribute__((address_space(3))) void *)(stage + cbase * 16), 16, 0, 0);
        }
    };

    i32x16 acc[LV][AM][AN];
#pragma clang loop unroll(full)
    for (int t = 0; t < LV; ++t) {
#pragma clang loop unroll(full)
        for (int i = 0; i < AM; ++i) {
#pragma clang loop unroll(full)
            for (int j = 0; j < AN; ++j) {
#pragma clang loop unroll(full)
                for (int v = 0; v < 16; ++v) {
                    acc[t][i][j][v] = 0;
                }
            }
        }
    }
    const int nst = K / QK;
    if (nst > 0) {
        issue(0);
    }
    if (nst > 1) {
        issue(1);
    }
    // fragment rows start at multiples of 32, so a row's swizzle bit is that of fr
    const int fr = lane % 32, fsw = (((lane / 32) ^ ((fr >> 3) & 1))) * 16;
    for (int st = 0; st < nst; ++st) {
        if (st + 1 < nst) {
            wait_vm<IPW>();
        }
        else {
            wait_vm<0>();
        }
        // every wave's loads of this stage have landed (its own vmcnt wait) and every wave is
        // done with the stage refilled next; no memory fence, so the loads two steps ahead
        // stay in flight
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        if (st + 2 < nst) {
            issue(st + 2);
        }
        const signed char *sa = lds + (st % NSTAGE) * STAGE;
        const signed char *sb = sa + CA * 16;
        i32x4 bf[SPLIT_DIGITS][AN];
#pragma clang loop unroll(full)
        for (int sd = 0; sd < SPLIT_DIGITS; ++sd) {
#pragma clang loop unroll(full)
            for (int j = 0; j < AN; ++j) {
                bf[sd][j] = *reinterpret_cast<const i32x4 *>(sb + sd * (BN * 32) + (wb + 32 * j + fr) * 32 + fsw);
            }
        }
#pragma clang loop unroll(full)
        for (int sd = 0; sd < SPLIT_DIGITS; ++sd) {
            i32x4 af[AM];
#pragma clang loop unroll(full)
            for (int i = 0; i < AM; ++i) {
                af[i] = *reinterpret_cast<const i32x4 *>(sa + sd * (BM * 32) + (wa + 32 * i + fr) * 32 + fsw);
            }
#pragma clang loop unroll(full)
            for (int r = 0; r + sd < LV && r < SPLIT_DIGITS; ++r) {
#pragma clang loop unroll(full)
                for (int i = 0; i < AM; ++i) {
#pragma clang loop unroll(full)
                    for (int j = 0; j < AN; ++j) {
                        acc[sd + r][i][j] =
                            __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[r][j], acc[sd + r][i][j], 0, 0, 0);
                    }
                }
            }
        }
    }
    // epilogue, per group of 8 output rows (of a 32-row block i): the level sums (their
    // accumulators end there), the row indices, then all of the group's C reads before its
    // writes.  Loads are unconditional at clamped addresses, so that loads of one kind are in
    // flight together (a wait per output would serialise their round trips).
    int ceb[AN];
    double csg[AN];
#pragma clang loop unroll(full)
    for (int j = 0; j < AN; ++j) {
        ceb[j] = ER[min(b0 + wb + 32 * j + fr, N - 1)];
        csg[j] = 1.0;
    }
    if (mode == 1 && sig != nullptr) {
#pragma clang loop unroll(full)
        for (int j = 0; j < AN; ++j) {
            csg[j] = sig[min(b0 + wb + 32 * j + fr, N - 1)];
        }
    }
    constexpr int EG = 8;
#pragma clang loop unroll(full)
    for (int i = 0; i < AM; ++i) {
#pragma clang loop unroll(full)
        for (int h = 0; h < 16 / EG; ++h) {
            double sm[EG][AN];
#pragma clang loop unroll(full)
            for (int u = 0; u < EG; ++u) {
                const int v = EG * h + u;
#pragma clang loop unroll(full)
                for (int j = 0; j < AN; ++j) {
                    double sum = (double)acc[LV - 1][i][j][v];
#pragma clang loop unroll(full)
                    for (int t = LV - 2; t >= 0; --t) {
                        sum = fma(sum, 1.0 / SPLIT_BASE, (double)acc[t][i][j][v]);
                    }
                    sm[u][j] = sum * (1.0 / (SPLIT_BASE * SPLIT_BASE));
                }
            }
            int rrow[EG], rea[EG];
#pragma clang loop unroll(full)
            for (int u = 0; u < EG; ++u) {
                const int v = EG * h + u;
                rrow[u] = min(a0 + wa + 32 * i + 8 * (v / 4) + 4 * (lane / 32) + (v % 4), M - 1);
                rea[u] = EL[rrow[u]];
            }
            if (outcol != nullptr) {
#pragma clang loop unroll(full)
                for (int u = 0; u < EG; ++u) {
                    rrow[u] = outcol[rrow[u]];
                }
            }
            double cv[EG][AN];
#pragma clang loop unroll(full)
            for (int u = 0; u < EG; ++u) {
#pragma clang loop unroll(full)
                for (int j = 0; j < AN; ++j) {
                    cv[u][j] = 0.0;
                }
            }
            if (mode == 1) {
#pragma clang loop unroll(full)
                for (int u = 0; u < EG; ++u) {
#pragma clang loop unroll(full)
                    for (int j = 0; j < AN; ++j) {
                        const int bcl = min(b0 + wb + 32 * j + fr, N - 1);
                        cv[u][j] = C[(long)rrow[u] * ldc + (long)bcl * cst];
                    }
                }
            }
#pragma clang loop unroll(full)
            for (int u = 0; u < EG; ++u) {
                const int v = EG * h + u;
                const int a = a0 + wa + 32 * i + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
#pragma clang loop unroll(full)
                for (int j = 0; j < AN; ++j) {
                    const int b = b0 + wb + 32 * j + fr;
                    if (a < M && b < N) {
                        double d = ldexp(sm[u][j], rea[u] + ceb[j] + 2);
                        if (mode == 1) {
                            d = fma(csg[j], cv[u][j], d);
                        }
                        C[(long)rrow[u] * ldc + (long)b * cst] = d;
                    }
                }
            }
        }
    }
}

#define XTG_I8_INST(BM, BN, WMW, WNW)                                                                               \
    template __global__ void k_gemm_i8<BM, BN, WMW, WNW>(int, int, int, int, const signed char *,                   \
                                                         const signed char *, long, long, const signed char *, long, \
                                                         long, const int *, const int *, double *, long, int,        \
                                                         const int *, int, const double *, const int *, int);
XTG_I8_INST(128, 128, 2, 4)
XTG_I8_INST(64, 128, 1, 4)
#undef XTG_I8_INST

// One launch; the tile is 64 x 128 when M is small or 128 x 128 tiles would not fill the GPU.
static int launch_i8(hipStream_t st, int M, int N, int K1, int K2, const signed char *L1, const signed char *L2,
                     long lslice, long kbl, const signed char *R, long rslice, long kbr, const int *EL, const int *ER,
                     double *C, long ldc, int cst, const int *outcol, int mode, const double *sig, const int *kmax,
                     int kofs)
{
    // 64-row tiles when 128-row tiles would leave CUs idle (few rows, or few columns)
    const long t128 = (long)((N + 127) / 128) * ((M + 127) / 128);
    if (M <= 64 || t128 < 256) {
        const dim3 grid(((N + 127) / 128) * ((M + 63) / 64));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_i8<64, 128, 1, 4>), grid, dim3(256), 0, st, M, N, K1, K2, L1, L2,
                           lslice, kbl, R, rslice, kbr, EL, ER, C, ldc, cst, outcol, mode, sig, kmax, kofs);
    }
    else {
        const dim3 grid(((N + 127) / 128) * ((M + 127) / 128));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gemm_i8<128, 128, 2, 4>), grid, dim3(512), 0, st, M, N, K1, K2, L1, L2,
                           lslice, kbl, R, rslice, kbr, EL, ER, C, ldc, cst, outcol, mode, sig, kmax, kofs);
    }
    return (int)hipGetLastError();
}

// Launches C = sig C + L R^T on the digits; returns 0 or a HIP error code.  A product longer
// than SPLIT_KMAX (single segment only) runs in K chunks accumulated into C.
int launch_gemm_i8(hipStream_t st, int M, int N, int K1, int K2, const signed char *L1, const signed char *L2,
                   long lslice, long kbl, const signed char *R, long rslice, long kbr, const int *EL, const int *ER,
                   double *C, long ldc, int cst, const int *outcol, int mode, const double *sig, const int *kmax)
{
    if (M <= 0 || N <= 0) {
        return 0;
    }
    if (K1 % QK != 0 || K2 % QK != 0 || K2 > SPLIT_KMAX) {
        return (int)hipErrorInvalidValue;
    }
    if (K1 + K2 > SPLIT_KMAX) {
        if (K2 != 0) {
            return (int)hipErrorInvalidValue;
        }
        for (int k0 = 0; k0 < K1; k0 += SPLIT_KMAX) {
            const int kc = std::min(SPLIT_KMAX, K1 - k0);
            const long lo = (long)(k0 / QK) * kbl, ro = (long)(k0 / QK) * kbr;
            const int rc = launch_i8(st, M, N, kc, 0, L1 + lo, L1 + lo, lslice, kbl, R + ro, rslice, kbr, EL, ER, C,
                                     ldc, cst, outcol, k0 == 0 ? mode : 1, k0 == 0 ? sig : nullptr, kmax, k0);
            if (rc != 0) {
                return rc;
            }
        }
        return 0;
    }
    return launch_i8(st, M, N, K1, K2, L1, L2, lslice, kbl, R, rslice, kbr, EL, ER, C, ldc, cst, outcol, mode, sig, kmax,
                     0);
}

} // namespace xtg
