// Shared definitions of the gfx950 xTGEVC3 path: column descriptors, launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace xtg {

// One eigenvector column of the working problem (acoeff*S_w - bcoeff*P_w) x = 0, solved
// bottom-up on the upper quasi-triangular working pencil.  Right eigenvectors use the
// pencil itself; left eigenvectors and the adjoint solve use the reversed transpose
// J S^T J (J = row/column reversal), so one solver serves every side.
//
// A complex-conjugate pair occupies two lanes: one holds the real part of the complex
// vector, the other the imaginary part; each lane sees its partner through a lane shuffle.
struct ColDesc {
    double a;    // acoeff (real, > 0 unless the eigenvalue is infinite)
    double bR;   // Re(bcoeff) of the working shift
    double bI;   // Im(bcoeff) of the working shift (0 for real eigenvalues)
    double pad;
    int d;       // first row of the column's own diagonal block (init rows); >= n: none
    int lo;      // rows < lo lie outside the column's system and stay zero
    int partner; // packed index of the partner lane column (== own index when real)
    int flags;   // bit0: imaginary-part lane, bit1: member of a pair, bits2-3: init mode
};

enum : int {
    INIT_NONE = 0,     // all rows come from the right-hand side held in X
    INIT_UNIT = 1,     // x_d = 1 (real eigenvalue, or complex interface)
    INIT_PAIR_ROW1 = 2, // x_{d..d+1} = null vector of the first row of the 2x2 block
    INIT_PAIR_ROW2 = 3, // x_{d..d+1} = null vector of the second row of the 2x2 block
};

// One eigenvalue of the real interface in the solve kernel: a real eigenvalue (one working
// column) or a complex-conjugate pair (two adjacent columns: real part, imaginary part).
struct UnitDesc {
    double a, bR, bI, pad; // shift coefficients (bI = 0 for a real eigenvalue)
    int col;               // first working column
    int w;                 // 1 or 2 columns
    int d;                 // first row of the unit's own diagonal block (>= n: none)
    int lo;                // rows < lo lie outside the unit's system and stay zero
    int init;              // INIT_* mode
    int idx;               // unit index (the solve's tile-ordered copy: -1 for an empty slot)
    int pad2, pad3;
};

// The row blocks of one update group of a sweep, top-down: block j covers local rows
// [off[j], off[j+1]) of the group (off[nblk] = wr, unused entries = wr) and is solved in step
// step[j]; step[0] is the group's last step.  At most GROUP_MAX_BLOCKS blocks per group.
constexpr int GROUP_MAX_BLOCKS = 8;
struct GroupInfo {
    int off[GROUP_MAX_BLOCKS + 1];
    int step[GROUP_MAX_BLOCKS];
    int wr;
    int last; // the launch is the group's last step
};

// Operand images of one step of the real sweep, packed once per sweep by k_pack_steps so that
// the solve kernel reads them as contiguous 16-byte runs (the pencil is column-major with a
// column pitch of many KB: a block's columns fall into few L2 channels, and every workgroup
// of a launch reads the same block):
//   block image at boff (PK_BLK {S,P} pairs): the diagonal block [i0, i0+nb)^2 in the solve
//   kernel's LDS layout, then its subdiagonal;
//   panel image at poff: [S | P][g0 + 16 t + fr, i0 + k] of the near update (the group's rows
//   above the block), row tiles t < nrt of 16 rows, in the lane order of the 16x16x4 FP64 MFMA
//   A operand: lane fr + 16 fk of k-step kk holds column k = 4 kk' + fk of part kk < kq ? S : P
//   (kk' = kk mod kq, kq = ks / 4, ks = nb rounded up to a multiple of 4); k-steps 2p and
//   2p + 1 of a lane are one {x, y} pair at pair index (t kq + p) 64 + lane.
struct StepPack {
    int i0, nb, g0, nrt;
    long boff, poff;
};
constexpr int PK_BLK = 2720; // {S,P} pairs of a block image (2640 staged + 65 subdiagonal, padded)
constexpr int PK_KSMAX = 68; // largest ks (blocks of at most 65 rows)

// Exponent bookkeeping of the mixed (FP32-product) form, written by the real solve kernel
// when on != 0: wgrp[col] bounds the group's W (2^wgrp; relative to the column's exponent
// until the group's last step, absolute after it), xrel[col] bounds the solution entries
// relative to the column's exponent.
struct MixedExp {
    int *wgrp;
    int *xrel;
    int on;
};

// Same layout for the complex interface: the working vector entries are complex.
struct ZColDesc {
    double aR, aI; // acoeff (complex for the complex interface)
    double bR, bI; // bcoeff
    int d, lo, flags, pad;
};

constexpr int WAVE = 64;

// Tile of workgroup l in a 1-D GEMM launch over tm_n x tn_n tiles.  Workgroup l runs on
// XCD l % 8, so each XCD takes a contiguous range of positions p; positions run over groups
// of 8 column tiles, within a group the row tiles from the last down (largest per-row K bound
// first), the group's columns fastest.  An XCD so works through one group of column tiles over
// all row tiles (even work for a triangular K bound) and its concurrent tiles share L2 panels.
__device__ __forceinline__ void xcd_tile(int l, int tm_n, int tn_n, int &tm, int &tn)
{
    constexpr int NX = 8;
    const int nblk = tm_n * tn_n;
    const int x = l % NX, slot = l / NX;
    const int q = nblk / NX, r = nblk % NX;
    const int p = x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
    const int grp = p / (8 * tm_n), gw = min(8, tn_n - grp * 8), rem = p - grp * 8 * tm_n;
    tm = tm_n - 1 - rem / gw;
    tn = grp * 8 + rem % gw;
}

// s_waitcnt with vmcnt = N (0..63), other counters at their maximum (gfx9 encoding)
template <int N> __device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Split-int8 form: digits per value and digit-product levels kept (i + j < SPLIT_LEVELS).
constexpr int SPLIT_DIGITS = 5;
constexpr int SPLIT_LEVELS = 5;

// Balanced base-255 digits of u (|u| <= 1/2): x = 255 u_s, d_s = x rounded to nearest with
// ties toward zero, u_{s+1} = x - d_s.  |x| <= 127.5, so d_s lies in [-127, 127] (the int8
// range) and |u_{s+1}| <= 1/2; u = sum_s d_s 255^-(s+1) up to 255^-D / 2 and the FP64
// rounding of the products 255 u_s (relative 2^-53 each).
constexpr double SPLIT_BASE = 255.0;
__device__ __forceinline__ void split_digits(double u, int d[SPLIT_DIGITS])
{
#pragma clang loop unroll(full)
    for (int s = 0; s < SPLIT_DIGITS; ++s) {
        const double x = u * SPLIT_BASE;
        const double r = x >= 0.0 ? ceil(x - 0.5) : floor(x + 0.5);
        d[s] = (int)fmin(fmax(r, -127.0), 127.0);
        u = x - r;
    }
}

// Split-int8 bookkeeping of the real solve kernel (bexp == nullptr: off): bexp and xbexp
// (offset to the step, row pitch ldx like blkexp) receive E of each column's block operand
// and of its block of X (2^E > max |.|, -1000 for a zero block); the group's last step
// writes gexp[p] = E of the column's whole group operand at its final scale.
struct SplitOut {
    int *bexp;
    int *gexp;
    int *xbexp;
};
// tile configurations of the FP64 GEMM (a x b)
enum : int { GEMM_128x128 = 0, GEMM_64x128 = 1, GEMM_128x64 = 2, GEMM_64x64 = 3, GEMM_32x64 = 4 };

} // namespace xtg

#define XTG_CHECK(x)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            xtg::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                           ":" + std::to_string(__LINE__));                                        \
            return -1000;                                                                          \
        }                                                                                          \
    } while (0)

namespace xtg {
// XTGEVC3_DEBUG_SYNC=1 in the environment: synchronise and check after every launch.
bool debug_sync();
} // namespace xtg

#define XTG_LAUNCH_CHECK(name)                                                                     \
    do {                                                                                           \
        hipError_t le_ = hipGetLastError();                                                        \
        if (le_ == hipSuccess && xtg::debug_sync()) {                                              \
            le_ = hipDeviceSynchronize();                                                          \
        }                                                                                          \
        if (le_ != hipSuccess) {                                                                   \
            xtg::set_error(std::string("launch of ") + (name) + ": " + hipGetErrorString(le_));    \
            return -1000;                                                                          \
        }                                                                                          \
    } while (0)
