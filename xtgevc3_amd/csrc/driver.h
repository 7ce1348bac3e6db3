// Internal interfaces between the C-ABI layer and the per-type drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <complex>
#include <functional>
#include <string>
#include <vector>

namespace xtg {

// One selected eigenvalue "unit": a real eigenvalue (w = 1) or a complex-conjugate pair
// (w = 2, columns k and k+1), with its first output column.
struct Unit {
    int k;
    int w;
    int outc;
};

void set_error(const std::string &s);
const char *last_error();

int reference_lwork(int n, int b);
template <class T>
bool check_args(char howmny, int n, int num_sel, int mm, int *m, T *work, int lwork, int *info, bool have_work);

int real_units(char howmny, const int *select, int n, const double *ai, std::vector<Unit> &units);
int complex_units(char howmny, const int *select, int n, std::vector<Unit> &units);

// Z for the back-transform (howmny = 'B') delivered beside the sweep instead of in VL / VR:
// zr / zl are device matrices (leading dimension n, the path's element type) that the caller
// fills on another stream.  ready(left) is called before that side's back-transform is queued;
// it returns an event recorded after the side's matrix is written, or null (error message set)
// if that failed.  FP64 and FP32-product forms only (the split form reads Z from the start of
// the sweep).
// Optionally (FP64 form, chunks > 1) the back-transform runs in up to `chunks` pieces of
// consecutive output columns; after each piece's vectors are final on st, done(left, c0, c1, st)
// is called (output columns [c0, c1)) so that the caller can queue their copy-out behind an
// event of its own.  A nonzero return aborts the call with that code.  out32[side] (right,
// left; optional): before done(), the piece's vectors are also rounded to FP32 into this
// device matrix (same columns, leading dimension n, 1 or 2 floats per element).
struct ZFeed {
    const void *zr, *zl;
    std::function<hipEvent_t(bool left)> ready;
    int chunks;
    std::function<int(bool left, int c0, int c1, hipStream_t st)> done;
    void *out32[2];
};

// prec (XTGEVC3_PREC_*): FP64 throughout; FP32-product or split-int8 far updates and
// back-transform (mixed forms).  safemin: smallest normal number of the interface precision
// (FLT_MIN for the FP32 / complex64 interfaces), the floor of the global scales and shift
// normalisers and the threshold below which |beta| makes an eigenvalue infinite.
constexpr double kDblMin = 2.2250738585072014e-308;
constexpr double kFltMin = 1.1754943508222875e-38;
int real_device(char side, char howmny, const int *select, int n, const double *S, long lds, const double *P,
                long ldp, const double *ar, const double *ai, const double *be, double *VL, long ldvl, double *VR,
                long ldvr, hipStream_t st, int prec = 0, const ZFeed *zf = nullptr, double safemin = kDblMin);

int complex_device(char side, char howmny, const int *select, int n, const std::complex<double> *S, long lds,
                   const std::complex<double> *P, long ldp, const std::complex<double> *al,
                   const std::complex<double> *be, std::complex<double> *VL, long ldvl, std::complex<double> *VR,
                   long ldvr, hipStream_t st, int prec = 0, const ZFeed *zf = nullptr, double safemin = kDblMin);

// allocation-free FP64 variant of real_device (static per-device workspace, n <= XTGEVC3_NOALLOC_MAX_N)
int real_device_noalloc(char side, char howmny, const int *select, int n, const double *S, long lds,
                        const double *P, long ldp, const double *ar, const double *ai, const double *be, double *VL,
                        long ldvl, double *VR, long ldvr, hipStream_t st);

int real_backward_device(int n, const double *S, long lds, const double *P, long ldp, const double *ar,
                         const double *ai, const double *be, const int *select, const double *V, long ldv,
                         const double *G, long ldg, int nbatch, double *dS, long ldds, double *dP, long lddp, int *info,
                         hipStream_t st, int prec = 0);

// backward for either side, optionally through the back-transform (Zb, and dL/dZb into dZ)
int real_backward_full_device(char side, int n, const double *S, long lds, const double *P, long ldp,
                              const double *ar, const double *ai, const double *be, const int *select,
                              const double *Zb, long ldz, const double *V, long ldv, const double *G, long ldg,
                              int nbatch, double *dS, long ldds, double *dP, long lddp, double *dZ, long lddz,
                              int *info, hipStream_t st);

int real_tgsna_device(char howmny, const int *select, int n, const double *S, long lds, const double *P, long ldp,
                      const double *ai, const double *VL, long ldvl, const double *VR, long ldvr, int m, double *s,
                      hipStream_t st);

int complex_tgsna_device(char howmny, const int *select, int n, const std::complex<double> *S, long lds,
                         const std::complex<double> *P, long ldp, const std::complex<double> *VL, long ldvl,
                         const std::complex<double> *VR, long ldvr, int m, double *s, hipStream_t st);

// live timing of kernel launches (profile.hip); begin returns -1 when disabled
int prof_begin(hipStream_t st);
void prof_end(int idx, hipStream_t st, int kind, double flops, double bytes);

} // namespace xtg
