/*
 * libxtgevc3 -- C ABI of the MI355X (gfx950) xTGEVC3 eigenvector path.
 *
 * Every entry point takes plain pointers and sizes.  Matrices are column-major.
 * The host-pointer entry points are drop-ins for the reference's interface
 * (vlovero/xTGEVC3 include/tgevc3.h:10-13) with C linkage; the C++-linkage symbols
 * dtgevc3/stgevc3/ztgevc3/ctgevc3 with the reference's exact declarations are exported by
 * the same library (see include/tgevc3.h).
 *
 * Output vectors are normalised as LAPACK xTGEVC documents: each vector is scaled by a
 * positive factor so that its largest component has |Re| + |Im| = 1 (for a complex pair of
 * the real interface, over both columns).  The reference leaves vectors unnormalised; they
 * differ from its output by that positive factor only.
 *
 * info: 0, -2 (bad howmny), -16 (mm < m; m already written), -19 (lwork too small) with
 * the reference's meaning (src/dtgevc3.cpp:1179-1242, src/ztgevc3.cpp:462-516);
 * XTGEVC3_INFO_DEVICE_ERROR when the GPU path fails (message: xtgevc3_last_error()).
 */
#ifndef XTGEVC3_H
#define XTGEVC3_H

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> xtgevc3_zcomplex;
typedef std::complex<float> xtgevc3_ccomplex;
extern "C" {
#else
#include <complex.h>
typedef double _Complex xtgevc3_zcomplex;
typedef float _Complex xtgevc3_ccomplex;
#endif

#define XTGEVC3_ABI_VERSION 1
#define XTGEVC3_INFO_DEVICE_ERROR (-1000)

/* precision selector of the *_ex entry points */
#define XTGEVC3_PREC_FP64 0      /* all arithmetic in FP64 (complex128) */
#define XTGEVC3_PREC_MIXED_F32 1 /* updates/back-transform on FP32 MFMA, FP64 solves */
#define XTGEVC3_PREC_SPLIT_I8 3  /* updates/back-transform on exact int8 MFMA digit products */
/* (any other value, 2 included, is rejected with XTGEVC3_INFO_UNSUPPORTED) */

int xtgevc3_abi_version(void);
const char *xtgevc3_last_error(void);

/* Host-pointer drop-ins.  Replace dtgevc3/stgevc3/ztgevc3/ctgevc3 of the reference
 * (include/tgevc3.h:10-13; src/dtgevc3.cpp:1150, src/stgevc3.cpp:1122,
 * src/ztgevc3.cpp:435, src/ctgevc3.cpp:435).  Same arguments, same info/m/lwork rules. */
void xtgevc3_dtgevc3(char side, char howmny, const int *select, int n, const double *S, int lds, const double *P,
                     int ldp, const double *alphar, const double *alphai, const double *beta, double *VL, int ldvl,
                     double *VR, int ldvr, int mm, int *m, double *work, int lwork, int *info);
void xtgevc3_stgevc3(char side, char howmny, const int *select, int n, const float *S, int lds, const float *P,
                     int ldp, const float *alphar, const float *alphai, const float *beta, float *VL, int ldvl,
                     float *VR, int ldvr, int mm, int *m, float *work, int lwork, int *info);
void xtgevc3_ztgevc3(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                     const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha, const xtgevc3_zcomplex *beta,
                     xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_zcomplex *work, int lwork, int *info);
void xtgevc3_ctgevc3(char side, char howmny, const int *select, int n, const xtgevc3_ccomplex *S, int lds,
                     const xtgevc3_ccomplex *P, int ldp, const xtgevc3_ccomplex *alpha, const xtgevc3_ccomplex *beta,
                     xtgevc3_ccomplex *VL, int ldvl, xtgevc3_ccomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_ccomplex *work, int lwork, int *info);

/* Device-pointer variants (no counterpart in the reference, which is host-only; they are
 * the same call as the drop-ins above without the staging copies).  S, P, VL, VR are device
 * arrays; select, alphar/alphai/beta (alpha/beta) are host arrays; stream is a hipStream_t
 * (NULL: the default stream).  work/lwork are not taken.  Returns 0, or
 * XTGEVC3_INFO_DEVICE_ERROR on a HIP failure. */
int xtgevc3_dtgevc3_device(char side, char howmny, const int *select, int n, const double *S, int lds,
                           const double *P, int ldp, const double *alphar, const double *alphai, const double *beta,
                           double *VL, int ldvl, double *VR, int ldvr, int mm, int *m, int *info, void *stream);
/* dtgevc3 on device with a precision selector (no counterpart in the reference: the mixed
 * form of the update and back-transform products).  XTGEVC3_PREC_FP64 is
 * xtgevc3_dtgevc3_device; XTGEVC3_PREC_MIXED_F32 runs the off-diagonal updates and the
 * back-transform on FP32 MFMA products accumulated in FP32 and added to the FP64 running
 * right-hand sides / output; XTGEVC3_PREC_SPLIT_I8 runs them on int8 MFMA products of
 * five-digit splits of the operands (7-bit digits, exponent per row of the pencil / of Z and
 * per column of the operands), accumulated exactly in int32 and combined in FP64 (error about
 * 2^-34 of the row-maximum times column-maximum products).  The block solves, scale factors
 * and normalisation stay FP64 in both.  Other selectors return XTGEVC3_INFO_UNSUPPORTED. */
int xtgevc3_dtgevc3_device_ex(char side, char howmny, const int *select, int n, const double *S, int lds,
                              const double *P, int ldp, const double *alphar, const double *alphai,
                              const double *beta, double *VL, int ldvl, double *VR, int ldvr, int mm, int *m,
                              int *info, void *stream, int precision);
int xtgevc3_ztgevc3_device(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                           const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha,
                           const xtgevc3_zcomplex *beta, xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR,
                           int ldvr, int mm, int *m, int *info, void *stream);

/* ztgevc3 on device with a precision selector (no counterpart in the reference):
 * XTGEVC3_PREC_FP64 is xtgevc3_ztgevc3_device; XTGEVC3_PREC_SPLIT_I8 runs the far updates and
 * the back-transform as two real int8 digit products each (Re = [Lr, -Li].[Rr, Ri],
 * Im = [Li, Lr].[Rr, Ri] over (re, im)-interleaved K), as in xtgevc3_dtgevc3_device_ex.
 * Other selectors return XTGEVC3_INFO_UNSUPPORTED. */
int xtgevc3_ztgevc3_device_ex(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                              const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha,
                              const xtgevc3_zcomplex *beta, xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR,
                              int ldvr, int mm, int *m, int *info, void *stream, int precision);

/* Backward (adjoint) of the selected right eigenvectors without back-transform (no
 * counterpart in the reference): for the normalised output V (n x m) of
 * xtgevc3_dtgevc3_device(side='R', howmny = select ? 'S' : 'A') and nbatch cotangent batches
 * G (column b*m + c = dL/dV[:, c] of batch b), writes dL/dS and dL/dP of every batch
 * (n x n column-major matrices at dS + b*n*ldds, dP + b*n*lddp).  Only the entries the forward
 * reads carry a gradient: dL/dS on the upper Hessenberg band (row i <= column j + 1), dL/dP on
 * the upper triangle (i <= j); every entry below is written as zero.  alpha_k = S_kk and
 * beta_k = P_kk are differentiated as functions of the diagonals.  Supports real, finite
 * selected eigenvalues only: info = XTGEVC3_INFO_UNSUPPORTED otherwise.  Device arrays,
 * host eigenvalues/select.  info > 0: that many (batch, column) adjoint vectors exceed the
 * double range (ill-conditioned eigenvectors: the exact gradient overflows); the entries of
 * dS / dP that depend on them are NaN and every other entry is finite and exact.  The call
 * synchronises `stream` at its end to read that count. */
#define XTGEVC3_INFO_UNSUPPORTED (-1001)
int xtgevc3_dtgevc3_backward_device(int n, const double *S, int lds, const double *P, int ldp, const double *alphar,
                                    const double *alphai, const double *beta, const int *select, const double *V,
                                    int ldv, const double *G, int ldg, int nbatch, double *dS, int ldds, double *dP,
                                    int lddp, int *info, void *stream);

/* xtgevc3_dtgevc3_backward_device with a precision argument (no counterpart in the reference):
 * XTGEVC3_PREC_FP64 as above; XTGEVC3_PREC_SPLIT_I8 / XTGEVC3_PREC_MIXED_F32 run the far
 * updates of the adjoint block sweep as in the corresponding forms of
 * xtgevc3_dtgevc3_device_ex (the diagonal-block solves, near updates, and the outer products
 * that assemble dS and dP stay FP64).  Other values: XTGEVC3_INFO_UNSUPPORTED, nothing written. */
int xtgevc3_dtgevc3_backward_device_ex(int n, const double *S, int lds, const double *P, int ldp,
                                       const double *alphar, const double *alphai, const double *beta,
                                       const int *select, const double *V, int ldv, const double *G, int ldg,
                                       int nbatch, double *dS, int ldds, double *dP, int lddp, int *info, void *stream,
                                       int precision);

/* Allocation-free variant of xtgevc3_dtgevc3_device (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * xtgevc3_dtgevc3_device (FP64), for n <= XTGEVC3_NOALLOC_MAX_N; larger n returns
 * XTGEVC3_INFO_UNSUPPORTED with a message and touches nothing.  The call makes no hipMalloc /
 * hipFree and creates no stream or event: its workspace is one static device array per GPU of
 * at most XTGEVC3_NOALLOC_SCRATCH_BYTES (held until the process ends) and every launch and copy
 * runs on `stream`.  All calls on one GPU share that array: calls that overlap in time on one
 * GPU (other streams, other host threads without ordering) race.  Like the other device
 * entries it synchronises `stream` once per call to read the pencil's column norms. */
#define XTGEVC3_NOALLOC_MAX_N 512
#define XTGEVC3_NOALLOC_SCRATCH_BYTES (16u << 20)
int xtgevc3_dtgevc3_device_noalloc(char side, char howmny, const int *select, int n, const double *S, int lds,
                                   const double *P, int ldp, const double *alphar, const double *alphai,
                                   const double *beta, double *VL, int ldvl, double *VR, int ldvr, int mm, int *m,
                                   int *info, void *stream);

/* Backward for either side and through the back-transform (no counterpart in the reference: the
 * adjoint of the left sweep, src/dtgevc3.cpp:1476-1663, and of the back-transform VR := Z X,
 * VL := Q Y, :1443-1454 and :1635-1657).  side 'R': V holds normalised right eigenvectors of
 * the selected real eigenvalues; side 'L': normalised left eigenvectors (y^H (beta S - alpha P) =
 * 0).  Zb NULL: V as xtgevc3_dtgevc3_device(side, howmny = select ? 'S' : 'A') returns it; Zb given:
 * V = the back-transformed vectors of those eigenvalues (Zb = Z for side R, Q for side L; the
 * columns howmny = 'B' returns for them), and dZ (optional, n x n per batch at dZ + b*n*lddz)
 * receives dL/dZb.  G, nbatch, dS, dP, info as in xtgevc3_dtgevc3_backward_device (FP64).
 * Returns 0, -1 for a bad side, or XTGEVC3_INFO_DEVICE_ERROR. */
int xtgevc3_dtgevc3_backward_full_device(char side, int n, const double *S, int lds, const double *P, int ldp,
                                         const double *alphar, const double *alphai, const double *beta,
                                         const int *select, const double *Zb, int ldz, const double *V, int ldv,
                                         const double *G, int ldg, int nbatch, double *dS, int ldds, double *dP,
                                         int lddp, double *dZ, int lddz, int *info, void *stream);

/* Eigenvalue condition numbers (no counterpart in the reference; the semantics of LAPACK
 * dtgsna with JOB = 'E'): for the eigenvalues chosen by howmny ('A' all, 'S' by select, with
 * dtgevc3's selection rule, complex pairs from alphai) and their left and right eigenvectors
 * VL, VR (n x m device arrays, the columns dtgevc3 returns for side = 'B', in the same basis as
 * the device arrays S and P), writes s[c] = |(u^H S v, u^H P v)| / (|u| |v|) for every column
 * c (both columns of a pair get the pair's value) to the device array s (m values).  m must be
 * the column count of the selection.  Returns 0, or a negative code with
 * xtgevc3_last_error() set: -1 bad howmny, -3 bad sizes, XTGEVC3_INFO_DEVICE_ERROR. */
int xtgevc3_dtgsna_e_device(char howmny, const int *select, int n, const double *S, int lds, const double *P,
                            int ldp, const double *alphai, const double *VL, int ldvl, const double *VR, int ldvr,
                            int m, double *s, int *info, void *stream);

/* The complex128 counterpart (LAPACK ztgsna, JOB = 'E'): one value per selected eigenvalue
 * column, s[c] = |(u^H S v, u^H P v)| / (|u| |v|), from the columns of
 * xtgevc3_ztgevc3_device(side = 'B', ...). */
int xtgevc3_ztgsna_e_device(char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                            const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *VL, int ldvl,
                            const xtgevc3_zcomplex *VR, int ldvr, int m, double *s, int *info, void *stream);

/* Live kernel timing (no counterpart in the reference).  When enabled, every launch of the
 * hot kernels is bracketed by HIP events on its stream; xtgevc3_profile_read synchronises,
 * returns per-kind totals (milliseconds, launches, algorithmic flops and bytes) and resets.
 * Kinds: */
#define XTGEVC3_PROF_SOLVE 0             /* per-column diagonal-block solves */
#define XTGEVC3_PROF_UPDATE 1            /* off-diagonal RHS update GEMM, 64x64 tiles (complex path: all) */
#define XTGEVC3_PROF_BACK 2              /* back-transform GEMM */
#define XTGEVC3_PROF_OTHER 3             /* copies, scale fix-up, normalisation */
#define XTGEVC3_PROF_UPDATE_128x128 4    /* off-diagonal RHS update GEMM, 128x128 tiles */
#define XTGEVC3_PROF_UPDATE_128x64 5     /* off-diagonal RHS update GEMM, 128x64 tiles */
#define XTGEVC3_PROF_UPDATE_64x128 6     /* off-diagonal RHS update GEMM, 64x128 tiles */
#define XTGEVC3_PROF_UPDATE_SPLIT 7      /* far RHS update (whole group), split-int8 digit products */
#define XTGEVC3_PROF_BACK_SPLIT 8        /* back-transform, split-int8 digit products */
#define XTGEVC3_PROF_SPLIT_DIGITS 9      /* digit splitting of operands (split-int8 form) */
#define XTGEVC3_PROF_KINDS 10
void xtgevc3_profile_enable(int on);
int xtgevc3_profile_read(double *ms, long *launches, double *flops, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
