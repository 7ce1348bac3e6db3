// TEST INFRASTRUCTURE ONLY.  C-linkage forwarding shims onto the reference's own C++ entry
// points (include/tgevc3.h:10-13), so that Python can load the reference library built
// from the sources under /root/reference by oracle/Makefile.  Contains no reference code.
#include <complex>
#include "tgevc3.h"

extern "C" {
void ref_dtgevc3(char side, char howmny, const int *select, int n, const double *S, int lds, const double *P, int ldp,
                 const double *ar, const double *ai, const double *be, double *VL, int ldvl, double *VR, int ldvr,
                 int mm, int *m, double *work, int lwork, int *info)
{
    dtgevc3(side, howmny, select, n, S, lds, P, ldp, ar, ai, be, VL, ldvl, VR, ldvr, mm, m, work, lwork, info);
}
void ref_stgevc3(char side, char howmny, const int *select, int n, const float *S, int lds, const float *P, int ldp,
                 const float *ar, const float *ai, const float *be, float *VL, int ldvl, float *VR, int ldvr, int mm,
                 int *m, float *work, int lwork, int *info)
{
    stgevc3(side, howmny, select, n, S, lds, P, ldp, ar, ai, be, VL, ldvl, VR, ldvr, mm, m, work, lwork, info);
}
void ref_ztgevc3(char side, char howmny, const int *select, int n, const std::complex<double> *S, int lds,
                 const std::complex<double> *P, int ldp, const std::complex<double> *al,
                 const std::complex<double> *be, std::complex<double> *VL, int ldvl, std::complex<double> *VR,
                 int ldvr, int mm, int *m, std::complex<double> *work, int lwork, int *info)
{
    ztgevc3(side, howmny, select, n, S, lds, P, ldp, al, be, VL, ldvl, VR, ldvr, mm, m, work, lwork, info);
}
void ref_ctgevc3(char side, char howmny, const int *select, int n, const std::complex<float> *S, int lds,
                 const std::complex<float> *P, int ldp, const std::complex<float> *al, const std::complex<float> *be,
                 std::complex<float> *VL, int ldvl, std::complex<float> *VR, int ldvr, int mm, int *m,
                 std::complex<float> *work, int lwork, int *info)
{
    ctgevc3(side, howmny, select, n, S, lds, P, ldp, al, be, VL, ldvl, VR, ldvr, mm, m, work, lwork, info);
}
}
