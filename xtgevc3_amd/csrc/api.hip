// C-ABI boundary of libxtgevc3 and the drop-in tgevc3.h symbols.
//
// Host-pointer entry points keep the reference's signatures and semantics exactly
// (include/tgevc3.h:10-13; info/m/lwork behaviour of src/dtgevc3.cpp:1173-1242): they stage
// S, P, VL, VR to the GPU, run the gfx950 path and copy the vectors back.  Device-pointer
// entry points take device arrays and a HIP stream and skip the workspace arguments.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/xtgevc3.h"
#include "common.h"
#include "driver.h"

namespace xtg {

namespace {

struct HostStage {
    void *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes)
    {
        if (bytes <= cap) {
            return true;
        }
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
            return false;
        }
        cap = std::max<size_t>(bytes, 256);
        return true;
    }
};

std::mutex g_stage_mutex;
HostStage g_dS, g_dP, g_dVL, g_dVR;

int fail(int *info, const char *what)
{
    std::string msg = std::string("xtgevc3: ") + what + ": " + last_error();
    std::fprintf(stderr, "%s\n", msg.c_str());
    if (info) {
        *info = XTGEVC3_INFO_DEVICE_ERROR;
    }
    return XTGEVC3_INFO_DEVICE_ERROR;
}

template <class T> hipError_t put(void *dst, const T *src, int rows, int cols, int ld)
{
    return hipMemcpy2D(dst, sizeof(T) * rows, src, sizeof(T) * ld, sizeof(T) * rows, cols, hipMemcpyHostToDevice);
}

template <class T> hipError_t get(T *dst, int ld, const void *src, int rows, int cols)
{
    return hipMemcpy2D(dst, sizeof(T) * ld, src, sizeof(T) * rows, sizeof(T) * rows, cols, hipMemcpyDeviceToHost);
}

// Host-pointer driver for the double-precision real interface.
void real_host(char side, char howmny, const int *select, int n, const double *S, int lds, const double *P, int ldp,
               const double *ar, const double *ai, const double *be, double *VL, int ldvl, double *VR, int ldvr,
               int mm, int *m, double *work, int lwork, int *info, bool have_work)
{
    std::vector<Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0 ? real_units(howmny, select, n, ai, units)
                                                                             : 0;
    if (!check_args<double>(howmny, n, num_sel, mm, m, work, lwork, info, have_work)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    size_t nn = (size_t)n * n * sizeof(double);
    size_t nm = (size_t)n * (back ? n : num_sel) * sizeof(double);
    if (!g_dS.ensure(nn) || !g_dP.ensure(nn) || (right && !g_dVR.ensure(nm)) || (left && !g_dVL.ensure(nm))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    if (put(g_dS.p, S, n, n, lds) != hipSuccess || put(g_dP.p, P, n, n, ldp) != hipSuccess) {
        set_error("host to device copy failed");
        fail(info, "staging");
        return;
    }
    if (back) {
        if ((right && put(g_dVR.p, VR, n, n, ldvr) != hipSuccess) ||
            (left && put(g_dVL.p, VL, n, n, ldvl) != hipSuccess)) {
            set_error("host to device copy failed");
            fail(info, "staging");
            return;
        }
    }
    int rc = real_device(side, howmny, select, n, (const double *)g_dS.p, n, (const double *)g_dP.p, n, ar, ai, be,
                         left ? (double *)g_dVL.p : nullptr, n, right ? (double *)g_dVR.p : nullptr, n, nullptr);
    if (rc == 0) {
        const hipError_t se = hipDeviceSynchronize();
        if (se != hipSuccess) {
            set_error(std::string("kernel execution: ") + hipGetErrorString(se));
            rc = -1;
        }
    }
    if (rc != 0) {
        fail(info, "device path");
        return;
    }
    if ((right && get(VR, ldvr, g_dVR.p, n, num_sel) != hipSuccess) ||
        (left && get(VL, ldvl, g_dVL.p, n, num_sel) != hipSuccess)) {
        set_error("device to host copy failed");
        fail(info, "staging");
        return;
    }
}


// Host-pointer driver for the complex128 interface.
void complex_host(char side, char howmny, const int *select, int n, const std::complex<double> *S, int lds,
                  const std::complex<double> *P, int ldp, const std::complex<double> *al,
                  const std::complex<double> *be, std::complex<double> *VL, int ldvl, std::complex<double> *VR,
                  int ldvr, int mm, int *m, std::complex<double> *work, int lwork, int *info, bool have_work)
{
    typedef std::complex<double> zt;
    std::vector<Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0 ? complex_units(howmny, select, n, units)
                                                                             : 0;
    if (!check_args<zt>(howmny, n, num_sel, mm, m, work, lwork, info, have_work)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    size_t nn = (size_t)n * n * sizeof(zt);
    size_t nm = (size_t)n * (back ? n : num_sel) * sizeof(zt);
    if (!g_dS.ensure(nn) || !g_dP.ensure(nn) || (right && !g_dVR.ensure(nm)) || (left && !g_dVL.ensure(nm))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    if (put(g_dS.p, S, n, n, lds) != hipSuccess || put(g_dP.p, P, n, n, ldp) != hipSuccess) {
        set_error("host to device copy failed");
        fail(info, "staging");
        return;
    }
    if (back) {
        if ((right && put(g_dVR.p, VR, n, n, ldvr) != hipSuccess) ||
            (left && put(g_dVL.p, VL, n, n, ldvl) != hipSuccess)) {
            set_error("host to device copy failed");
            fail(info, "staging");
            return;
        }
    }
    int rc = complex_device(side, howmny, select, n, (const zt *)g_dS.p, n, (const zt *)g_dP.p, n, al, be,
                            left ? (zt *)g_dVL.p : nullptr, n, right ? (zt *)g_dVR.p : nullptr, n, nullptr);
    if (rc == 0) {
        const hipError_t se = hipDeviceSynchronize();
        if (se != hipSuccess) {
            set_error(std::string("kernel execution: ") + hipGetErrorString(se));
            rc = -1;
        }
    }
    if (rc != 0) {
        fail(info, "device path");
        return;
    }
    if ((right && get(VR, ldvr, g_dVR.p, n, num_sel) != hipSuccess) ||
        (left && get(VL, ldvl, g_dVL.p, n, num_sel) != hipSuccess)) {
        set_error("device to host copy failed");
        fail(info, "staging");
        return;
    }
}

} // namespace
} // namespace xtg

using xtg::complex_host;
using xtg::real_host;

extern "C" {

const char *xtgevc3_last_error(void) { return xtg::last_error(); }

int xtgevc3_abi_version(void) { return XTGEVC3_ABI_VERSION; }

void xtgevc3_dtgevc3(char side, char howmny, const int *select, int n, const double *S, int lds, const double *P,
                     int ldp, const double *alphar, const double *alphai, const double *beta, double *VL, int ldvl,
                     double *VR, int ldvr, int mm, int *m, double *work, int lwork, int *info)
{
    real_host(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR, ldvr, mm, m, work, lwork,
              info, true);
}

void xtgevc3_stgevc3(char side, char howmny, const int *select, int n, const float *S, int lds, const float *P,
                     int ldp, const float *alphar, const float *alphai, const float *beta, float *VL, int ldvl,
                     float *VR, int ldvr, int mm, int *m, float *work, int lwork, int *info)
{
    // FP32 interface: same argument semantics (stgevc3.cpp:1122-1242); computed on the FP64
    // path and rounded back, so its vectors are at least as accurate as the FP32 reference's.
    std::vector<xtg::Unit> units;
    std::vector<double> ai_d(n > 0 ? n : 0);
    for (int k = 0; k < n; ++k) {
        ai_d[k] = alphai[k];
    }
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0
                      ? xtg::real_units(howmny, select, n, ai_d.data(), units)
                      : 0;
    if (!xtg::check_args<float>(howmny, n, num_sel, mm, m, work, lwork, info, true)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::vector<double> Sd((size_t)n * n), Pd((size_t)n * n), ar_d(n), be_d(n);
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r) {
            Sd[r + (size_t)c * n] = S[r + (size_t)c * lds];
            Pd[r + (size_t)c * n] = P[r + (size_t)c * ldp];
        }
        ar_d[c] = alphar[c];
        be_d[c] = beta[c];
    }
    int vcols = back ? n : num_sel;
    std::vector<double> VLd(left ? (size_t)n * vcols : 0), VRd(right ? (size_t)n * vcols : 0);
    if (back) {
        for (int c = 0; c < n; ++c) {
            for (int r = 0; r < n; ++r) {
                if (left) {
                    VLd[r + (size_t)c * n] = VL[r + (size_t)c * ldvl];
                }
                if (right) {
                    VRd[r + (size_t)c * n] = VR[r + (size_t)c * ldvr];
                }
            }
        }
    }
    int md = 0, infod = 0;
    double wq = 0.0;
    real_host(side, howmny, select, n, Sd.data(), n, Pd.data(), n, ar_d.data(), ai_d.data(), be_d.data(),
              left ? VLd.data() : nullptr, n, right ? VRd.data() : nullptr, n, vcols, &md, &wq, 0, &infod, false);
    if (infod != 0) {
        *info = infod;
        return;
    }
    for (int c = 0; c < num_sel; ++c) {
        for (int r = 0; r < n; ++r) {
            if (left) {
                VL[r + (size_t)c * ldvl] = (float)VLd[r + (size_t)c * n];
            }
            if (right) {
                VR[r + (size_t)c * ldvr] = (float)VRd[r + (size_t)c * n];
            }
        }
    }
}

int xtgevc3_dtgevc3_device_ex(char side, char howmny, const int *select, int n, const double *S, int lds,
                              const double *P, int ldp, const double *alphar, const double *alphai,
                              const double *beta, double *VL, int ldvl, double *VR, int ldvr, int mm, int *m,
                              int *info, void *stream, int precision)
{
    if (precision != XTGEVC3_PREC_FP64 && precision != XTGEVC3_PREC_MIXED_F32 &&
        precision != XTGEVC3_PREC_SPLIT_I8) {
        xtg::set_error("precision " + std::to_string(precision) + " is not available");
        if (info) {
            *info = XTGEVC3_INFO_UNSUPPORTED;
        }
        return XTGEVC3_INFO_UNSUPPORTED;
    }
    std::vector<xtg::Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0
                      ? xtg::real_units(howmny, select, n, alphai, units)
                      : 0;
    if (!xtg::check_args<double>(howmny, n, num_sel, mm, m, (double *)nullptr, 0, info, false)) {
        return 0;
    }
    int rc = xtg::real_device(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR, ldvr,
                              (hipStream_t)stream, precision);
    if (rc != 0) {
        return xtg::fail(info, "device path");
    }
    return 0;
}

int xtgevc3_dtgevc3_device(char side, char howmny, const int *select, int n, const double *S, int lds,
                           const double *P, int ldp, const double *alphar, const double *alphai, const double *beta,
                           double *VL, int ldvl, double *VR, int ldvr, int mm, int *m, int *info, void *stream)
{
    return xtgevc3_dtgevc3_device_ex(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR,
                                     ldvr, mm, m, info, stream, XTGEVC3_PREC_FP64);
}


void xtgevc3_ztgevc3(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                     const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha, const xtgevc3_zcomplex *beta,
                     xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_zcomplex *work, int lwork, int *info)
{
    complex_host(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr, mm, m, work, lwork, info,
                 true);
}

void xtgevc3_ctgevc3(char side, char howmny, const int *select, int n, const xtgevc3_ccomplex *S, int lds,
                     const xtgevc3_ccomplex *P, int ldp, const xtgevc3_ccomplex *alpha, const xtgevc3_ccomplex *beta,
                     xtgevc3_ccomplex *VL, int ldvl, xtgevc3_ccomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_ccomplex *work, int lwork, int *info)
{
    // complex64 interface: argument semantics of ctgevc3.cpp:435-516, computed on the
    // complex128 path and rounded back.
    typedef std::complex<double> zt;
    std::vector<xtg::Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0
                      ? xtg::complex_units(howmny, select, n, units)
                      : 0;
    if (!xtg::check_args<xtgevc3_ccomplex>(howmny, n, num_sel, mm, m, work, lwork, info, true)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::vector<zt> Sd((size_t)n * n), Pd((size_t)n * n), al(n), be(n);
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r) {
            Sd[r + (size_t)c * n] = zt(S[r + (size_t)c * lds]);
            Pd[r + (size_t)c * n] = zt(P[r + (size_t)c * ldp]);
        }
        al[c] = zt(alpha[c]);
        be[c] = zt(beta[c]);
    }
    int vcols = back ? n : num_sel;
    std::vector<zt> VLd(left ? (size_t)n * vcols : 0), VRd(right ? (size_t)n * vcols : 0);
    if (back) {
        for (int c = 0; c < n; ++c) {
            for (int r = 0; r < n; ++r) {
                if (left) {
                    VLd[r + (size_t)c * n] = zt(VL[r + (size_t)c * ldvl]);
                }
                if (right) {
                    VRd[r + (size_t)c * n] = zt(VR[r + (size_t)c * ldvr]);
                }
            }
        }
    }
    int md = 0, infod = 0;
    zt wq;
    complex_host(side, howmny, select, n, Sd.data(), n, Pd.data(), n, al.data(), be.data(), left ? VLd.data() : nullptr,
                 n, right ? VRd.data() : nullptr, n, vcols, &md, &wq, 0, &infod, false);
    if (infod != 0) {
        *info = infod;
        return;
    }
    for (int c = 0; c < num_sel; ++c) {
        for (int r = 0; r < n; ++r) {
            if (left) {
                VL[r + (size_t)c * ldvl] = xtgevc3_ccomplex(VLd[r + (size_t)c * n]);
            }
            if (right) {
                VR[r + (size_t)c * ldvr] = xtgevc3_ccomplex(VRd[r + (size_t)c * n]);
            }
        }
    }
}

int xtgevc3_ztgevc3_device(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                           const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha,
                           const xtgevc3_zcomplex *beta, xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR,
                           int ldvr, int mm, int *m, int *info, void *stream)
{
    return xtgevc3_ztgevc3_device_ex(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr, mm, m,
                                     info, stream, XTGEVC3_PREC_FP64);
}

int xtgevc3_ztgevc3_device_ex(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                              const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha,
                              const xtgevc3_zcomplex *beta, xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR,
                              int ldvr, int mm, int *m, int *info, void *stream, int precision)
{
    if (precision != XTGEVC3_PREC_FP64 && precision != XTGEVC3_PREC_SPLIT_I8) {
        xtg::set_error("precision " + std::to_string(precision) + " is not available for ztgevc3");
        if (info) {
            *info = XTGEVC3_INFO_UNSUPPORTED;
        }
        return XTGEVC3_INFO_UNSUPPORTED;
    }
    std::vector<xtg::Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0
                      ? xtg::complex_units(howmny, select, n, units)
                      : 0;
    if (!xtg::check_args<xtgevc3_zcomplex>(howmny, n, num_sel, mm, m, (xtgevc3_zcomplex *)nullptr, 0, info, false)) {
        return 0;
    }
    int rc = xtg::complex_device(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr,
                                 (hipStream_t)stream, precision);
    if (rc != 0) {
        return xtg::fail(info, "device path");
    }
    return 0;
}


int xtgevc3_dtgevc3_backward_device(int n, const double *S, int lds, const double *P, int ldp, const double *alphar,
                                    const double *alphai, const double *beta, const int *select, const double *V,
                                    int ldv, const double *G, int ldg, int nbatch, double *dS, int ldds, double *dP,
                                    int lddp, int *info, void *stream)
{
    int rc = xtg::real_backward_device(n, S, lds, P, ldp, alphar, alphai, beta, select, V, ldv, G, ldg, nbatch, dS,
                                       ldds, dP, lddp, info, (hipStream_t)stream);
    if (rc != 0) {
        return xtg::fail(info, "backward");
    }
    return 0;
}

} // extern "C"

// Drop-in C++-linkage symbols of include/tgevc3.h:10-11.
void dtgevc3(char side, char howmny, const int *select, int n, const double *S, int lds, const double *P, int ldp,
             const double *alphar, const double *alphai, const double *beta, double *__restrict VL, int ldvl,
             double *__restrict VR, int ldvr, int mm, int *m, double *__restrict work, int lwork, int *info)
{
    xtgevc3_dtgevc3(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR, ldvr, mm, m, work,
                    lwork, info);
}

void stgevc3(char side, char howmny, const int *select, int n, const float *S, int lds, const float *P, int ldp,
             const float *alphar, const float *alphai, const float *beta, float *__restrict VL, int ldvl,
             float *__restrict VR, int ldvr, int mm, int *m, float *__restrict work, int lwork, int *info)
{
    xtgevc3_stgevc3(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR, ldvr, mm, m, work,
                    lwork, info);
}

void ztgevc3(char side, char howmny, const int *select, int n, const std::complex<double> *S, int lds,
             const std::complex<double> *P, int ldp, const std::complex<double> *alpha,
             const std::complex<double> *beta, std::complex<double> *__restrict VL, int ldvl,
             std::complex<double> *__restrict VR, int ldvr, int mm, int *m, std::complex<double> *__restrict work,
             int lwork, int *info)
{
    xtgevc3_ztgevc3(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr, mm, m, work, lwork,
                    info);
}

void ctgevc3(char side, char howmny, const int *select, int n, const std::complex<float> *S, int lds,
             const std::complex<float> *P, int ldp, const std::complex<float> *alpha, const std::complex<float> *beta,
             std::complex<float> *__restrict VL, int ldvl, std::complex<float> *__restrict VR, int ldvr, int mm,
             int *m, std::complex<float> *__restrict work, int lwork, int *info)
{
    xtgevc3_ctgevc3(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr, mm, m, work, lwork,
                    info);
}
