// C-ABI boundary of libxtgevc3 and the drop-in tgevc3.h symbols.
//
// Host-pointer entry points keep the reference's signatures and semantics exactly
// (include/tgevc3.h:10-13; info/m/lwork behaviour of src/dtgevc3.cpp:1173-1242): they stage
// S, P, VL, VR to the GPU, run the gfx950 path and copy the vectors back.  Device-pointer
// entry points take device arrays and a HIP stream and skip the workspace arguments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstring>
#include <future>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/xtgevc3.h"
#include "common.h"
#include "driver.h"

namespace xtg {

namespace {

// Grow-only device array of the host-pointer entry points.
struct DevBuf {
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

// Matrices move between user memory and the device in column chunks of at most STAGE_CHUNK
// bytes through two pinned host buffers: packing one chunk on the host overlaps the DMA of the
// other.  FP32 chunks land in one of two device buffers and are widened (or narrowed, on the way
// back) by a kernel, so PCIe carries FP32 bytes for stgevc3/ctgevc3.
constexpr size_t STAGE_CHUNK = (size_t)32 << 20;

struct Ring {
    void *host[2] = {nullptr, nullptr};
    void *dev[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool armed[2] = {false, false};
    size_t cap = 0;

    // waits until the last copy through buffer b has finished
    hipError_t wait(int b)
    {
        if (!armed[b]) {
            return hipSuccess;
        }
        armed[b] = false;
        return hipEventSynchronize(ev[b]);
    }
    hipError_t mark(int b, hipStream_t st)
    {
        const hipError_t e = hipEventRecord(ev[b], st);
        armed[b] = e == hipSuccess;
        return e;
    }
    bool ensure(size_t colbytes)
    {
        const size_t want = std::max(STAGE_CHUNK, colbytes);
        if (want <= cap) {
            return true;
        }
        if (wait(0) != hipSuccess || wait(1) != hipSuccess) {
            return false;
        }
        cap = 0;
        for (int b = 0; b < 2; ++b) {
            if (host[b]) {
                (void)hipHostFree(host[b]);
                host[b] = nullptr;
            }
            if (dev[b]) {
                (void)hipFree(dev[b]);
                dev[b] = nullptr;
            }
            if (!ev[b] && hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess) {
                ev[b] = nullptr;
                return false;
            }
        }
        for (int b = 0; b < 2; ++b) {
            if (hipHostMalloc(&host[b], want, hipHostMallocDefault) != hipSuccess) {
                host[b] = nullptr;
                return false;
            }
            if (hipMalloc(&dev[b], want) != hipSuccess) {
                dev[b] = nullptr;
                return false;
            }
        }
        cap = want;
        return true;
    }
};

std::mutex g_stage_mutex;
DevBuf g_dS, g_dP, g_dVL, g_dVR, g_dZR, g_dZL, g_dVL32, g_dVR32;
Ring g_ring;

// Per-device copy stream (non-blocking, default priority) and the events that mark the end of
// the right and left Z uploads on it.
struct CopyLane {
    hipStream_t s = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};
    hipEvent_t piece[2 * 8] = {}; // end of each back-transform piece (right, left)
};
constexpr int OUT_PIECES = 8; // back-transform pieces per side whose copy-out overlaps the rest
CopyLane g_lane[64];

int fail(int *info, const char *what)
{
    std::string msg = std::string("xtgevc3: ") + what + ": " + last_error();
    std::fprintf(stderr, "%s\n", msg.c_str());
    if (info) {
        *info = XTGEVC3_INFO_DEVICE_ERROR;
    }
    return XTGEVC3_INFO_DEVICE_ERROR;
}

bool hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return false;
    }
    return true;
}

__global__ void k_widen(const float *__restrict__ src, double *__restrict__ dst, long len)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x) {
        dst[i] = (double)src[i];
    }
}

__global__ void k_narrow(const double *__restrict__ src, float *__restrict__ dst, long len)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x) {
        dst[i] = (float)src[i];
    }
}

unsigned flat_grid(long len) { return (unsigned)std::min<long>((len + 255) / 256, 4096); }

// Copies `cols` columns of `colbytes` bytes between strided host arrays, on up to maxt threads
// when the chunk is large (packing pageable memory is bound by per-thread copy bandwidth).
void copy_cols(char *dst, size_t dstride, const char *src, size_t sstride, size_t colbytes, int cols, int maxt = 8)
{
    const int nt = colbytes * cols >= ((size_t)4 << 20) ? std::min(cols, maxt) : 1;
    auto part = [=](int t) {
        const int c1 = (int)((long)cols * (t + 1) / nt);
        for (int c = (int)((long)cols * t / nt); c < c1; ++c) {
            std::memcpy(dst + c * dstride, src + c * sstride, colbytes);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) {
        try {
            pool.emplace_back(part, t);
        } catch (...) {
            part(t);
        }
    }
    part(0);
    for (auto &th : pool) {
        th.join();
    }
}

// Moves a rows x cols column-major host matrix (leading dimension ld, elements of w scalars of
// type T = double or float) into a packed device matrix of w doubles per element, queued on st.
template <class T>
bool upload(double *dev, const T *host, int rows, int cols, int ld, int w, hipStream_t st, int maxt = 8)
{
    if (rows <= 0 || cols <= 0) {
        return true;
    }
    const size_t colb = (size_t)rows * w * sizeof(T);
    if (!g_ring.ensure(colb)) {
        set_error("pinned staging allocation failed");
        return false;
    }
    const int cc = (int)std::min<size_t>(cols, g_ring.cap / colb);
    for (int c0 = 0, i = 0; c0 < cols; c0 += cc, ++i) {
        const int b = i & 1, nc = std::min(cc, cols - c0);
        if (!hip_ok(g_ring.wait(b), "staging wait")) {
            return false;
        }
        copy_cols((char *)g_ring.host[b], colb, (const char *)(host + (size_t)c0 * ld * w), (size_t)ld * w * sizeof(T),
                  colb, nc, maxt);
        double *d = dev + (size_t)c0 * rows * w;
        if (std::is_same<T, double>::value) {
            if (!hip_ok(hipMemcpyAsync(d, g_ring.host[b], colb * nc, hipMemcpyHostToDevice, st), "host to device copy")) {
                return false;
            }
        } else {
            if (!hip_ok(hipMemcpyAsync(g_ring.dev[b], g_ring.host[b], colb * nc, hipMemcpyHostToDevice, st),
                        "host to device copy")) {
                return false;
            }
            const long len = (long)nc * rows * w;
            hipLaunchKernelGGL(k_widen, dim3(flat_grid(len)), dim3(256), 0, st, (const float *)g_ring.dev[b], d, len);
            if (!hip_ok(hipGetLastError(), "k_widen launch")) {
                return false;
            }
        }
        if (!hip_ok(g_ring.mark(b, st), "staging event")) {
            return false;
        }
    }
    return true;
}

// Moves the first cols columns of a packed device matrix (w scalars of type D per element)
// into a host matrix with leading dimension ld and scalars of type T, rounding to FP32 on the
// device for T = float, D = double; returns when the host copy is complete.
template <class T, class D> bool download(T *host, int ld, const D *dev, int rows, int cols, int w, hipStream_t st)
{
    if (rows <= 0 || cols <= 0) {
        return true;
    }
    const size_t colb = (size_t)rows * w * sizeof(T);
    if (!g_ring.ensure(colb)) {
        set_error("pinned staging allocation failed");
        return false;
    }
    const int cc = (int)std::min<size_t>(cols, g_ring.cap / colb);
    const int nchunks = (cols + cc - 1) / cc;
    auto issue = [&](int i) -> bool {
        const int b = i & 1, c0 = i * cc, nc = std::min(cc, cols - c0);
        if (!hip_ok(g_ring.wait(b), "staging wait")) {
            return false;
        }
        const D *d = dev + (size_t)c0 * rows * w;
        if constexpr (std::is_same<T, D>::value) {
            if (!hip_ok(hipMemcpyAsync(g_ring.host[b], d, colb * nc, hipMemcpyDeviceToHost, st), "device to host copy")) {
                return false;
            }
        } else {
            static_assert(std::is_same<T, float>::value && std::is_same<D, double>::value, "narrowing copy");
            const long len = (long)nc * rows * w;
            hipLaunchKernelGGL(k_narrow, dim3(flat_grid(len)), dim3(256), 0, st, d, (float *)g_ring.dev[b], len);
            if (!hip_ok(hipGetLastError(), "k_narrow launch") ||
                !hip_ok(hipMemcpyAsync(g_ring.host[b], g_ring.dev[b], colb * nc, hipMemcpyDeviceToHost, st),
                        "device to host copy")) {
                return false;
            }
        }
        return hip_ok(g_ring.mark(b, st), "staging event");
    };
    if (!issue(0) || (nchunks > 1 && !issue(1))) {
        return false;
    }
    for (int i = 0; i < nchunks; ++i) {
        const int b = i & 1, c0 = i * cc, nc = std::min(cc, cols - c0);
        if (!hip_ok(g_ring.wait(b), "staging wait")) {
            return false;
        }
        copy_cols((char *)(host + (size_t)c0 * ld * w), (size_t)ld * w * sizeof(T), (const char *)g_ring.host[b], colb,
                  colb, nc);
        if (i + 2 < nchunks && !issue(i + 2)) {
            return false;
        }
    }
    return true;
}

// Z of the back-transform (the input VR / VL of howmny = 'B') crosses PCIe on the device's copy
// stream from a worker thread while the sweep runs on the calling stream: the right matrix
// first, then the left, each followed by an event.  feed.ready(side) blocks until the worker
// has queued that side's copies and returns its event, so the back-transform waits on the
// device only for its own matrix.  finish() joins the worker; every path calls it before the
// host arrays are released to the caller.  With pieces_on the back-transform runs in
// OUT_PIECES pieces per side (rounded to FP32 on the device when feed.out32 is set) and
// feed.done records an event after each;
// copy_out() then brings the pieces back in device order on the copy stream, each behind its
// event, so the PCIe copies of finished columns overlap the device work that is still queued.
constexpr int ZSTAGE_THREADS = 4; // packing threads beside the sweep's launching thread
template <class T> struct ZStage {
    std::thread th;
    std::promise<bool> done[2];
    std::shared_future<bool> fut[2];
    std::string err;
    CopyLane *lane = nullptr;
    ZFeed feed{};
    struct Piece {
        bool left;
        int c0, c1;
        hipEvent_t ev;
    };
    std::vector<Piece> pieces; // back-transform pieces queued so far, in device order
    int npiece[2] = {0, 0};

    bool start(int dev, double *dzr, const T *zr, int ldzr, double *dzl, const T *zl, int ldzl, int n, int w,
               bool pieces_on)
    {
        CopyLane &l = g_lane[dev & 63];
        if (l.s == nullptr) {
            if (!hip_ok(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking), "copy stream")) {
                return false;
            }
            for (hipEvent_t &e : l.done) {
                if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "copy event")) {
                    return false;
                }
            }
            for (hipEvent_t &e : l.piece) {
                if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "copy event")) {
                    return false;
                }
            }
        }
        lane = &l;
        for (int i = 0; i < 2; ++i) {
            fut[i] = done[i].get_future().share();
        }
        feed.zr = dzr;
        feed.zl = dzl;
        feed.ready = [this](bool left) -> hipEvent_t {
            if (!fut[left ? 1 : 0].get()) {
                set_error(err);
                return nullptr;
            }
            return lane->done[left ? 1 : 0];
        };
        feed.chunks = pieces_on ? OUT_PIECES : 1;
        feed.done = [this](bool left, int c0, int c1, hipStream_t st) -> int {
            int &np = npiece[left ? 1 : 0];
            if (np >= OUT_PIECES) {
                set_error("more back-transform pieces than events");
                return -1000;
            }
            const hipEvent_t ev = lane->piece[(left ? OUT_PIECES : 0) + np++];
            if (!hip_ok(hipEventRecord(ev, st), "piece event")) {
                return -1000;
            }
            pieces.push_back(Piece{left, c0, c1, ev});
            return 0;
        };
        auto body = [this, dev, zr, zl, dzr, dzl, ldzr, ldzl, n, w]() {
            bool ok = hip_ok(hipSetDevice(dev), "copy thread device");
            const T *src[2] = {zr, zl};
            double *dst[2] = {dzr, dzl};
            const int ld[2] = {ldzr, ldzl};
            for (int i = 0; i < 2; ++i) {
                ok = ok && (src[i] == nullptr || upload(dst[i], src[i], n, n, ld[i], w, lane->s, ZSTAGE_THREADS)) &&
                     hip_ok(hipEventRecord(lane->done[i], lane->s), "copy event");
                if (!ok && err.empty()) {
                    err = last_error();
                }
                done[i].set_value(ok);
            }
        };
        try {
            th = std::thread(body);
        } catch (...) {
            body();
        }
        return true;
    }
    void finish()
    {
        if (th.joinable()) {
            th.join();
        }
    }
    // Copies the queued pieces of side `left` (device matrix dv of FP64 or, rounded by the
    // device path, of T; n rows, leading dimension n) into the host matrix hv.  Returns false,
    // copying nothing, unless the pieces cover all cols columns (the caller then copies the
    // whole matrix); on a copy error it returns true with ok cleared.
    template <class D> bool copy_out(bool left, T *hv, int ldv, const D *dv, int n, int w, int cols, bool &ok)
    {
        int covered = 0;
        for (const Piece &pc : pieces) {
            if (pc.left == left) {
                covered += pc.c1 - pc.c0;
            }
        }
        if (covered != cols) {
            return false;
        }
        for (const Piece &pc : pieces) {
            if (pc.left != left) {
                continue;
            }
            if (!hip_ok(hipStreamWaitEvent(lane->s, pc.ev, 0), "piece wait") ||
                !download(hv + (size_t)pc.c0 * ldv * w, ldv, dv + (size_t)pc.c0 * n * w, n, pc.c1 - pc.c0, w, lane->s)) {
                ok = false;
                return true;
            }
        }
        return true;
    }
    ~ZStage() { finish(); }
};

// Host-pointer driver for the real interfaces.  T (double, or float for stgevc3) is the host
// scalar type; the eigenvalue arrays are FP64 and the device computes in FP64.
template <class T>
void real_host(char side, char howmny, const int *select, int n, const T *S, int lds, const T *P, int ldp,
               const double *ar, const double *ai, const double *be, T *VL, int ldvl, T *VR, int ldvr, int mm, int *m,
               T *work, int lwork, int *info)
{
    std::vector<Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0 ? real_units(howmny, select, n, ai, units)
                                                                             : 0;
    if (!check_args<T>(howmny, n, num_sel, mm, m, work, lwork, info, true)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    const hipStream_t st = nullptr;
    int dev = 0;
    size_t nn = (size_t)n * n * sizeof(double);
    size_t nm = (size_t)n * num_sel * sizeof(double);
    if (!hip_ok(hipGetDevice(&dev), "device query")) {
        fail(info, "staging");
        return;
    }
    if (!g_dS.ensure(nn) || !g_dP.ensure(nn) || (right && !g_dVR.ensure(nm)) || (left && !g_dVL.ensure(nm)) ||
        (back && right && !g_dZR.ensure(nn)) || (back && left && !g_dZL.ensure(nn))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    double *dS = (double *)g_dS.p, *dP = (double *)g_dP.p, *dVL = (double *)g_dVL.p, *dVR = (double *)g_dVR.p;
    if (!upload(dS, S, n, n, lds, 1, st) || !upload(dP, P, n, n, ldp, 1, st)) {
        fail(info, "staging");
        return;
    }
    ZStage<T> zs;
    // FP32 host arrays: the device path rounds each piece into an FP32 matrix on its own stream
    // (narrowing kernels on the copy stream would wait for CUs behind the running pieces), and
    // the copy stream moves FP32 bytes only
    constexpr bool narrow = !std::is_same<T, double>::value;
    const size_t n32 = (size_t)n * num_sel * sizeof(float);
    if (narrow && back && ((right && !g_dVR32.ensure(n32)) || (left && !g_dVL32.ensure(n32)))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    if (back && !zs.start(dev, (double *)g_dZR.p, right ? VR : nullptr, ldvr, (double *)g_dZL.p, left ? VL : nullptr,
                          ldvl, n, 1, true)) {
        fail(info, "staging");
        return;
    }
    if (narrow) {
        zs.feed.out32[0] = right ? g_dVR32.p : nullptr;
        zs.feed.out32[1] = left ? g_dVL32.p : nullptr;
    }
    int rc = real_device(side, howmny, select, n, dS, n, dP, n, ar, ai, be, left ? dVL : nullptr, n,
                         right ? dVR : nullptr, n, st, XTGEVC3_PREC_FP64, back ? &zs.feed : nullptr,
                         narrow ? kFltMin : kDblMin);
    zs.finish();
    if (rc != 0) {
        (void)hipDeviceSynchronize();
        fail(info, "device path");
        return;
    }
    // back-transform pieces first (their copies overlap the queued device work), then the rest
    bool ok = true, got_r = false, got_l = false;
    if (back && right) {
        got_r = narrow ? zs.copy_out(false, VR, ldvr, (const T *)g_dVR32.p, n, 1, num_sel, ok)
                       : zs.copy_out(false, VR, ldvr, dVR, n, 1, num_sel, ok);
    }
    if (ok && back && left) {
        got_l = narrow ? zs.copy_out(true, VL, ldvl, (const T *)g_dVL32.p, n, 1, num_sel, ok)
                       : zs.copy_out(true, VL, ldvl, dVL, n, 1, num_sel, ok);
    }
    if (!ok || !hip_ok(hipDeviceSynchronize(), "kernel execution")) {
        fail(info, "device path");
        return;
    }
    if ((right && !got_r && !download(VR, ldvr, dVR, n, num_sel, 1, st)) ||
        (left && !got_l && !download(VL, ldvl, dVL, n, num_sel, 1, st))) {
        fail(info, "staging");
        return;
    }
}

// Host-pointer driver for the complex interfaces.  T (double, or float for ctgevc3) is the
// scalar type of the host std::complex<T> arrays; the device computes in complex128.
template <class T>
void complex_host(char side, char howmny, const int *select, int n, const std::complex<T> *S, int lds,
                  const std::complex<T> *P, int ldp, const std::complex<double> *al, const std::complex<double> *be,
                  std::complex<T> *VL, int ldvl, std::complex<T> *VR, int ldvr, int mm, int *m, std::complex<T> *work,
                  int lwork, int *info)
{
    typedef std::complex<double> zt;
    std::vector<Unit> units;
    int num_sel = (howmny == 'A' || howmny == 'B' || howmny == 'S') && n > 0 ? complex_units(howmny, select, n, units)
                                                                             : 0;
    if (!check_args<std::complex<T>>(howmny, n, num_sel, mm, m, work, lwork, info, true)) {
        return;
    }
    bool right = (side == 'R' || side == 'B') && VR != nullptr;
    bool left = (side == 'L' || side == 'B') && VL != nullptr;
    if (num_sel == 0 || (!right && !left)) {
        return;
    }
    bool back = howmny == 'B';
    std::lock_guard<std::mutex> lock(g_stage_mutex);
    const hipStream_t st = nullptr;
    int dev = 0;
    size_t nn = (size_t)n * n * sizeof(zt);
    size_t nm = (size_t)n * num_sel * sizeof(zt);
    if (!hip_ok(hipGetDevice(&dev), "device query")) {
        fail(info, "staging");
        return;
    }
    if (!g_dS.ensure(nn) || !g_dP.ensure(nn) || (right && !g_dVR.ensure(nm)) || (left && !g_dVL.ensure(nm)) ||
        (back && right && !g_dZR.ensure(nn)) || (back && left && !g_dZL.ensure(nn))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    double *dS = (double *)g_dS.p, *dP = (double *)g_dP.p, *dVL = (double *)g_dVL.p, *dVR = (double *)g_dVR.p;
    const T *hS = reinterpret_cast<const T *>(S), *hP = reinterpret_cast<const T *>(P);
    T *hVL = reinterpret_cast<T *>(VL), *hVR = reinterpret_cast<T *>(VR);
    if (!upload(dS, hS, n, n, lds, 2, st) || !upload(dP, hP, n, n, ldp, 2, st)) {
        fail(info, "staging");
        return;
    }
    ZStage<T> zs;
    // complex64 host arrays: pieces rounded on the device path's stream, as in real_host
    constexpr bool narrow = !std::is_same<T, double>::value;
    const size_t n32 = (size_t)n * num_sel * 2 * sizeof(float);
    if (narrow && back && ((right && !g_dVR32.ensure(n32)) || (left && !g_dVL32.ensure(n32)))) {
        set_error("device allocation failed");
        fail(info, "staging");
        return;
    }
    if (back && !zs.start(dev, (double *)g_dZR.p, right ? hVR : nullptr, ldvr, (double *)g_dZL.p, left ? hVL : nullptr,
                          ldvl, n, 2, true)) {
        fail(info, "staging");
        return;
    }
    if (narrow) {
        zs.feed.out32[0] = right ? g_dVR32.p : nullptr;
        zs.feed.out32[1] = left ? g_dVL32.p : nullptr;
    }
    int rc = complex_device(side, howmny, select, n, (const zt *)dS, n, (const zt *)dP, n, al, be,
                            left ? (zt *)dVL : nullptr, n, right ? (zt *)dVR : nullptr, n, st, XTGEVC3_PREC_FP64,
                            back ? &zs.feed : nullptr, narrow ? kFltMin : kDblMin);
    zs.finish();
    if (rc != 0) {
        (void)hipDeviceSynchronize();
        fail(info, "device path");
        return;
    }
    bool ok = true, got_r = false, got_l = false;
    if (back && right) {
        got_r = narrow ? zs.copy_out(false, hVR, ldvr, (const T *)g_dVR32.p, n, 2, num_sel, ok)
                       : zs.copy_out(false, hVR, ldvr, dVR, n, 2, num_sel, ok);
    }
    if (ok && back && left) {
        got_l = narrow ? zs.copy_out(true, hVL, ldvl, (const T *)g_dVL32.p, n, 2, num_sel, ok)
                       : zs.copy_out(true, hVL, ldvl, dVL, n, 2, num_sel, ok);
    }
    if (!ok || !hip_ok(hipDeviceSynchronize(), "kernel execution")) {
        fail(info, "device path");
        return;
    }
    if ((right && !got_r && !download(hVR, ldvr, dVR, n, num_sel, 2, st)) ||
        (left && !got_l && !download(hVL, ldvl, dVL, n, num_sel, 2, st))) {
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
    real_host<double>(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR, ldvr, mm, m, work,
                      lwork, info);
}

void xtgevc3_stgevc3(char side, char howmny, const int *select, int n, const float *S, int lds, const float *P,
                     int ldp, const float *alphar, const float *alphai, const float *beta, float *VL, int ldvl,
                     float *VR, int ldvr, int mm, int *m, float *work, int lwork, int *info)
{
    // FP32 interface: same argument semantics (stgevc3.cpp:1122-1242); computed on the FP64
    // path and rounded back, so its vectors are at least as accurate as the FP32 reference's.
    // Matrices cross PCIe in FP32 and are widened and rounded on the device.
    const int nv = n > 0 ? n : 0;
    std::vector<double> ar_d(nv), ai_d(nv), be_d(nv);
    for (int k = 0; k < nv; ++k) {
        ar_d[k] = alphar[k];
        ai_d[k] = alphai[k];
        be_d[k] = beta[k];
    }
    real_host<float>(side, howmny, select, n, S, lds, P, ldp, ar_d.data(), ai_d.data(), be_d.data(), VL, ldvl, VR,
                     ldvr, mm, m, work, lwork, info);
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

int xtgevc3_dtgevc3_device_noalloc(char side, char howmny, const int *select, int n, const double *S, int lds,
                                   const double *P, int ldp, const double *alphar, const double *alphai,
                                   const double *beta, double *VL, int ldvl, double *VR, int ldvr, int mm, int *m,
                                   int *info, void *stream)
{
    if (n > XTGEVC3_NOALLOC_MAX_N) {
        xtg::set_error("n = " + std::to_string(n) + " exceeds XTGEVC3_NOALLOC_MAX_N = " +
                       std::to_string(XTGEVC3_NOALLOC_MAX_N) + " of the allocation-free entry");
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
    int rc = xtg::real_device_noalloc(side, howmny, select, n, S, lds, P, ldp, alphar, alphai, beta, VL, ldvl, VR,
                                      ldvr, (hipStream_t)stream);
    if (rc == XTGEVC3_INFO_UNSUPPORTED) {
        if (info) {
            *info = XTGEVC3_INFO_UNSUPPORTED;
        }
        return rc;
    }
    if (rc != 0) {
        return xtg::fail(info, "device path");
    }
    return 0;
}


void xtgevc3_ztgevc3(char side, char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                     const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *alpha, const xtgevc3_zcomplex *beta,
                     xtgevc3_zcomplex *VL, int ldvl, xtgevc3_zcomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_zcomplex *work, int lwork, int *info)
{
    complex_host<double>(side, howmny, select, n, S, lds, P, ldp, alpha, beta, VL, ldvl, VR, ldvr, mm, m, work, lwork,
                         info);
}

void xtgevc3_ctgevc3(char side, char howmny, const int *select, int n, const xtgevc3_ccomplex *S, int lds,
                     const xtgevc3_ccomplex *P, int ldp, const xtgevc3_ccomplex *alpha, const xtgevc3_ccomplex *beta,
                     xtgevc3_ccomplex *VL, int ldvl, xtgevc3_ccomplex *VR, int ldvr, int mm, int *m,
                     xtgevc3_ccomplex *work, int lwork, int *info)
{
    // complex64 interface: argument semantics of ctgevc3.cpp:435-516, computed on the
    // complex128 path and rounded back; matrices cross PCIe in complex64.
    typedef std::complex<double> zt;
    const int nv = n > 0 ? n : 0;
    std::vector<zt> al(nv), be(nv);
    for (int k = 0; k < nv; ++k) {
        al[k] = zt(alpha[k]);
        be[k] = zt(beta[k]);
    }
    complex_host<float>(side, howmny, select, n, S, lds, P, ldp, al.data(), be.data(), VL, ldvl, VR, ldvr, mm, m, work,
                        lwork, info);
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
    return xtgevc3_dtgevc3_backward_device_ex(n, S, lds, P, ldp, alphar, alphai, beta, select, V, ldv, G, ldg, nbatch,
                                              dS, ldds, dP, lddp, info, stream, XTGEVC3_PREC_FP64);
}

int xtgevc3_dtgevc3_backward_device_ex(int n, const double *S, int lds, const double *P, int ldp,
                                       const double *alphar, const double *alphai, const double *beta,
                                       const int *select, const double *V, int ldv, const double *G, int ldg,
                                       int nbatch, double *dS, int ldds, double *dP, int lddp, int *info, void *stream,
                                       int precision)
{
    if (precision != XTGEVC3_PREC_FP64 && precision != XTGEVC3_PREC_MIXED_F32 &&
        precision != XTGEVC3_PREC_SPLIT_I8) {
        xtg::set_error("precision " + std::to_string(precision) + " is not available for the backward");
        if (info) {
            *info = XTGEVC3_INFO_UNSUPPORTED;
        }
        return XTGEVC3_INFO_UNSUPPORTED;
    }
    int rc = xtg::real_backward_device(n, S, lds, P, ldp, alphar, alphai, beta, select, V, ldv, G, ldg, nbatch, dS,
                                       ldds, dP, lddp, info, (hipStream_t)stream, precision);
    if (rc != 0) {
        return xtg::fail(info, "backward");
    }
    return 0;
}

int xtgevc3_dtgevc3_backward_full_device(char side, int n, const double *S, int lds, const double *P, int ldp,
                                         const double *alphar, const double *alphai, const double *beta,
                                         const int *select, const double *Zb, int ldz, const double *V, int ldv,
                                         const double *G, int ldg, int nbatch, double *dS, int ldds, double *dP,
                                         int lddp, double *dZ, int lddz, int *info, void *stream)
{
    if (side != 'R' && side != 'L') {
        xtg::set_error(std::string("side must be 'R' or 'L', got '") + side + "'");
        if (info) {
            *info = -1;
        }
        return -1;
    }
    int rc = xtg::real_backward_full_device(side, n, S, lds, P, ldp, alphar, alphai, beta, select, Zb, ldz, V, ldv,
                                            G, ldg, nbatch, dS, ldds, dP, lddp, dZ, lddz, info,
                                            (hipStream_t)stream);
    if (rc != 0) {
        return xtg::fail(info, "backward");
    }
    return 0;
}

int xtgevc3_dtgsna_e_device(char howmny, const int *select, int n, const double *S, int lds, const double *P,
                            int ldp, const double *alphai, const double *VL, int ldvl, const double *VR, int ldvr,
                            int m, double *s, int *info, void *stream)
{
    if (info) {
        *info = 0;
    }
    if (howmny != 'A' && howmny != 'S') {
        xtg::set_error("howmny must be 'A' or 'S'");
        if (info) {
            *info = -1;
        }
        return -1;
    }
    if (n < 0 || m < 0 || (n > 0 && (lds < n || ldp < n || ldvl < n || ldvr < n))) {
        xtg::set_error("invalid sizes");
        if (info) {
            *info = -3;
        }
        return -3;
    }
    if (n == 0) {
        return 0;
    }
    int rc = xtg::real_tgsna_device(howmny, select, n, S, lds, P, ldp, alphai, VL, ldvl, VR, ldvr, m, s,
                                    (hipStream_t)stream);
    if (rc != 0) {
        return xtg::fail(info, "dtgsna_e");
    }
    return 0;
}

int xtgevc3_ztgsna_e_device(char howmny, const int *select, int n, const xtgevc3_zcomplex *S, int lds,
                            const xtgevc3_zcomplex *P, int ldp, const xtgevc3_zcomplex *VL, int ldvl,
                            const xtgevc3_zcomplex *VR, int ldvr, int m, double *s, int *info, void *stream)
{
    if (info) {
        *info = 0;
    }
    if (howmny != 'A' && howmny != 'S') {
        xtg::set_error("howmny must be 'A' or 'S'");
        if (info) {
            *info = -1;
        }
        return -1;
    }
    if (n < 0 || m < 0 || (n > 0 && (lds < n || ldp < n || ldvl < n || ldvr < n))) {
        xtg::set_error("invalid sizes");
        if (info) {
            *info = -3;
        }
        return -3;
    }
    if (n == 0) {
        return 0;
    }
    int rc = xtg::complex_tgsna_device(howmny, select, n, S, lds, P, ldp, VL, ldvl, VR, ldvr, m, s,
                                       (hipStream_t)stream);
    if (rc != 0) {
        return xtg::fail(info, "ztgsna_e");
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
