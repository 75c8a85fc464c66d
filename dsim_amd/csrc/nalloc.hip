// Allocation-free per-direction readout (include/dsim.h,
// dsim_signal_moment_nalloc).  It runs the kernels of dsim_signal_moment over
// device arrays the caller owns and keeps its partial sums in a fixed device
// array of this library instead of memory it would allocate, so that callers
// that cannot call hipMalloc (code under HIP graph capture, serving loops,
// harnesses that forbid it) can form signals.
//
// The array is one per GPU, shared by every call on that GPU and holding the
// last call's contents; calls that overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/dsim.h"
#include "dsim_internal.h"
#include "host_common.h"

namespace {

constexpr uint32_t kMaxDirs = DSIM_NALLOC_MAX_DIRECTIONS;
constexpr uint32_t kBlocks = 256;  // partial-sum blocks per direction, as dsim_signal_moment
constexpr size_t kPartialDoubles = (size_t)kMaxDirs * kBlocks * 15;
constexpr size_t kScratchDoubles = kPartialDoubles + (size_t)kMaxDirs * 15;  // + final sums

// 256 x 256 x 15 + 256 x 15 doubles = 7.9 MB per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in; kernels
// get the device address found by hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up once per device,
// with that device current).
int scratch_ptr(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return dsim_host::fail(DSIM_ERR_STATE, "device %d", dev);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return dsim_host::fail(DSIM_ERR_HIP, "hipGetSymbolAddress(scratch): %s",
                                   hipGetErrorString(e));
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return DSIM_OK;
}

// out[j*5 + k] = |sum m e^{i phase}| / count from the final sums, on the
// device (dsim_signal_moment does the same division on the host).
__global__ void k_finish(const double *sums, uint32_t ndir, double *out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ndir * 5) return;
    const uint32_t j = t / 5, k = t % 5;
    const double *q = sums + (size_t)j * 15;
    out[t] = q[10 + k] == 0.0 ? 0.0 : sqrt(q[k] * q[k] + q[5 + k] * q[5 + k]) / q[10 + k];
}

}  // namespace

extern "C" int dsim_signal_moment_nalloc(uint32_t n, const float *mag, const float *moment,
                                         const float *x0, float wsum, const uint32_t *label,
                                         uint32_t ndir, const float *grads_scaled,
                                         float phase_constant, double *out, void *stream) {
    if (ndir == 0) return DSIM_OK;
    if (ndir > kMaxDirs)
        return dsim_host::fail(DSIM_ERR_ARG,
                               "dsim_signal_moment_nalloc: %u directions, at most %u (fixed "
                               "scratch)", ndir, kMaxDirs);
    if (!out || !grads_scaled || (n > 0 && (!mag || !moment || !label)))
        return dsim_host::fail(DSIM_ERR_ARG, "dsim_signal_moment_nalloc: NULL argument");
    if (wsum != 0.0f && n > 0 && !x0)
        return dsim_host::fail(DSIM_ERR_ARG, "dsim_signal_moment_nalloc: x0 needed when wsum != 0");
    double *base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *partial = base, *sums = base + kPartialDoubles;
    hipError_t e = dsim::launch_moment_signal(n, mag, moment, x0, wsum, label, ndir, grads_scaled,
                                              phase_constant, partial, kBlocks, sums, s);
    if (e != hipSuccess)
        return dsim_host::fail(DSIM_ERR_HIP, "moment signal: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_finish, dim3((ndir * 5 + 255) / 256), dim3(256), 0, s, sums, ndir, out);
    e = hipGetLastError();
    if (e != hipSuccess) return dsim_host::fail(DSIM_ERR_HIP, "finish: %s", hipGetErrorString(e));
    return DSIM_OK;
}
