// Waveform phases of the dsim C ABI (include/dsim_phase.h): one walk of a
// simulation handle, its displacement recorded chunk by chunk and multiplied
// by the waveforms' gradients on the matrix cores (k_phase_gemm), FP64
// readout (k_phase_partial).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/dsim.h"
#include "../../include/dsim_phase.h"
#include "dsim_internal.h"
#include "host_common.h"
#include "sim_internal.h"

using dsim_host::DevBuf;
using dsim_host::fail;
using dsim_host::upload;

struct dsim_waveforms {
    dsim_sim *sim = nullptr;
    uint32_t n = 0, n_pad = 0, n_wave = 0, n_steps = 0, chunk = 0, done = 0;
    uint32_t n_wpad = 0;          // waveforms padded with zero ones to a multiple of kPhaseTileN
    uint32_t buffer = 0;          // steps walked per trajectory buffer (a multiple of chunk)
    DevBuf<uint16_t> g_hi, g_lo;  // [n_wpad][3*n_steps] bf16 bits, pc folded in
    DevBuf<double> gsum;          // [n_wpad][3]: sum_t pc * G_w(t)
    DevBuf<float> traj;           // n_pad / kTrajTile tiles of [3*len][kTrajTile] (len <= buffer)
    DevBuf<float> phase;          // [n_pad][n_wpad] (rows past n: padding)
    DevBuf<uint32_t> comp;        // [n_pad / 2][n_wpad]: Kahan compensation of spins 2m, 2m + 1
                                  // as bf16 in the low and high half
    DevBuf<double> partial, out;
    double walk_ms = 0.0, gemm_ms = 0.0;
    ~dsim_waveforms() {
        g_hi.release(); g_lo.release(); gsum.release(); traj.release(); phase.release();
        comp.release();
        partial.release(); out.release();
    }
};

namespace {

constexpr uint32_t kPhaseBlocks = 512;  // spin blocks of the readout

uint16_t bf16_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_val(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

extern "C" int dsim_waveforms_create(dsim_sim *sim, uint32_t n_wave, uint32_t n_steps,
                                     const float *grads, float pc, uint32_t chunk,
                                     dsim_waveforms **out) {
    if (!sim || !grads || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (n_wave == 0) return fail(DSIM_ERR_ARG, "n_wave must be positive");
    if (chunk == 0 || chunk % dsim::kPhaseSliceK || n_steps == 0 || n_steps % dsim::kPhaseSliceK)
        return fail(DSIM_ERR_ARG, "chunk_steps and n_steps must be multiples of %u",
                    dsim::kPhaseSliceK);
    auto *W = new dsim_waveforms();
    W->sim = sim;
    W->n = dsim_host::sim_spins(sim);
    W->n_pad = (W->n + dsim::kPhaseTileM - 1) / dsim::kPhaseTileM * dsim::kPhaseTileM;
    W->n_wave = n_wave;
    // the GEMM's waveform tiles take 128 waveforms (256 where they divide into
    // those); zero waveforms fill the last tile and are never read out
    const uint32_t nwp = (n_wave + dsim::kPhaseTileN - 1) / dsim::kPhaseTileN * dsim::kPhaseTileN;
    W->n_wpad = nwp;
    W->n_steps = n_steps;
    W->chunk = chunk;
    W->buffer = chunk;
    const size_t K = (size_t)3 * n_steps;
    std::vector<uint16_t> hi(nwp * K, 0), lo(nwp * K, 0);
    std::vector<double> gs((size_t)nwp * 3, 0.0);
    for (uint32_t t = 0; t < n_steps; ++t)
        for (uint32_t w = 0; w < n_wave; ++w)
            for (uint32_t a = 0; a < 3; ++a) {
                const float v = pc * grads[((size_t)t * n_wave + w) * 3 + a];
                const uint16_t h = bf16_rne(v);
                hi[w * K + 3 * t + a] = h;
                lo[w * K + 3 * t + a] = bf16_rne(v - bf16_val(h));
                gs[(size_t)w * 3 + a] += (double)v;
            }
    int rc;
    const size_t nw = (size_t)std::max<uint32_t>(W->n_pad, 1) * nwp;
    if ((rc = upload(W->g_hi, hi)) || (rc = upload(W->g_lo, lo)) || (rc = upload(W->gsum, gs)) ||
        (rc = W->traj.alloc((size_t)3 * chunk * W->n_pad)) || (rc = W->phase.alloc(nw)) ||
        (rc = W->comp.alloc((size_t)std::max<uint32_t>(W->n_pad / 2, 1) * nwp)) ||
        (rc = W->partial.alloc((size_t)kPhaseBlocks * nwp * 15)) ||
        (rc = W->out.alloc((size_t)nwp * 15))) {
        delete W;
        return rc;
    }
    // padded trajectory columns stay zero; phases start at zero
    hipError_t e = hipMemset(W->traj.p, 0, W->traj.n * 4);
    if (e == hipSuccess) e = hipMemset(W->phase.p, 0, nw * 4);
    if (e == hipSuccess) e = hipMemset(W->comp.p, 0, W->comp.n * 4);
    if (e != hipSuccess) {
        delete W;
        return fail(DSIM_ERR_HIP, "waveform init: %s", hipGetErrorString(e));
    }
    *out = W;
    return DSIM_OK;
}

extern "C" int dsim_waveforms_set_buffer(dsim_waveforms *W, uint32_t buffer_steps) {
    if (!W) return fail(DSIM_ERR_ARG, "NULL waveforms");
    if (buffer_steps == 0 || buffer_steps % W->chunk)
        return fail(DSIM_ERR_ARG, "buffer_steps must be a positive multiple of chunk_steps (%u)",
                    W->chunk);
    if (buffer_steps == W->buffer) return DSIM_OK;
    HIP_TRY(hipStreamSynchronize(dsim_host::sim_stream(W->sim)));
    W->traj.release();
    if (int rc = W->traj.alloc((size_t)3 * buffer_steps * W->n_pad)) return rc;
    // padded trajectory columns stay zero
    HIP_TRY(hipMemset(W->traj.p, 0, W->traj.n * 4));
    W->buffer = buffer_steps;
    return DSIM_OK;
}

extern "C" int dsim_waveforms_destroy(dsim_waveforms *W) {
    if (!W) return DSIM_OK;
    (void)hipStreamSynchronize(dsim_host::sim_stream(W->sim));
    delete W;
    return DSIM_OK;
}

extern "C" int dsim_waveforms_run(dsim_waveforms *W, uint32_t steps, float *traj_out) {
    if (!W) return fail(DSIM_ERR_ARG, "NULL waveforms");
    if (steps % dsim::kPhaseSliceK || W->done + steps > W->n_steps)
        return fail(DSIM_ERR_ARG, "steps must be a multiple of %u within the waveforms' %u",
                    dsim::kPhaseSliceK, W->n_steps);
    hipStream_t st = dsim_host::sim_stream(W->sim);
    hipEvent_t ev[3];
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    int rc = DSIM_OK;
    for (uint32_t c = 0; c < steps && rc == DSIM_OK;) {
        const uint32_t len = std::min(W->buffer, steps - c);  // a multiple of 32
        HIP_TRY(hipEventRecord(ev[0], st));
        if ((rc = dsim_host::sim_walk_traj(W->sim, len, W->traj.p))) break;
        HIP_TRY(hipEventRecord(ev[1], st));
        hipError_t e = dsim::launch_phase_gemm(W->traj.p, W->n, W->n_pad, 3 * len, 3 * W->chunk,
                                               W->g_hi.p, W->g_lo.p, 3 * W->n_steps,
                                               3 * W->done, W->n_wpad, W->phase.p, W->comp.p,
                                               st);
        if (e != hipSuccess) {
            rc = fail(DSIM_ERR_HIP, "phase GEMM: %s", hipGetErrorString(e));
            break;
        }
        HIP_TRY(hipEventRecord(ev[2], st));
        if (traj_out) {
            // rows [3 * len][n] out of the spin tiles' 16-byte k chunks
            const size_t tile = (size_t)3 * len * dsim::kTrajTile;  // floats per tile
            std::vector<float> h(tile * (W->n_pad / dsim::kTrajTile));
            HIP_TRY(hipMemcpyAsync(h.data(), W->traj.p, h.size() * 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (uint32_t r = 0; r < 3 * len; ++r) {
                float *dst = traj_out + ((size_t)3 * c + r) * W->n;
                const size_t kc = (size_t)(r / dsim::kTrajChunk) * dsim::kTrajTile * dsim::kTrajChunk +
                                  r % dsim::kTrajChunk;
                for (uint32_t s = 0; s < W->n; ++s)
                    dst[s] = h[(s / dsim::kTrajTile) * tile + kc +
                               (size_t)(s % dsim::kTrajTile) * dsim::kTrajChunk];
            }
        }
        HIP_TRY(hipEventSynchronize(ev[2]));
        float a = 0.f, b = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[1], ev[2]));
        W->walk_ms += a;
        W->gemm_ms += b;
        W->done += len;
        c += len;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    return dsim_sync(W->sim);
}

extern "C" int dsim_waveforms_signal(dsim_waveforms *W, double *out) {
    if (!W || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    const float *mag, *x0;
    const uint32_t *label, *id;
    dsim_host::sim_readout_fields(W->sim, &mag, &label, &x0, &id);
    hipStream_t st = dsim_host::sim_stream(W->sim);
    const uint32_t nb = std::max<uint32_t>(1, std::min<uint32_t>(kPhaseBlocks, W->n));
    hipError_t e = dsim::launch_phase_signal(W->n, W->phase.p, W->n_wpad, W->gsum.p, mag, label, x0,
                                             id, W->partial.p, nb, W->out.p, st);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "phase readout: %s", hipGetErrorString(e));
    std::vector<double> v((size_t)W->n_wpad * 15);
    HIP_TRY(hipMemcpyAsync(v.data(), W->out.p, v.size() * 8, hipMemcpyDeviceToHost, st));
    if (int rc = dsim_sync(W->sim)) return rc;
    for (uint32_t w = 0; w < W->n_wave; ++w) {
        const double *q = &v[(size_t)w * 15];
        for (int k = 0; k < 5; ++k)
            out[(size_t)w * 5 + k] =
                q[10 + k] == 0.0 ? 0.0 : std::sqrt(q[k] * q[k] + q[5 + k] * q[5 + k]) / q[10 + k];
    }
    return DSIM_OK;
}

extern "C" int dsim_waveforms_get_phase(dsim_waveforms *W, float *phase, float *comp) {
    if (!W) return fail(DSIM_ERR_ARG, "NULL waveforms");
    if (int rc = dsim_sync(W->sim)) return rc;
    const size_t m = (size_t)W->n * W->n_wave;
    const uint32_t nwp = W->n_wpad;  // row stride of the device arrays
    if (phase && m)
        HIP_TRY(hipMemcpy2D(phase, (size_t)W->n_wave * 4, W->phase.p, (size_t)nwp * 4,
                            (size_t)W->n_wave * 4, W->n, hipMemcpyDeviceToHost));
    if (comp && m) {
        std::vector<uint32_t> v((size_t)(W->n + 1) / 2 * nwp);
        HIP_TRY(hipMemcpy(v.data(), W->comp.p, v.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < W->n; ++s)
            for (uint32_t w = 0; w < W->n_wave; ++w) {
                const uint32_t word = v[(size_t)(s / 2) * nwp + w];
                const uint32_t bits = (s & 1u) ? (word & 0xffff0000u) : (word << 16);
                std::memcpy(&comp[(size_t)s * W->n_wave + w], &bits, 4);
            }
    }
    return DSIM_OK;
}

extern "C" int dsim_waveforms_timing(dsim_waveforms *W, double *walk_ms, double *gemm_ms) {
    if (!W) return fail(DSIM_ERR_ARG, "NULL waveforms");
    if (walk_ms) *walk_ms = W->walk_ms;
    if (gemm_ms) *gemm_ms = W->gemm_ms;
    return DSIM_OK;
}
