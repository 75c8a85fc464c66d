// Simulation handles of the dsim C ABI (include/dsim.h): device state,
// host transfers in the reference's AoS layout, step dispatch with the
// spatial reordering of spins, readout, labelling and timing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/dsim.h"
#include "dsim_internal.h"
#include "host_common.h"
#include "sim_internal.h"

using dsim_host::DevBuf;
using dsim_host::fail;

namespace {

// Per-spin state is one field-major block of 4-byte words, so reordering the
// spins is a single gather into a second block followed by a pointer swap.
enum Field {
    F_X, F_Y, F_Z, F_MAG, F_PHASE, F_LABEL, F_ID,
    F_DISP, F_MOM = F_DISP + 3, F_MOMC = F_MOM + 3, F_X0 = F_MOMC + 3,
    F_HITS = F_X0 + 3,  // membrane hits since the last reorder (orders the next one)
    F_WRAP,             // periodic voxel crossings (phase of the unwrapped position)
    F_COUNT
};

// Finished step-launch timings are folded into ms_done once this many event
// pairs are outstanding, so a long run of integrate calls holds a bounded
// number of events.
const size_t kMaxEventPairs = 64;

const uint32_t kDefaultStepsPerLaunch = 1024;
// Longest launch of a walk that records trajectories (sim_walk_traj).
const uint32_t kTrajLaunchSteps = 128;

}  // namespace

struct dsim_sim {
    dsim_substrate *sub = nullptr;
    dsim_params prm{};
    uint32_t n = 0;
    uint64_t step = 0;
    hipStream_t stream = nullptr;
    DevBuf<uint32_t> state, alt;  // F_COUNT * n words each (alt only when sorting)
    DevBuf<uint32_t> seed, flags, tries;
    DevBuf<double> partial, out;
    DevBuf<float> grads;
    DevBuf<float4> inject;
    DevBuf<uint32_t> trace;
    DevBuf<unsigned long long> stats;  // walk work counters while counting is on
    DevBuf<uint32_t> queue;            // slot counters of the persistent walk
    bool counting = false;
    DevBuf<uint32_t> keys_a, keys_b, vals_a, vals_b;
    DevBuf<uint8_t> sort_tmp;
    uint64_t steps_since_sort = ~0ull;
    bool permuted = false;
    bool moment_ready = false;
    float wsum = 0.0f;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    double ms_done = 0.0;
    uint64_t launches = 0, sorts = 0;

    uint32_t *f(int k) { return state.p + (size_t)k * n; }
    float *ff(int k) { return reinterpret_cast<float *>(f(k)); }
    bool sorting() const {
        return prm.sort_interval > 0 && prm.mode == DSIM_MODE_PHYSICAL && prm.rng == DSIM_RNG_PHILOX;
    }
    ~dsim_sim() {
        for (auto &e : ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (auto *b : {&state, &alt, &seed, &flags, &tries, &trace, &keys_a, &keys_b, &vals_a,
                        &vals_b})
            b->release();
        partial.release(); out.release(); grads.release(); inject.release(); sort_tmp.release();
        stats.release(); queue.release();
        if (stream) (void)hipStreamDestroy(stream);
        if (sub) sub->refs--;
    }
};

static int check_sim(dsim_sim *s) { return s ? DSIM_OK : fail(DSIM_ERR_ARG, "NULL sim"); }

static int upload_iota(dsim_sim *s) {
    std::vector<uint32_t> id(s->n);
    std::iota(id.begin(), id.end(), 0u);
    if (s->n) HIP_TRY(hipMemcpy(s->f(F_ID), id.data(), (size_t)s->n * 4, hipMemcpyHostToDevice));
    s->permuted = false;
    return DSIM_OK;
}

extern "C" int dsim_sim_create(dsim_substrate *sub, const dsim_params *p, uint32_t n,
                               dsim_sim **out) {
    if (!sub || !p || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (!sub->on_device) return fail(DSIM_ERR_NODEV, "substrate was created without a HIP device");
    if (p->mode > 1 || p->boundary > 1 || (p->rng != DSIM_RNG_MWC && p->rng != DSIM_RNG_PHILOX))
        return fail(DSIM_ERR_ARG, "bad mode/boundary/rng");
    bool voxel_set = false;
    for (int k = 0; k < 3; ++k) voxel_set |= p->voxel_lo[k] != 0.0f || p->voxel_hi[k] != 0.0f;
    if (voxel_set)
        for (int k = 0; k < 3; ++k)
            if (!(p->voxel_hi[k] > p->voxel_lo[k]) || !std::isfinite(p->voxel_lo[k]) ||
                !std::isfinite(p->voxel_hi[k]))
                return fail(DSIM_ERR_ARG, "voxel axis %d: [%g, %g) is empty", k, p->voxel_lo[k],
                            p->voxel_hi[k]);
    if (p->mode == DSIM_MODE_REFERENCE && p->boundary != DSIM_BOUNDARY_CLAMP)
        return fail(DSIM_ERR_ARG, "the reference mode clamps (spinKernelCpu.cpp:469-474)");
    if (!(p->dt > 0.0f) || !(p->permeability >= 0.0f && p->permeability <= 1.0f))
        return fail(DSIM_ERR_ARG, "dt must be > 0 and permeability in [0,1]");
    if ((uint64_t)n * F_COUNT * 4 > (64ull << 30)) return fail(DSIM_ERR_ARG, "too many spins");
    auto *S = new dsim_sim();
    S->sub = sub;
    sub->refs++;
    S->prm = *p;
    if (!voxel_set)
        for (int k = 0; k < 3; ++k) { S->prm.voxel_lo[k] = -1.0f; S->prm.voxel_hi[k] = 1.0f; }
    if (S->prm.steps_per_launch == 0) S->prm.steps_per_launch = kDefaultStepsPerLaunch;
    S->n = n;
    hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete S;
        return fail(DSIM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    const size_t m = std::max<uint32_t>(n, 1);
    int rc;
    if ((rc = S->state.alloc(m * F_COUNT)) || (rc = S->flags.alloc(4)) || (rc = S->queue.alloc(8)) ||
        (rc = S->partial.alloc((size_t)dsim::kSignalBlocks * 15)) || (rc = S->out.alloc(15)) ||
        (rc = S->tries.alloc(m)) ||
        (p->rng == DSIM_RNG_MWC && (rc = S->seed.alloc(2 * m)))) {
        delete S;
        return rc;
    }
    if (S->sorting()) {
        size_t tb = 0;
        e = dsim::sort_temp_bytes(n, &tb);
        if (e != hipSuccess) {
            delete S;
            return fail(DSIM_ERR_HIP, "sort sizing: %s", hipGetErrorString(e));
        }
        if ((rc = S->alt.alloc(m * F_COUNT)) || (rc = S->keys_a.alloc(m)) ||
            (rc = S->keys_b.alloc(m)) || (rc = S->vals_a.alloc(m)) || (rc = S->vals_b.alloc(m)) ||
            (rc = S->sort_tmp.alloc(std::max<size_t>(tb, 1)))) {
            delete S;
            return rc;
        }
    }
    e = hipMemsetAsync(S->state.p, 0, m * F_COUNT * 4, S->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->flags.p, 0, 16, S->stream);
    // MWC seeds: a draw keyed by the Philox key, so a fresh handle walks a
    // random stream (an all-zero MWC state repeats one draw forever); callers
    // with the reference's seeds replace them with dsim_set_state_aos
    if (e == hipSuccess && S->seed.p)
        e = dsim::launch_init_seeds(n, S->seed.p, (uint32_t)p->seed, (uint32_t)(p->seed >> 32),
                                    S->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    if (e != hipSuccess) {
        delete S;
        return fail(DSIM_ERR_HIP, "init: %s", hipGetErrorString(e));
    }
    if ((rc = upload_iota(S))) {
        delete S;
        return rc;
    }
    *out = S;
    return DSIM_OK;
}

extern "C" int dsim_sim_destroy(dsim_sim *s) {
    if (!s) return DSIM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s;
    return DSIM_OK;
}

static int sync_and_check(dsim_sim *s) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    uint32_t fl[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpy(fl, s->flags.p, 16, hipMemcpyDeviceToHost));
    if (fl[0] & 1u) {
        HIP_TRY(hipMemset(s->flags.p, 0, 16));
        return fail(DSIM_ERR_OVERFLOW, "a spin exceeded 4096 collisions in one step");
    }
    if (fl[0] & 4u) {
        HIP_TRY(hipMemset(s->flags.p, 0, 16));
        return fail(DSIM_ERR_STATE, "collision search numbered a grid entry out of range");
    }
    return DSIM_OK;
}

// Copies field k to the host in spin order (undoing the spatial reorder).
static int download_field(dsim_sim *s, int k, const std::vector<uint32_t> &id, uint32_t *dst) {
    const uint32_t n = s->n;
    if (!s->permuted) {
        HIP_TRY(hipMemcpy(dst, s->f(k), (size_t)n * 4, hipMemcpyDeviceToHost));
        return DSIM_OK;
    }
    std::vector<uint32_t> t(n);
    HIP_TRY(hipMemcpy(t.data(), s->f(k), (size_t)n * 4, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < n; ++j) dst[id[j]] = t[j];
    return DSIM_OK;
}

static int spin_ids(dsim_sim *s, std::vector<uint32_t> &id) {
    id.resize(s->n);
    if (s->permuted && s->n)
        HIP_TRY(hipMemcpy(id.data(), s->f(F_ID), (size_t)s->n * 4, hipMemcpyDeviceToHost));
    return DSIM_OK;
}

// Restores spin order on the device (slot == spin id).
static int unpermute(dsim_sim *s) {
    if (!s->permuted) return DSIM_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> id, all((size_t)s->n * F_COUNT);
    if (int rc = spin_ids(s, id)) return rc;
    for (int k = 0; k < F_COUNT; ++k)
        if (int rc = download_field(s, k, id, &all[(size_t)k * s->n])) return rc;
    HIP_TRY(hipMemcpy(s->state.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
    s->permuted = false;
    return DSIM_OK;
}

extern "C" int dsim_set_state_aos(dsim_sim *s, const float *pos, const dsim_spin_data *spins,
                                  const uint32_t *seeds) {
    if (int rc = check_sim(s)) return rc;
    const uint32_t n = s->n;
    if (n == 0) return DSIM_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));  // the copies below use the null stream
    if (int rc = unpermute(s)) return rc;
    if (pos) {
        std::vector<float> a(n), b(n), c(n);
        for (uint32_t i = 0; i < n; ++i) { a[i] = pos[3 * i]; b[i] = pos[3 * i + 1]; c[i] = pos[3 * i + 2]; }
        HIP_TRY(hipMemcpy(s->ff(F_X), a.data(), n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->ff(F_Y), b.data(), n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->ff(F_Z), c.data(), n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(s->f(F_WRAP), 0, (size_t)n * 4));  // new positions: no crossings yet
    }
    if (spins) {
        std::vector<float> m(n), ph(n);
        std::vector<uint32_t> lab(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (spins[i].compartmentType > 3)
                return fail(DSIM_ERR_ARG, "spin %u compartment %u > 3", i, spins[i].compartmentType);
            m[i] = spins[i].signalMagnitude;
            ph[i] = spins[i].signalPhase;
            lab[i] = (uint32_t)spins[i].compartmentType | ((uint32_t)spins[i].insideFiber << 8);
        }
        HIP_TRY(hipMemcpy(s->ff(F_MAG), m.data(), n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->ff(F_PHASE), ph.data(), n * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->f(F_LABEL), lab.data(), n * 4, hipMemcpyHostToDevice));
    }
    if (seeds) {
        if (!s->seed.p) return fail(DSIM_ERR_STATE, "seeds given but rng is not MWC");
        HIP_TRY(hipMemcpy(s->seed.p, seeds, (size_t)n * 8, hipMemcpyHostToDevice));
    }
    s->moment_ready = false;
    s->steps_since_sort = ~0ull;
    return DSIM_OK;
}

extern "C" int dsim_get_state_aos(dsim_sim *s, float *pos, dsim_spin_data *spins,
                                  uint32_t *seeds) {
    if (int rc = check_sim(s)) return rc;
    if (int rc = sync_and_check(s)) return rc;
    const uint32_t n = s->n;
    if (n == 0) return DSIM_OK;
    std::vector<uint32_t> id;
    if (int rc = spin_ids(s, id)) return rc;
    if (pos) {
        std::vector<float> a(n), b(n), c(n);
        int rc;
        if ((rc = download_field(s, F_X, id, (uint32_t *)a.data())) ||
            (rc = download_field(s, F_Y, id, (uint32_t *)b.data())) ||
            (rc = download_field(s, F_Z, id, (uint32_t *)c.data())))
            return rc;
        for (uint32_t i = 0; i < n; ++i) { pos[3 * i] = a[i]; pos[3 * i + 1] = b[i]; pos[3 * i + 2] = c[i]; }
    }
    if (spins) {
        std::vector<float> m(n), ph(n);
        std::vector<uint32_t> lab(n);
        int rc;
        if ((rc = download_field(s, F_MAG, id, (uint32_t *)m.data())) ||
            (rc = download_field(s, F_PHASE, id, (uint32_t *)ph.data())) ||
            (rc = download_field(s, F_LABEL, id, lab.data())))
            return rc;
        for (uint32_t i = 0; i < n; ++i) {
            memset(&spins[i], 0, sizeof(dsim_spin_data));
            spins[i].signalMagnitude = m[i];
            spins[i].signalPhase = ph[i];
            spins[i].compartmentType = (uint8_t)(lab[i] & 0xFFu);
            spins[i].insideFiber = (uint16_t)(lab[i] >> 8);
        }
    }
    if (seeds) {
        if (!s->seed.p) return fail(DSIM_ERR_STATE, "rng is not MWC");
        HIP_TRY(hipMemcpy(seeds, s->seed.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    }
    return DSIM_OK;
}

extern "C" int dsim_set_step(dsim_sim *s, uint64_t step) {
    if (int rc = check_sim(s)) return rc;
    s->step = step;
    return DSIM_OK;
}

extern "C" int dsim_get_step(dsim_sim *s, uint64_t *step) {
    if (int rc = check_sim(s)) return rc;
    if (!step) return fail(DSIM_ERR_ARG, "NULL step");
    *step = s->step;
    return DSIM_OK;
}

static dsim::StepParams step_params(dsim_sim *s) {
    dsim::StepParams P{};
    dsim_substrate *b = s->sub;
    P.tri_ref = b->tri_ref.p;
    P.tri_mt = b->tri_mt.p;
    P.cell_start = b->d_cell_start.p;
    P.cell_tri = b->d_cell_tri.p;
    P.cell_plane = b->cell_plane.p;
    P.cell_mt = b->cell_mt.p;
    P.n_entries = (uint32_t)b->d_cell_tri.n;
    P.cell_clear = s->prm.mode == DSIM_MODE_PHYSICAL ? b->d_clearance.p : nullptr;
    P.fiber_myelin = b->fiber_myelin.p;
    P.n_fibers = b->n_fibers;
    for (int k = 0; k < 3; ++k) {
        P.dims[k] = b->dims[k]; P.lo[k] = b->lo[k]; P.cell[k] = b->cell[k]; P.inv_cell[k] = b->inv_cell[k];
    }
    for (int c = 0; c < 4; ++c) {
        P.std_dev[c] = s->prm.std_dev[c];
        P.t2[c] = s->prm.t2[c];
        P.decay[c] = s->prm.dt / s->prm.t2[c];
    }
    P.dt = s->prm.dt;
    P.perm = s->prm.permeability;
    P.reflection_type = s->prm.reflection_type;
    P.clamp_hi = 1.0f - s->prm.spin_radius;
    P.clamp_lo = -1.0f + s->prm.spin_radius;
    for (int k = 0; k < 3; ++k) {
        P.wrap_lo[k] = s->prm.voxel_lo[k];
        P.wrap_hi[k] = s->prm.voxel_hi[k];
        P.wrap_len[k] = s->prm.voxel_hi[k] - s->prm.voxel_lo[k];
    }
    P.n = s->n;
    P.x = s->ff(F_X); P.y = s->ff(F_Y); P.z = s->ff(F_Z);
    P.mag = s->ff(F_MAG); P.phase = s->ff(F_PHASE);
    P.label = s->f(F_LABEL);
    P.spin_id = s->f(F_ID);
    P.seed = s->seed.p;
    P.disp = s->ff(F_DISP); P.mom = s->ff(F_MOM); P.momc = s->ff(F_MOMC);
    P.hits = s->sorting() ? s->f(F_HITS) : nullptr;
    P.wrap = s->f(F_WRAP);
    P.key0 = (uint32_t)s->prm.seed;
    P.key1 = (uint32_t)(s->prm.seed >> 32);
    P.flags = s->flags.p;
    P.queue = s->prm.fixed_blocks ? nullptr : s->queue.p;
    return P;
}

// Reorders the spins by their recent membrane-hit rate, then by the Morton
// index of their cell (the key is built on the device).
static int spatial_sort(dsim_sim *s) {
    dsim::StepParams P = step_params(s);
    // hit counts cover the steps since the last reorder (none before the first)
    P.hit_window = s->steps_since_sort < (1ull << 31) ? (uint32_t)s->steps_since_sort : 0u;
    const uint32_t *perm = nullptr;
    hipError_t e = dsim::launch_sort_perm(P, s->keys_a.p, s->keys_b.p, s->vals_a.p, s->vals_b.p,
                                          s->sort_tmp.p, s->sort_tmp.n, s->stream, &perm);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "sort: %s", hipGetErrorString(e));
    dsim::Fields fl{};
    fl.count = F_COUNT;
    for (int k = 0; k < F_COUNT; ++k) {
        fl.src[k] = s->state.p + (size_t)k * s->n;
        fl.dst[k] = s->alt.p + (size_t)k * s->n;
    }
    e = dsim::launch_gather(s->n, perm, fl, s->stream);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "gather: %s", hipGetErrorString(e));
    std::swap(s->state, s->alt);
    s->permuted = true;
    s->steps_since_sort = 0;
    s->sorts++;
    return DSIM_OK;
}

// Adds the elapsed times of the oldest `count` event pairs to ms_done and
// destroys them (waiting for the last of them to complete).
static int fold_events(dsim_sim *s, size_t count) {
    count = std::min(count, s->ev.size());
    if (count == 0) return DSIM_OK;
    HIP_TRY(hipEventSynchronize(s->ev[count - 1].second));
    for (size_t k = 0; k < count; ++k) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, s->ev[k].first, s->ev[k].second));
        s->ms_done += t;
        (void)hipEventDestroy(s->ev[k].first);
        (void)hipEventDestroy(s->ev[k].second);
    }
    s->ev.erase(s->ev.begin(), s->ev.begin() + (std::ptrdiff_t)count);
    return DSIM_OK;
}

static int run_steps(dsim_sim *s, dsim::StepParams P, uint32_t iters, int rng, bool moment,
                     const float *n3u, uint32_t *trace) {
    const uint32_t n = s->n;
    if (n3u) {
        if ((uint64_t)iters * n * 16 > (1ull << 32))
            return fail(DSIM_ERR_ARG, "injected sequence too large (%u x %u)", iters, n);
        if (s->inject.n < (size_t)iters * n) {
            s->inject.release();
            if (int rc = s->inject.alloc((size_t)iters * n)) return rc;
        }
        HIP_TRY(hipMemcpyAsync(s->inject.p, n3u, (size_t)iters * n * 16, hipMemcpyHostToDevice,
                               s->stream));
    }
    if (trace) {
        if ((uint64_t)iters * n * 20 > (1ull << 32))
            return fail(DSIM_ERR_ARG, "trace too large (%u x %u)", iters, n);
        if (s->trace.n < (size_t)iters * n * 5) {
            s->trace.release();
            if (int rc = s->trace.alloc((size_t)iters * n * 5)) return rc;
        }
    }
    int instr = trace ? dsim::kInstrTrace : dsim::kInstrNone;
    if (s->counting && !trace) {
        if (s->prm.mode != DSIM_MODE_PHYSICAL || rng == dsim::kRngMwc)
            return fail(DSIM_ERR_ARG,
                        "walk counters need the physical mode with Philox or injected draws");
        instr = dsim::kInstrStats;
    }
    uint32_t done = 0;
    const uint32_t chunk = s->sorting() ? std::min(s->prm.steps_per_launch, s->prm.sort_interval)
                                        : s->prm.steps_per_launch;
    float *const traj0 = P.traj;  // row 0 of the trajectory buffer's tiles
    while (done < iters) {
        if (s->sorting() && s->steps_since_sort >= s->prm.sort_interval) {
            if (int rc = spatial_sort(s)) return rc;
            P = [&] {
                dsim::StepParams Q = step_params(s);
                Q.gx = P.gx; Q.gy = P.gy; Q.gz = P.gz; Q.pc = P.pc; Q.weight = P.weight;
                Q.traj_stride = P.traj_stride;  // traj itself is set per launch below
                return Q;
            }();
        }
        uint32_t k = std::min(chunk, iters - done);
        P.step0 = s->step + done;
        P.inject = n3u ? s->inject.p + (size_t)done * n : nullptr;
        P.trace = trace ? s->trace.p + (size_t)done * n * 5 : nullptr;
        // a launch writes its steps' rows 3 * done .. of every tile
        P.traj = traj0 ? traj0 + (size_t)3 * done * dsim::kTrajTile : nullptr;
        P.stats = instr == dsim::kInstrStats ? s->stats.p : nullptr;
        // each step launch is bracketed by its own events: the timing covers
        // the step kernel only, not the reorders between launches
        if (s->ev.size() >= kMaxEventPairs)
            if (int rc = fold_events(s, s->ev.size() / 2)) return rc;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
        if (e != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            return fail(DSIM_ERR_HIP, "step events: %s", hipGetErrorString(e));
        }
        s->ev.emplace_back(e0, e1);  // owned by the handle from here on
        e = dsim::launch_step(P, k, (int)s->prm.mode, s->prm.boundary == DSIM_BOUNDARY_PERIODIC,
                              rng, instr, moment, s->stream);
        if (e != hipSuccess) return fail(DSIM_ERR_HIP, "step launch: %s", hipGetErrorString(e));
        HIP_TRY(hipEventRecord(e1, s->stream));
        done += k;
        s->steps_since_sort += k;
        s->launches++;
    }
    s->step += iters;
    if (trace) {
        HIP_TRY(hipMemcpyAsync(trace, s->trace.p, (size_t)iters * n * 20, hipMemcpyDeviceToHost,
                               s->stream));
        if (int rc = sync_and_check(s)) return rc;
    }
    return DSIM_OK;
}

static int rng_kind(dsim_sim *s, const float *n3u) {
    if (n3u) return dsim::kRngInject;
    return s->prm.rng == DSIM_RNG_MWC ? dsim::kRngMwc : dsim::kRngPhilox;
}

extern "C" int dsim_integrate(dsim_sim *s, const float g[3], float pc, uint32_t iters) {
    return dsim_integrate_injected(s, g, pc, nullptr, iters, nullptr);
}

extern "C" int dsim_integrate_injected(dsim_sim *s, const float g[3], float pc, const float *n3u,
                                       uint32_t iters, uint32_t *trace) {
    if (int rc = check_sim(s)) return rc;
    if (!g) return fail(DSIM_ERR_ARG, "NULL gradient");
    dsim::StepParams P = step_params(s);
    P.gx = g[0]; P.gy = g[1]; P.gz = g[2]; P.pc = pc;
    return run_steps(s, P, iters, rng_kind(s, n3u), false, n3u, trace);
}

extern "C" int dsim_signal(dsim_sim *s, double out[5]) {
    if (int rc = check_sim(s)) return rc;
    if (!out) return fail(DSIM_ERR_ARG, "NULL out");
    hipError_t e = dsim::launch_signal(s->n, s->ff(F_MAG), s->ff(F_PHASE), s->f(F_LABEL),
                                       s->partial.p, dsim::kSignalBlocks, s->out.p, s->stream);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "signal launch: %s", hipGetErrorString(e));
    double v[15];
    HIP_TRY(hipMemcpyAsync(v, s->out.p, sizeof v, hipMemcpyDeviceToHost, s->stream));
    if (int rc = sync_and_check(s)) return rc;
    for (int k = 0; k < 5; ++k)
        out[k] = v[10 + k] == 0.0 ? 0.0 : std::sqrt(v[k] * v[k] + v[5 + k] * v[5 + k]) / v[10 + k];
    return DSIM_OK;
}

extern "C" int dsim_moment_reset(dsim_sim *s) {
    if (int rc = check_sim(s)) return rc;
    hipError_t e = dsim::launch_moment_reset(s->n, s->ff(F_X), s->ff(F_Y), s->ff(F_Z), s->ff(F_X0),
                                             s->ff(F_DISP), s->ff(F_MOM), s->ff(F_MOMC),
                                             s->ff(F_MAG), s->stream);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "moment reset: %s", hipGetErrorString(e));
    s->wsum = 0.0f;
    s->moment_ready = true;
    return DSIM_OK;
}

extern "C" int dsim_integrate_moment(dsim_sim *s, float w, uint32_t iters) {
    return dsim_integrate_moment_injected(s, w, nullptr, iters, nullptr);
}

extern "C" int dsim_integrate_moment_injected(dsim_sim *s, float w, const float *n3u,
                                              uint32_t iters, uint32_t *trace) {
    if (int rc = check_sim(s)) return rc;
    if (!s->moment_ready) return fail(DSIM_ERR_STATE, "call dsim_moment_reset first");
    dsim::StepParams P = step_params(s);
    P.weight = w;
    int rc = run_steps(s, P, iters, rng_kind(s, n3u), true, n3u, trace);
    if (rc == DSIM_OK) s->wsum += w * (float)iters;
    return rc;
}

extern "C" int dsim_signal_moment(dsim_sim *s, uint32_t ndir, const float *g, float pc,
                                  double *out) {
    if (int rc = check_sim(s)) return rc;
    if (!s->moment_ready) return fail(DSIM_ERR_STATE, "no moments accumulated");
    if (ndir == 0) return DSIM_OK;
    if (!g || !out) return fail(DSIM_ERR_ARG, "NULL argument");
    const uint32_t nb = 256;
    if (s->grads.n < (size_t)ndir * 3) {
        s->grads.release();
        if (int rc = s->grads.alloc((size_t)ndir * 3)) return rc;
    }
    if (s->partial.n < (size_t)ndir * nb * 15) {
        s->partial.release();
        if (int rc = s->partial.alloc(
                (size_t)std::max<size_t>(ndir * (size_t)nb, dsim::kSignalBlocks) * 15))
            return rc;
    }
    DevBuf<double> o;
    if (int rc = o.alloc((size_t)ndir * 15)) return rc;
    HIP_TRY(hipMemcpyAsync(s->grads.p, g, (size_t)ndir * 12, hipMemcpyHostToDevice, s->stream));
    hipError_t e = dsim::launch_moment_signal(s->n, s->ff(F_MAG), s->ff(F_MOM), s->ff(F_X0),
                                              s->wsum, s->f(F_LABEL), ndir, s->grads.p, pc,
                                              s->partial.p, nb, o.p, s->stream);
    if (e != hipSuccess) {
        o.release();
        return fail(DSIM_ERR_HIP, "moment signal: %s", hipGetErrorString(e));
    }
    std::vector<double> v((size_t)ndir * 15);
    e = hipMemcpyAsync(v.data(), o.p, v.size() * 8, hipMemcpyDeviceToHost, s->stream);
    int rc = sync_and_check(s);
    o.release();
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "copy: %s", hipGetErrorString(e));
    if (rc) return rc;
    for (uint32_t j = 0; j < ndir; ++j) {
        const double *q = &v[(size_t)j * 15];
        for (int k = 0; k < 5; ++k)
            out[j * 5 + k] = q[10 + k] == 0.0 ? 0.0 : std::sqrt(q[k] * q[k] + q[5 + k] * q[5 + k]) / q[10 + k];
    }
    return DSIM_OK;
}

extern "C" int dsim_get_moment(dsim_sim *s, float *moment, float *x0) {
    if (int rc = check_sim(s)) return rc;
    if (!s->moment_ready) return fail(DSIM_ERR_STATE, "no moments accumulated");
    if (int rc = sync_and_check(s)) return rc;
    const uint32_t n = s->n;
    std::vector<uint32_t> id;
    if (int rc = spin_ids(s, id)) return rc;
    std::vector<float> t(n);
    for (int which = 0; which < 2; ++which) {
        float *dst = which == 0 ? moment : x0;
        if (!dst) continue;
        int base = which == 0 ? F_MOM : F_X0;
        for (int k = 0; k < 3; ++k) {
            if (int rc = download_field(s, base + k, id, (uint32_t *)t.data())) return rc;
            for (uint32_t i = 0; i < n; ++i) dst[3 * i + k] = t[i];
        }
    }
    return DSIM_OK;
}

extern "C" int dsim_init_spins_uniform(dsim_sim *s, float box) {
    const float lo[3] = {-box, -box, -box}, hi[3] = {box, box, box};
    return dsim_init_spins_box(s, lo, hi);
}

extern "C" int dsim_init_spins_box(dsim_sim *s, const float lo[3], const float hi[3]) {
    if (int rc = check_sim(s)) return rc;
    if (!lo || !hi) return fail(DSIM_ERR_ARG, "NULL box");
    for (int k = 0; k < 3; ++k)
        if (!(hi[k] >= lo[k]) || !std::isfinite(lo[k]) || !std::isfinite(hi[k]))
            return fail(DSIM_ERR_ARG, "box axis %d: [%g, %g]", k, lo[k], hi[k]);
    hipError_t e = dsim::launch_init_uniform(s->n, s->ff(F_X), s->ff(F_Y), s->ff(F_Z), s->ff(F_MAG),
                                             s->ff(F_PHASE), s->f(F_LABEL), s->f(F_ID),
                                             s->f(F_WRAP), lo, hi, (uint32_t)s->prm.seed,
                                             (uint32_t)(s->prm.seed >> 32), s->step, s->stream);
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "init launch: %s", hipGetErrorString(e));
    s->permuted = false;
    s->moment_ready = false;
    s->steps_since_sort = ~0ull;
    return DSIM_OK;
}

extern "C" int dsim_label_spins(dsim_sim *s, float xmax, float ymax, float zmax,
                                uint32_t max_tries, uint32_t *n_ambiguous) {
    if (int rc = check_sim(s)) return rc;
    dsim::LabelParams L{};
    L.tri_ref = s->sub->tri_ref.p;
    L.tri_mt = s->sub->tri_mt.p;
    L.use_mt = s->prm.mode == DSIM_MODE_PHYSICAL ? 1u : 0u;
    L.tri_aabb = s->sub->tri_aabb.p;
    L.n_tri = s->sub->n_tri;
    L.n = s->n;
    L.x = s->ff(F_X); L.y = s->ff(F_Y); L.z = s->ff(F_Z);
    L.label = s->f(F_LABEL);
    L.tries = s->tries.p;
    L.xmax = xmax; L.ymax = ymax; L.zmax = zmax;
    L.max_tries = max_tries;
    // the physical mode casts along the substrate's ray axis (+x unless the
    // mesh's long walls run along x), the reference mode along the spin's
    // dominant axis
    L.ray_axis = s->prm.mode == DSIM_MODE_PHYSICAL ? s->sub->label_axis + 1u : 0u;
    // axis rays over an AABB or exact-membership grid take the grid-row
    // labeller first; the rest, and every re-cast, go through the tiled pass
    // over all triangles
    const dsim_substrate *sub = s->sub;
    const bool rows = L.ray_axis && sub->kind != DSIM_GRID_REFERENCE && sub->n_cells > 0 &&
                      s->n > 0;
    dsim_host::DevBuf<uint32_t> work;  // retry queue, counters, outside list
    hipError_t e;
    if (rows) {
        const size_t q = 2 * (size_t)s->n;  // the queue's uint2 entries sit at the aligned start
        if (int rc = work.alloc(q + 2 + (size_t)sub->n_tri)) return rc;
        L.retry = reinterpret_cast<uint2 *>(work.p);
        L.n_outside = work.p + q;
        L.n_retry = work.p + q + 1;
        L.outside = work.p + q + 2;
        e = hipMemsetAsync(L.n_outside, 0, 8, s->stream);
        if (e != hipSuccess) {
            work.release();
            return fail(DSIM_ERR_HIP, "label: %s", hipGetErrorString(e));
        }
        L.cell_start = sub->d_cell_start.p;
        L.cell_tri = sub->d_cell_tri.p;
        L.exact_grid = sub->kind == DSIM_GRID_TRIBOX ? 1u : 0u;
        for (int k = 0; k < 3; ++k) {
            L.dims[k] = sub->dims[k];
            L.lo[k] = sub->lo[k];
            L.cell[k] = sub->cell[k];
            L.inv_cell[k] = sub->inv_cell[k];
            L.hi[k] = sub->lo[k] + (float)sub->dims[k] * sub->cell[k];
        }
        e = dsim::launch_label_row(L, s->stream);
    } else {
        e = dsim::launch_label(L, s->stream);
    }
    if (e != hipSuccess) {
        work.release();
        return fail(DSIM_ERR_HIP, "label launch: %s", hipGetErrorString(e));
    }
    e = hipStreamSynchronize(s->stream);
    work.release();
    if (e != hipSuccess) return fail(DSIM_ERR_HIP, "label: %s", hipGetErrorString(e));
    if (n_ambiguous) {
        std::vector<uint32_t> t(s->n);
        if (s->n) HIP_TRY(hipMemcpy(t.data(), s->tries.p, (size_t)s->n * 4, hipMemcpyDeviceToHost));
        uint32_t a = 0;
        for (uint32_t v : t) a += v > max_tries;
        *n_ambiguous = a;
    }
    s->moment_ready = false;
    return DSIM_OK;
}

extern "C" int dsim_sync(dsim_sim *s) {
    if (int rc = check_sim(s)) return rc;
    return sync_and_check(s);
}

extern "C" int dsim_step_timing(dsim_sim *s, double *ms, uint64_t *launches) {
    if (int rc = check_sim(s)) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (int rc = fold_events(s, s->ev.size())) return rc;
    if (ms) *ms = s->ms_done;
    if (launches) *launches = s->launches;
    return DSIM_OK;
}

extern "C" int dsim_step_timing_reset(dsim_sim *s) {
    if (int rc = check_sim(s)) return rc;
    double t;
    if (int rc = dsim_step_timing(s, &t, nullptr)) return rc;
    s->ms_done = 0.0;
    s->launches = 0;
    return DSIM_OK;
}

extern "C" int dsim_walk_stats_enable(dsim_sim *s, int enable) {
    if (int rc = check_sim(s)) return rc;
    if (enable) {
        if (!s->stats.p)
            if (int rc = s->stats.alloc(dsim::kWalkStats)) return rc;
        HIP_TRY(hipMemsetAsync(s->stats.p, 0, dsim::kWalkStats * sizeof(unsigned long long),
                               s->stream));
    }
    s->counting = enable != 0;
    return DSIM_OK;
}

extern "C" int dsim_walk_stats_read(dsim_sim *s, uint64_t counts[DSIM_WALK_STATS]) {
    if (int rc = check_sim(s)) return rc;
    if (!counts) return fail(DSIM_ERR_ARG, "NULL counts");
    static_assert(DSIM_WALK_STATS == dsim::kWalkStats, "counter count");
    if (!s->stats.p) {
        for (int k = 0; k < DSIM_WALK_STATS; ++k) counts[k] = 0;
        return DSIM_OK;
    }
    HIP_TRY(hipMemcpyAsync(counts, s->stats.p, dsim::kWalkStats * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DSIM_OK;
}

// ---------------------------------------------------------------------------
// Hooks for the batched server
// ---------------------------------------------------------------------------

int dsim_host::sim_batch_params(dsim_sim *s, float w, dsim::StepParams *P) {
    if (int rc = check_sim(s)) return rc;
    if (s->prm.mode != DSIM_MODE_PHYSICAL || s->prm.rng != DSIM_RNG_PHILOX || s->sorting())
        return fail(DSIM_ERR_STATE, "batched steps need the physical mode, Philox, no reordering");
    if (!s->moment_ready) return fail(DSIM_ERR_STATE, "call dsim_moment_reset first");
    *P = step_params(s);
    P->weight = w;
    P->step0 = s->step;  // the Philox counter continues where the handle stands
    P->inject = nullptr;
    P->trace = nullptr;
    return DSIM_OK;
}

void dsim_host::sim_batch_advance(dsim_sim *s, float w, uint32_t iters) {
    s->wsum += w * (float)iters;
    s->step += iters;
    s->launches++;
}

hipStream_t dsim_host::sim_stream(const dsim_sim *s) { return s->stream; }

uint64_t dsim_host::sim_footprint(uint32_t n, const dsim_params *p, uint32_t n_lines) {
    const uint64_t m = std::max<uint32_t>(n, 1);
    uint64_t b = m * F_COUNT * 4 + m * 4 + 16 + (uint64_t)dsim::kSignalBlocks * 15 * 8 + 15 * 8;
    if (p->rng == DSIM_RNG_MWC) b += m * 8;
    if (p->sort_interval > 0 && p->mode == DSIM_MODE_PHYSICAL && p->rng == DSIM_RNG_PHILOX)
        b += m * F_COUNT * 4 + m * 16 + m * 8;  // second block, keys, values, sort scratch (about)
    b += (uint64_t)n_lines * (12 + 256 * 15 * 8 + 15 * 8);  // readout
    return b;
}

int dsim_host::sim_walk_traj(dsim_sim *s, uint32_t iters, float *traj) {
    if (int rc = check_sim(s)) return rc;
    if (s->prm.mode != DSIM_MODE_PHYSICAL)
        return fail(DSIM_ERR_STATE, "waveform runs need the physical mode");
    if (!s->moment_ready) return fail(DSIM_ERR_STATE, "call dsim_moment_reset first");
    dsim::StepParams P = step_params(s);
    P.weight = 0.0f;
    P.traj = traj;
    P.traj_stride = 3 * iters * dsim::kTrajTile;
    // Launches of at most kTrajLaunchSteps steps, each writing its steps'
    // rows: the resident waves of one launch then stand within that many
    // steps of each other, so their trajectory stores stay in a narrow band
    // of every tile (one launch per 1024-step buffer walked 15 % slower, and
    // the gap grows with the buffer).
    const uint32_t spl = s->prm.steps_per_launch;
    s->prm.steps_per_launch = std::min(spl, kTrajLaunchSteps);
    const int rc = run_steps(s, P, iters, rng_kind(s, nullptr), true, nullptr, nullptr);
    s->prm.steps_per_launch = spl;
    return rc;
}

void dsim_host::sim_readout_fields(dsim_sim *s, const float **mag, const uint32_t **label,
                                   const float **x0, const uint32_t **id) {
    *id = s->f(F_ID);
    *mag = s->ff(F_MAG);
    *label = s->f(F_LABEL);
    *x0 = s->ff(F_X0);
}

uint32_t dsim_host::sim_spins(const dsim_sim *s) { return s->n; }
