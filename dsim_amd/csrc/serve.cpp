// Server of the dsim C ABI (include/dsim_serve.h): resident substrates,
// first-come first-served admission within an HBM budget, and rounds in which
// one batched step launch (a replayed HIP graph over a device job table)
// advances every admitted request.
//
// Each admitted request is an ordinary simulation handle (sim.cpp): spins
// initialised and labelled on its own stream, then advanced by the batched
// kernel k_walk_batch, then read out with dsim_signal_moment.  The handle does
// the same arithmetic as a plain dsim_integrate_moment run, so a request's
// signals are the same alone, in any mix, and through the plain path.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/dsim.h"
#include "../../include/dsim_serve.h"
#include "dsim_internal.h"
#include "host_common.h"
#include "sim_internal.h"

using dsim_host::fail;

namespace {

// kActive covers admitted requests from preparation to readout; budget is
// held from admission until the handle is released after the round.
enum Status { kQueued = 0, kActive = 1, kDone = 2, kFailed = 3 };

struct Req {
    uint64_t ticket = 0;
    dsim_request r{};
    dsim_substrate *sub = nullptr;
    std::vector<float> lobe_w, grads;
    std::vector<uint32_t> lobe_n;
    uint64_t bytes = 0;
    dsim_sim *sim = nullptr;
    uint32_t lobe = 0, lobe_left = 0;
    uint32_t chunk = 0;  // steps of the current round
    bool ready = false;     // spins prepared: takes part in rounds
    bool read_out = false;  // lobes done: signals formed at the next round
    bool released = false;
    int status = kQueued;
    int err = DSIM_OK;
    std::string msg;
    std::vector<double> signals;
    dsim_request_stats st{};
};

// One launch kind of the batched step (boundary, and whether the jobs' grids
// keep cell-major records): its job table in pinned host memory and on the
// device, and the captured launch.
struct Batch {
    dsim::WalkTable *host = nullptr;
    dsim::WalkTable *dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // namespace

struct dsim_server {
    uint64_t budget = 0, used = 0;
    uint32_t round_steps = 256, max_spins = 0, active_spins = 0, grid_blocks = 0;
    uint32_t small_first = 0;  // dsim_server_set_policy
    std::vector<dsim_substrate *> subs;
    std::deque<Req *> queue;
    std::vector<Req *> active;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> all;
    uint64_t next_ticket = 1;
    size_t n_active = 0;  // active.size(), read by the loop under mu
    std::mutex mu;        // queue, all, statuses, n_active
    std::mutex round_mu;  // one round at a time
    std::condition_variable cv_work, cv_done;
    std::thread loop;
    bool running = false, stopping = false;
    int loop_err = DSIM_OK;
    hipStream_t stream = nullptr;
    Batch batch[4];  // [boundary * 2 + (grid entries read through cell_tri)]
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

size_t table_bytes(uint32_t n_jobs) {
    return offsetof(dsim::WalkTable, job) + (size_t)n_jobs * sizeof(dsim::WalkJob);
}

void finish(dsim_server *S, Req *r, int status, int err, const std::string &msg) {
    std::lock_guard<std::mutex> l(S->mu);
    r->status = status;
    r->err = err;
    r->msg = msg;
    r->st.done_s = S->clock();
    S->cv_done.notify_all();
}

// Creates the request's handle and prepares its spins (on its own stream).
int start_request(dsim_server *S, Req *r) {
    dsim_params p{};
    p.dt = r->r.dt;
    for (int c = 0; c < 4; ++c) { p.std_dev[c] = r->r.std_dev[c]; p.t2[c] = r->r.t2[c]; }
    p.permeability = r->r.permeability;
    p.reflection_type = 1;
    p.mode = DSIM_MODE_PHYSICAL;
    p.boundary = r->r.boundary;
    p.rng = DSIM_RNG_PHILOX;
    p.seed = r->r.seed;
    p.spin_radius = 0.005f;
    p.sort_interval = 0;
    p.steps_per_launch = 0;
    bool voxel = false;
    for (int k = 0; k < 3; ++k) {
        p.voxel_lo[k] = r->r.voxel_lo[k];
        p.voxel_hi[k] = r->r.voxel_hi[k];
        voxel |= p.voxel_lo[k] != 0.0f || p.voxel_hi[k] != 0.0f;
    }
    int rc;
    if ((rc = dsim_sim_create(r->sub, &p, r->r.n_spins, &r->sim))) return rc;
    uint32_t amb = 0;
    if ((rc = voxel ? dsim_init_spins_box(r->sim, p.voxel_lo, p.voxel_hi)
                    : dsim_init_spins_uniform(r->sim, r->r.start_box)) ||
        (rc = dsim_label_spins(r->sim, r->r.label_extent[0], r->r.label_extent[1],
                               r->r.label_extent[2], 20, &amb)) ||
        (rc = dsim_moment_reset(r->sim)) || (rc = dsim_sync(r->sim)))
        return rc;
    r->st.ambiguous = amb;
    r->lobe = 0;
    r->lobe_left = r->lobe_n.empty() ? 0 : r->lobe_n[0];
    while (r->lobe_left == 0 && r->lobe + 1 < r->lobe_n.size()) r->lobe_left = r->lobe_n[++r->lobe];
    return DSIM_OK;
}

// Frees a finished request's handle and returns its budget (called with the
// device idle for the server stream, so the frees do not stall a round).
void release(dsim_server *S, Req *r) {
    if (r->released) return;
    if (r->sim) {
        dsim_sim_destroy(r->sim);
        r->sim = nullptr;
    }
    std::lock_guard<std::mutex> l(S->mu);
    S->used -= r->bytes;
    S->active_spins -= r->r.n_spins;
    r->released = true;
}

// Workgroups of the step kernel resident at once on the device (256 CUs, four
// one-wave workgroups per SIMD at its register use).
constexpr uint32_t kResidentBlocks = 4096;

uint32_t walk_blocks(uint32_t n_spins) {
    return (n_spins + dsim::kWalkBlockSpins - 1) / dsim::kWalkBlockSpins;
}

int ensure_batch(dsim_server *S, int kind) {
    Batch &B = S->batch[kind];
    if (B.exec) return DSIM_OK;
    HIP_TRY(hipHostMalloc((void **)&B.host, sizeof(dsim::WalkTable), hipHostMallocDefault));
    HIP_TRY(hipMalloc((void **)&B.dev, sizeof(dsim::WalkTable)));
    // the launch shape never changes: blocks past the table's total exit at
    // once, so one captured launch serves every round
    HIP_TRY(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = dsim::launch_walk_batch(
        B.dev, S->grid_blocks, (kind >> 1) == DSIM_BOUNDARY_PERIODIC,
        (kind & 1) ? dsim::kRecsIndirect : dsim::kRecsCellMajor, S->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(S->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(DSIM_ERR_HIP, "capture of the batched step: %s",
                    hipGetErrorString(e != hipSuccess ? e : e2));
    B.graph = g;
    HIP_TRY(hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0));
    return DSIM_OK;
}

// One round: (1) launch the batched step for every prepared request; (2)
// while it runs, form the signals of requests whose lobes ended last round
// and prepare newly admitted requests (their handles' own streams); (3) wait
// for the step; (4) advance the bookkeeping; (5) release finished requests.
int run_round(dsim_server *S, uint32_t *n_active, uint32_t *n_queued) {
    std::lock_guard<std::mutex> rl(S->round_mu);
    // (1) one batched launch per kind (boundary, record layout of the grid)
    // carries every prepared request of that kind.
    // A round lasts about total_blocks / kResidentBlocks times as long as one
    // block's round_steps steps, so a job of few blocks can walk that many
    // more steps in it without making it longer: small jobs get longer chunks
    // (fewer rounds, lower latency) and go first in the table, so their
    // blocks are dispatched first.
    bool launched = false;
    // small-first policy: a round with a small prepared request runs only the
    // small ones
    uint32_t cap = UINT32_MAX;
    if (S->small_first)
        for (Req *r : S->active)
            if (r->ready && !r->read_out && r->lobe_left && walk_blocks(r->r.n_spins) <= S->small_first)
                cap = S->small_first;
    for (int kind = 0; kind < 4; ++kind) {
        std::vector<Req *> jobs;
        uint32_t total = 0;
        for (Req *r : S->active) {
            if (!r->ready || r->read_out || r->r.boundary != (uint32_t)(kind >> 1) ||
                r->lobe_left == 0 || (r->sub->cell_mt.p == nullptr) != ((kind & 1) != 0) ||
                walk_blocks(r->r.n_spins) > cap)
                continue;
            jobs.push_back(r);
            total += walk_blocks(r->r.n_spins);
        }
        std::stable_sort(jobs.begin(), jobs.end(),
                         [](const Req *a, const Req *b) { return a->r.n_spins < b->r.n_spins; });
        const uint32_t stretch = std::max<uint32_t>(1, total / kResidentBlocks);
        uint32_t nj = 0, blocks = 0;
        dsim::WalkTable *T = nullptr;
        for (Req *r : jobs) {
            if (!T) {
                if (int rc = ensure_batch(S, kind)) return rc;
                T = S->batch[kind].host;
            }
            const uint32_t nb = walk_blocks(r->r.n_spins);
            const uint32_t want = nb * 8 <= total ? S->round_steps * stretch : S->round_steps;
            r->chunk = std::min(want, r->lobe_left);
            dsim::WalkJob &J = T->job[nj];
            if (int rc = dsim_host::sim_batch_params(r->sim, r->lobe_w[r->lobe], &J.p)) return rc;
            J.iters = r->chunk;
            T->block_start[nj] = blocks;
            blocks += nb;
            nj++;
        }
        if (!nj) continue;
        T->block_start[nj] = blocks;
        T->n_jobs = nj;
        T->total_blocks = blocks;
        HIP_TRY(hipMemcpyAsync(S->batch[kind].dev, T, table_bytes(nj), hipMemcpyHostToDevice,
                               S->stream));
        HIP_TRY(hipGraphLaunch(S->batch[kind].exec, S->stream));
        launched = true;
    }
    // (2a) signals of the requests whose lobes ended last round
    for (Req *r : S->active) {
        if (!r->read_out || r->status != kActive) continue;
        r->signals.assign((size_t)r->r.n_lines * 5, 0.0);
        int rc = r->r.n_lines ? dsim_signal_moment(r->sim, r->r.n_lines, r->grads.data(),
                                                   r->r.phase_constant, r->signals.data())
                              : dsim_sync(r->sim);
        if (rc) finish(S, r, kFailed, rc, dsim_last_error());
        else finish(S, r, kDone, DSIM_OK, "");
    }
    // (2b) admission: first come, first served, while budget, spin limit and
    // job table room allow; spins prepared while the step runs
    std::vector<Req *> admitted;
    {
        std::lock_guard<std::mutex> l(S->mu);
        uint32_t per_kind[2] = {0, 0};
        for (Req *r : S->active) per_kind[r->r.boundary]++;
        while (!S->queue.empty()) {
            Req *r = S->queue.front();
            if (S->used + r->bytes > S->budget || S->active_spins + r->r.n_spins > S->max_spins ||
                per_kind[r->r.boundary] >= dsim::kMaxJobs)
                break;
            S->queue.pop_front();
            S->used += r->bytes;
            S->active_spins += r->r.n_spins;
            per_kind[r->r.boundary]++;
            r->status = kActive;
            admitted.push_back(r);
        }
    }
    for (Req *r : admitted) {
        r->st.admit_s = S->clock();
        S->active.push_back(r);
        if (int rc = start_request(S, r)) finish(S, r, kFailed, rc, dsim_last_error());
        else r->ready = true;
    }
    // (3) the step
    if (launched) HIP_TRY(hipStreamSynchronize(S->stream));
    // (4) bookkeeping; (5) release what is finished
    std::vector<Req *> still;
    for (Req *r : S->active) {
        if (r->chunk) {
            dsim_host::sim_batch_advance(r->sim, r->lobe_w[r->lobe], r->chunk);
            r->st.rounds++;
            r->st.spin_steps += (uint64_t)r->r.n_spins * r->chunk;
            r->lobe_left -= r->chunk;
            r->chunk = 0;
            while (r->lobe_left == 0 && r->lobe + 1 < r->lobe_n.size())
                r->lobe_left = r->lobe_n[++r->lobe];
        }
        if (r->ready && r->lobe_left == 0) r->read_out = true;
        if (r->status == kDone || r->status == kFailed) release(S, r);
        else still.push_back(r);
    }
    S->active.swap(still);
    std::lock_guard<std::mutex> l(S->mu);
    S->n_active = S->active.size();
    if (n_active) *n_active = (uint32_t)S->active.size();
    if (n_queued) *n_queued = (uint32_t)S->queue.size();
    return DSIM_OK;
}

void loop_main(dsim_server *S) {
    for (;;) {
        {
            std::unique_lock<std::mutex> l(S->mu);
            S->cv_work.wait(l, [S] { return S->stopping || !S->queue.empty() || S->n_active; });
            if (S->stopping) return;
        }
        int rc = run_round(S, nullptr, nullptr);
        if (rc) {
            // a failed launch takes every active request with it
            std::string msg = dsim_last_error();
            std::lock_guard<std::mutex> rl(S->round_mu);
            (void)hipStreamSynchronize(S->stream);
            for (Req *r : S->active) {
                finish(S, r, kFailed, rc, msg);
                release(S, r);
            }
            S->active.clear();
            std::lock_guard<std::mutex> l(S->mu);
            S->n_active = 0;
            S->loop_err = rc;
        }
    }
}

}  // namespace

extern "C" int dsim_server_create(uint64_t hbm_budget, uint32_t round_steps,
                                  uint32_t max_active_spins, dsim_server **out) {
    if (!out) return fail(DSIM_ERR_ARG, "NULL argument");
    *out = nullptr;
    int ndev = 0;
    if (int rc = dsim_device_count(&ndev)) return rc;
    if (ndev == 0) return fail(DSIM_ERR_NODEV, "no HIP device visible");
    if (hbm_budget == 0) return fail(DSIM_ERR_ARG, "zero HBM budget");
    auto *S = new dsim_server();
    S->budget = hbm_budget;
    S->round_steps = round_steps ? round_steps : 256;
    S->max_spins = max_active_spins ? max_active_spins : (64u << 20);
    S->grid_blocks = walk_blocks(S->max_spins) + dsim::kMaxJobs;
    hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete S;
        return fail(DSIM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = S;
    return DSIM_OK;
}

extern "C" int dsim_server_stop(dsim_server *S) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    {
        std::lock_guard<std::mutex> l(S->mu);
        if (!S->running) return DSIM_OK;
        S->stopping = true;
    }
    S->cv_work.notify_all();
    S->loop.join();
    std::lock_guard<std::mutex> l(S->mu);
    S->running = false;
    S->stopping = false;
    return DSIM_OK;
}

extern "C" int dsim_server_destroy(dsim_server *S) {
    if (!S) return DSIM_OK;
    dsim_server_stop(S);
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    for (auto &kv : S->all)
        if (kv.second->sim) dsim_sim_destroy(kv.second->sim);
    for (Batch &B : S->batch) {
        if (B.exec) (void)hipGraphExecDestroy(B.exec);
        if (B.graph) (void)hipGraphDestroy(B.graph);
        if (B.dev) (void)hipFree(B.dev);
        if (B.host) (void)hipHostFree(B.host);
    }
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return DSIM_OK;
}

extern "C" int dsim_server_add_substrate(dsim_server *S, dsim_substrate *sub, uint32_t *id) {
    if (!S || !sub || !id) return fail(DSIM_ERR_ARG, "NULL argument");
    if (!sub->on_device) return fail(DSIM_ERR_NODEV, "substrate was created without a HIP device");
    std::lock_guard<std::mutex> l(S->mu);
    S->subs.push_back(sub);
    *id = (uint32_t)(S->subs.size() - 1);
    return DSIM_OK;
}

extern "C" int dsim_server_submit(dsim_server *S, const dsim_request *q, uint64_t *ticket) {
    if (!S || !q || !ticket) return fail(DSIM_ERR_ARG, "NULL argument");
    if (q->boundary > 1) return fail(DSIM_ERR_ARG, "bad boundary");
    if (q->n_lobes && (!q->lobe_weight || !q->lobe_steps)) return fail(DSIM_ERR_ARG, "NULL lobes");
    if (q->n_lines && !q->grads) return fail(DSIM_ERR_ARG, "NULL grads");
    auto r = std::make_unique<Req>();
    r->r = *q;
    r->lobe_w.assign(q->lobe_weight, q->lobe_weight + q->n_lobes);
    r->lobe_n.assign(q->lobe_steps, q->lobe_steps + q->n_lobes);
    r->grads.assign(q->grads, q->grads + (size_t)q->n_lines * 3);
    r->r.lobe_weight = nullptr;
    r->r.lobe_steps = nullptr;
    r->r.grads = nullptr;
    dsim_params p{};
    p.mode = DSIM_MODE_PHYSICAL;
    p.rng = DSIM_RNG_PHILOX;
    r->bytes = dsim_host::sim_footprint(q->n_spins, &p, q->n_lines);
    std::lock_guard<std::mutex> l(S->mu);
    if (q->substrate >= S->subs.size()) return fail(DSIM_ERR_ARG, "unknown substrate %u", q->substrate);
    if (r->bytes > S->budget || q->n_spins > S->max_spins)
        return fail(DSIM_ERR_ARG, "request of %u spins exceeds the server's budget", q->n_spins);
    r->sub = S->subs[q->substrate];
    r->ticket = S->next_ticket++;
    r->st.submit_s = S->clock();
    *ticket = r->ticket;
    S->queue.push_back(r.get());
    S->all.emplace(r->ticket, std::move(r));
    S->cv_work.notify_all();
    return DSIM_OK;
}

extern "C" int dsim_server_set_policy(dsim_server *S, uint32_t small_first_blocks) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> rl(S->round_mu);
    S->small_first = small_first_blocks;
    return DSIM_OK;
}

extern "C" int dsim_server_round(dsim_server *S, uint32_t *n_active, uint32_t *n_queued) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    {
        std::lock_guard<std::mutex> l(S->mu);
        if (S->running) return fail(DSIM_ERR_STATE, "the background loop is running");
    }
    return run_round(S, n_active, n_queued);
}

extern "C" int dsim_server_start(dsim_server *S) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> l(S->mu);
    if (S->running) return DSIM_OK;
    S->running = true;
    S->stopping = false;
    S->loop = std::thread(loop_main, S);
    return DSIM_OK;
}

extern "C" int dsim_server_wait(dsim_server *S, uint64_t ticket, double timeout_s) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> l(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return fail(DSIM_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Req *r = it->second.get();
    auto done = [r] { return r->status == kDone || r->status == kFailed; };
    if (timeout_s < 0) S->cv_done.wait(l, done);
    else if (!S->cv_done.wait_for(l, std::chrono::duration<double>(timeout_s), done))
        return fail(DSIM_ERR_STATE, "request %llu not finished", (unsigned long long)ticket);
    if (r->status == kFailed) return fail(r->err, "%s", r->msg.c_str());
    return DSIM_OK;
}

extern "C" int dsim_server_result(dsim_server *S, uint64_t ticket, double *signals,
                                  dsim_request_stats *stats) {
    if (!S) return fail(DSIM_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> l(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return fail(DSIM_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Req *r = it->second.get();
    if (r->status == kFailed) return fail(r->err, "%s", r->msg.c_str());
    if (r->status != kDone) return fail(DSIM_ERR_STATE, "request not finished");
    if (signals) std::copy(r->signals.begin(), r->signals.end(), signals);
    if (stats) *stats = r->st;
    return DSIM_OK;
}

extern "C" double dsim_server_clock(const dsim_server *S) { return S ? S->clock() : 0.0; }
