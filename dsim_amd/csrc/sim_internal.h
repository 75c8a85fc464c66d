// Hooks of the simulation handle (sim.cpp) for the server (serve.cpp): the
// server advances many handles with one batched step launch and then has to
// do the bookkeeping a dsim_integrate_moment call would have done.
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>

#include "../../include/dsim.h"
#include "dsim_internal.h"

namespace dsim_host {

// Step parameters of `s` for one batched launch at gradient weight w.
// Requires the physical mode, the Philox stream, no spin reordering and
// accumulated moments (dsim_moment_reset called).
int sim_batch_params(dsim_sim *s, float w, dsim::StepParams *P);
// Records that a batched launch advanced `s` by `iters` steps at weight w.
void sim_batch_advance(dsim_sim *s, float w, uint32_t iters);
hipStream_t sim_stream(const dsim_sim *s);
// Advances `s` by `iters` steps (moment mode, weight 0: no moment change)
// writing the unwrapped displacement after every step to traj by spin id, in
// tiles of kTrajTile spins of 3 * iters rows each ([s / kTrajTile][3 * step +
// axis][s % kTrajTile]).  Requires dsim_moment_reset; spins may be reordered
// between launches (the rows follow the spin id, not the slot).
int sim_walk_traj(dsim_sim *s, uint32_t iters, float *traj);
// Device pointers of the state fields the waveform readout needs.
// (slot order; id: the spin id of each slot)
void sim_readout_fields(dsim_sim *s, const float **mag, const uint32_t **label, const float **x0,
                        const uint32_t **id);
uint32_t sim_spins(const dsim_sim *s);
// Device bytes a handle of n spins holds, readout of n_lines lines included.
uint64_t sim_footprint(uint32_t n, const dsim_params *p, uint32_t n_lines);

}  // namespace dsim_host
