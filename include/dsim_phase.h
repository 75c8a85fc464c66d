/*
 * dsim waveforms: the phase of many arbitrary gradient waveforms from one walk
 * (C ABI).
 *
 * No single counterpart in the reference: its phase update takes one
 * gradient vector per integrate call (spinKernelCpu.cpp:490, spinKernel.cu:811)
 * and its sequencer plays only piecewise PGSE lines (diffusion.cpp:564-585),
 * re-walking the spins for each.  Here one walk of a simulation handle serves
 * every waveform: the unwrapped displacement of every spin is recorded chunk by
 * chunk and multiplied on the matrix cores by the waveforms' gradients, the
 * phase of each (spin, waveform) pair kept in FP32 with Kahan compensation
 * across chunks (hi/lo bf16 split products, FP32 accumulation; the compensation
 * held as bf16), and the signals are summed in FP64.
 *
 * The handle must use the physical mode (spin reordering is allowed, and the
 * phases follow the spin id, but a reordered wave's trajectory stores scatter
 * over the buffer: the walk runs several times slower), and
 * dsim_moment_reset must have been called: the phase
 * is pc * sum_t G_w(t) . (x0 + u(t)) with x0 the positions at that reset and u
 * the unwrapped displacement since.
 */
#ifndef DSIM_PHASE_H
#define DSIM_PHASE_H

#include <stdint.h>

#include "dsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsim_waveforms dsim_waveforms;

/* n_wave: any positive count (the device arrays are padded with zero
 * waveforms to a multiple of 128); n_steps and chunk_steps: multiples of 32.
 * chunk_steps is the Kahan interval: the products of each chunk (GEMM K =
 * 3 * chunk_steps) are summed in FP32 and then added to the phase with
 * compensation; the last chunk of a run may be shorter.  grads: host array
 * [n_steps][n_wave][3] of gradScaled values (include/dsim.h units),
 * phase_constant = 2*pi*gamma*dt.  Device memory: 6 bytes per (spin,
 * waveform) for the phase (FP32) and its compensation (bf16), spins and
 * waveforms padded to multiples of 128, plus 4 * n_wave * 3 * n_steps bytes of
 * bf16 gradients and one trajectory buffer (12 bytes per spin and buffered
 * step). */
int dsim_waveforms_create(dsim_sim *sim, uint32_t n_wave, uint32_t n_steps, const float *grads,
                          float phase_constant, uint32_t chunk_steps, dsim_waveforms **out);
int dsim_waveforms_destroy(dsim_waveforms *wf);
/* Steps walked into the trajectory buffer before each GEMM launch (a
 * multiple of chunk_steps; chunk_steps at creation).  One launch folds every
 * chunk of the buffer into phases held in registers, so the phase pairs are
 * read and written once per buffer rather than once per chunk; the results
 * do not depend on it. */
int dsim_waveforms_set_buffer(dsim_waveforms *wf, uint32_t buffer_steps);
/* Walks `steps` more steps (a multiple of 32; the total may not pass n_steps)
 * and accumulates the phases, one trajectory buffer at a time.  traj_out (may be NULL): host array
 * [steps][3][n_spins] that receives the recorded displacements (for checks). */
int dsim_waveforms_run(dsim_waveforms *wf, uint32_t steps, float *traj_out);
/* Signals per waveform, [n_wave*5]: total, then compartments 0-3, as
 * getMrSignal (spinSystem.cpp:1452-1500), from the FP64 sums. */
int dsim_waveforms_signal(dsim_waveforms *wf, double *out);
/* Host copy of the accumulated phases without the x0 term, [n_spins][n_wave],
 * and their Kahan compensations (values of bf16; either may be NULL). */
int dsim_waveforms_get_phase(dsim_waveforms *wf, float *phase, float *comp);
/* Milliseconds spent in the walk and in the phase GEMM since creation. */
int dsim_waveforms_timing(dsim_waveforms *wf, double *walk_ms, double *gemm_ms);

#ifdef __cplusplus
}
#endif

#endif
