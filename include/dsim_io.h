/*
 * dsim I/O: the reference's input and output formats (C ABI).
 *
 * Replaces the host-program half of the reference (diffusion.cpp): readInputs and the gradient
 * reader (dSim3/src/diffusion.cpp:124-269), the mesh reader of
 * SpinSystem::initFibers (dSim3/src/spinSystem.cpp:290-610), the PGSE
 * state machine of update() (diffusion.cpp:564-624), computeTensor
 * (diffusion.cpp:286-511) and the MATLAB .m writer (diffusion.cpp:662-728).
 * Status codes as in dsim.h.
 */
#ifndef DSIM_IO_H
#define DSIM_IO_H

#include <stdint.h>

#include "dsim.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DSIM_PATH_MAX 1024

/* Message for the last failed call of this header on the calling thread. */
const char *dsim_io_last_error(void);

/* Settings read from sim.cfg (libconfig groups app / sim / fibers) with the
 * defaults of diffusion.cpp:89-107. */
typedef struct {
    int32_t stepsPerUpdate;
    int32_t useGpu;
    int32_t useDisplay;
    int32_t writeOut;
    uint32_t numSpins;
    float gyroMagneticRatio; /* kHz/T */
    float timeStep;          /* ms */
    float extraAdc, intraAdc, myelinAdc; /* um^2/ms */
    float extraT2, intraT2, myelinT2;    /* ms */
    float spaceScale;        /* um */
    float startBox;
    float permeability;
    uint32_t searchMethod;
    char outFile[DSIM_PATH_MAX];
    char gradsFile[DSIM_PATH_MAX];
    char fiberFile[DSIM_PATH_MAX];
    char configFile[DSIM_PATH_MAX];
    uint64_t seed;           /* not in the reference (srand(time(0))); 0 = time-seeded */
} dsim_config;

/* Defaults of diffusion.cpp:89-107. */
int dsim_config_defaults(dsim_config *cfg);
/* readInputs' cfg lookups (diffusion.cpp:148-210).  Unknown keys are ignored
 * as libconfig lookups ignore them; dSim2 key names (sim.timestep,
 * sim.gyromagneticRatio, fibers.file) are accepted as aliases.  A missing or
 * unparsable file is reported (the reference prints and continues with the
 * defaults, diffusion.cpp:136-143); cfg keeps what was read. */
int dsim_config_read(const char *path, dsim_config *cfg);
/* Command-line overrides (diffusion.cpp:219-268, 1081): -config=, -disp,
 * -nodisp, -gpu, -cpu, -w, -nw, -stepsPerUpdate=, -numSpins=, -searchMethod=,
 * -perm=, -startBox=, -fiberFile=, and -seed= (build addition).  Arguments
 * are matched with one or two leading dashes. */
int dsim_config_apply_args(dsim_config *cfg, int argc, const char *const *argv);

/* Gradient file (diffusion.cpp:177-202): lines of six numbers
 * "delta Delta readout gx gy gz" (ms, mT/m).  rows receives up to cap lines
 * (6 floats each); *n is the number found. */
int dsim_grads_read(const char *path, float *rows, uint32_t cap, uint32_t *n);

/* Mesh file (V / T / A n / M n / G n sections, 0-based indices; the format
 * written by dSimTriangulateFibers.m:90-112).  Vertices are scaled by
 * 1/space_scale.  A `Number_of_... <n>` / `Max_triangles_on_membrane <n>`
 * header is skipped (the reference misreads it and rejects the file).  Fails
 * with DSIM_ERR_FORMAT if the axon and myelin membrane counts differ, as the
 * reference does (spinSystem.cpp:593-596). */
typedef struct dsim_mesh_file dsim_mesh_file;
typedef struct {
    uint32_t n_vertices, n_triangles;
    uint32_t n_axon, n_myelin, n_glia; /* membrane sections */
    uint32_t n_fibers;                 /* n_axon + n_glia (spinSystem.cpp:384) */
    uint32_t max_triangles_on_surface;
    float xmax, ymax, zmax;            /* max(1, |coordinate|) after scaling */
} dsim_mesh_info;
int dsim_mesh_read(const char *path, float space_scale, dsim_mesh_file **out);
/* Borrowed view of the arrays (valid until dsim_mesh_free). */
int dsim_mesh_view(const dsim_mesh_file *m, dsim_mesh *view, dsim_mesh_info *info);
int dsim_mesh_free(dsim_mesh_file *m);
/* Writes a mesh in the same format (coordinates multiplied by space_scale). */
int dsim_mesh_write(const char *path, const dsim_mesh *mesh, float space_scale);

/* The on/off gradient state machine of update() (diffusion.cpp:564-608) for
 * one gradient line: L, B and R in updates of steps_per_update steps, the
 * lobe sign of every update (sign[k] for k < L+B+R, cap entries), and whether
 * the readout happens at update L+B+R. */
int dsim_pgse_schedule(float ldelta, float bdelta, float readout, float time_step,
                       int32_t steps_per_update, int32_t *sign, uint32_t cap, int32_t *L,
                       int32_t *B, int32_t *R, int32_t *readout_happens);

/* computeTensor (diffusion.cpp:286-511): rows are the gradient lines (6
 * floats) and signal[j] their measured signal for one compartment; the
 * Stejskal-Tanner least-squares tensor through a rounded pseudo-inverse.
 * eigval ascending (JAMA order), eigvec columns match eigval. */
typedef struct {
    double tensor[3][3];
    double eigvec[3][3];
    float eigval[3];
    float md, fa, rd;
} dsim_tensor_fit;
int dsim_tensor_fit_compute(uint32_t n_grads, const float *rows, const float *signal,
                            dsim_tensor_fit *out);

/* Appends one run to a MATLAB .m file in the reference's layout
 * (diffusion.cpp:666-728).  signals is [n_grads][5] (total, compartments
 * 0..3), fits is [5] (total, compartments 0..3). */
int dsim_write_m(const char *path, const dsim_config *cfg, uint32_t n_grads, const float *rows,
                 const float *signals, const dsim_tensor_fit *fits);

#ifdef __cplusplus
}
#endif
#endif
