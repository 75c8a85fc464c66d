/* Oracle driver for the AddressSanitizer / UndefinedBehaviorSanitizer run
 * (tests/sanitize/run.sh): parses the sample meshes and gradient file, builds
 * the reference grid and both AABB grids, labels spins and integrates in the
 * reference mode (clamp, MWC, phase) and the physical mode (periodic voxel
 * with wrap counts, Philox, moments, work counters), then reads out. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dsim_oracle.h"

static int fails = 0;

static void run(const char *name, const orc_mesh *m, int mode, int grid_kind)
{
    const uint32_t nt = m->n_tri;
    float *helpers = malloc(sizeof(float) * 12 * (nt + 1));
    float *mt = malloc(sizeof(float) * 16 * (nt + 1));
    orc_tri_helpers(m->vertices, m->trgls, nt, helpers);
    orc_tri_mt_records(m->vertices, m->trgls, nt, mt);
    uint32_t dims[3] = {20, 20, 20};
    float lo[3] = {-1.2f, -1.2f, -1.2f}, cell[3];
    if (grid_kind == 0) { lo[0] = lo[1] = lo[2] = -1.0f; }
    for (int k = 0; k < 3; ++k) cell[k] = (grid_kind == 0 ? 2.0f : 2.4f) / (float)dims[k];
    const uint32_t nc = dims[0] * dims[1] * dims[2];
    uint32_t *cnt = calloc(nc, 4), maxc;
    if (grid_kind == 0) maxc = orc_populate_cubes(m->vertices, m->trgls, nt, 20, cnt, NULL, 0);
    else if (grid_kind == 1) maxc = orc_populate_aabb(m->vertices, m->trgls, nt, dims, lo, cell, 1e-5f, cnt, NULL, 0);
    else maxc = orc_populate_aabb_plane(m->vertices, m->trgls, nt, dims, lo, cell, 1e-5f, cnt, NULL, 0);
    uint32_t *lst = calloc((size_t)nc * (maxc ? maxc : 1), 4);
    if (grid_kind == 0) orc_populate_cubes(m->vertices, m->trgls, nt, 20, cnt, lst, maxc);
    else if (grid_kind == 1) orc_populate_aabb(m->vertices, m->trgls, nt, dims, lo, cell, 1e-5f, cnt, lst, maxc);
    else orc_populate_aabb_plane(m->vertices, m->trgls, nt, dims, lo, cell, 1e-5f, cnt, lst, maxc);
    uint8_t fm[4096] = {0};
    orc_world w;
    memset(&w, 0, sizeof w);
    w.n_tri = nt; w.helpers = helpers; w.vertices = m->vertices; w.tri_info = m->tri_info;
    w.trgls = m->trgls; w.tri_mt = mt;
    for (int k = 0; k < 3; ++k) { w.dims[k] = dims[k]; w.lo[k] = lo[k]; w.cell[k] = cell[k]; w.inv_cell[k] = 1.0f / cell[k]; }
    w.ref_cell_math = grid_kind == 0; w.max_per_cell = maxc; w.cell_count = cnt; w.cell_list = lst;
    for (int c = 0; c < 4; ++c) { w.std_dev[c] = 0.03f; w.t2[c] = 80.0f; w.decay[c] = 0.005f / 80.0f; }
    w.dt = 0.005f; w.permeability = 0.3f; w.reflection_type = 1; w.spin_radius = 0.005f;
    w.boundary = (uint32_t)mode; w.test = (uint32_t)mode; w.traversal = (uint32_t)mode;
    w.transitions = (uint32_t)mode; w.fiber_has_myelin = fm; w.n_fibers = m->n_fibers;
    if (mode) for (int k = 0; k < 3; ++k) { w.wrap_lo[k] = -0.9f; w.wrap_hi[k] = 0.95f; }
    const uint32_t ns = 400;
    float *pos = malloc(sizeof(float) * 3 * ns);
    srand(5);
    for (uint32_t i = 0; i < 3 * ns; ++i) pos[i] = 1.8f * (float)rand() / (float)RAND_MAX - 0.9f;
    uint8_t *comp = calloc(ns, 1);
    uint16_t *fib = calloc(ns, 2);
    uint32_t *tries = calloc(ns, 4);
    orc_find_spins(&w, pos, ns, m->xmax, m->ymax, m->zmax, 20, mode ? 2u : 0u, comp, fib, tries);
    float *mag = malloc(4 * ns), *ph = calloc(ns, 4);
    float *disp = calloc(3 * (size_t)ns, 4), *mom = calloc(3 * (size_t)ns, 4), *momc = calloc(3 * (size_t)ns, 4);
    int32_t *wrap = calloc(3 * (size_t)ns, 4);
    uint32_t *seed = malloc(8 * ns), *trace = malloc(sizeof(uint32_t) * 5 * ns * 40);
    float *x0 = malloc(sizeof(float) * 3 * ns);
    memcpy(x0, pos, sizeof(float) * 3 * ns);
    for (uint32_t i = 0; i < ns; ++i) { mag[i] = 1.0f; seed[2 * i] = 12345u + i; seed[2 * i + 1] = 777u * i + 1u; }
    orc_state st;
    memset(&st, 0, sizeof st);
    st.n = ns; st.pos = pos; st.mag = mag; st.phase = ph; st.comp = comp; st.fiber = fib;
    st.seed = mode ? NULL : seed;
    if (mode) { st.wrap = wrap; }
    orc_step_args a;
    memset(&a, 0, sizeof a);
    uint64_t stats[3] = {0, 0, 0};
    a.rng = mode ? ORC_RNG_PHILOX : ORC_RNG_MWC; a.philox_key[0] = 3; a.philox_key[1] = 4;
    a.grad.x = 1e-4f; a.grad.z = -2e-4f; a.phase_constant = 1337.0f; a.n_threads = 2;
    a.trace = trace; a.stats = stats;
    orc_integrate(&w, &st, 40, &a);
    double out[5];
    orc_mr_signal(ns, mag, ph, comp, out);
    if (!(out[0] > 0.0 && out[0] <= 1.0)) fails++;
    if (mode) {
        /* moment mode on the same state */
        orc_state sm = st;
        sm.phase = NULL; sm.wrap = NULL; sm.disp = disp; sm.moment = mom; sm.moment_c = momc;
        a.trace = NULL; a.weight = 1.0f; a.step0 = 40;
        orc_integrate(&w, &sm, 30, &a);
        float g[6] = {1e-4f, 0, 0, 0, 1e-4f, 0};
        double md[10];
        orc_moment_signal(ns, mag, mom, x0, 30.0f, comp, 2, g, 1337.0f, md);
        if (!(md[0] > 0.0 && md[0] <= 1.0)) fails++;
    }
    printf("  %-24s mode %d grid %d: %u triangles, %u per cell, signal %.5f, hits %llu\n", name,
           mode, grid_kind, nt, maxc, out[0], (unsigned long long)stats[2]);
    free(pos); free(comp); free(fib); free(tries); free(mag); free(ph); free(seed); free(trace);
    free(disp); free(mom); free(momc); free(wrap); free(x0);
    free(helpers); free(mt); free(cnt); free(lst);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const char *names[] = {"sim.fibers_1_r004_f09", "sim.triangles_curved", "sim.triangles_8f_1g"};
    char path[4096];
    snprintf(path, sizeof path, "%s/sim.grads", argv[1]);
    float rows[6 * 64];
    int ng = orc_parse_grads(path, rows, 64);
    if (ng != 13) fails++;
    int sign[4096], lbr[4];
    for (int j = 0; j < ng && j < 64; ++j)
        if (orc_pgse_schedule(rows[6 * j], rows[6 * j + 1], rows[6 * j + 2], 0.005f, 10, sign, 4096, lbr) <= 0) fails++;
    for (int q = 0; q < 3; ++q) {
        snprintf(path, sizeof path, "%s/%s", argv[1], names[q]);
        orc_mesh m;
        memset(&m, 0, sizeof m);
        int rc = orc_parse_mesh(path, 6.0f, &m);
        /* the 8-fibre sample has no myelin sections: the reference rejects it
         * (-2) but the arrays are read, and the walk runs on them */
        if (rc == -1 || (rc == -2) != (q == 2)) { fprintf(stderr, "parse %s: %d\n", path, rc); fails++; }
        if (rc == -1) { orc_free_mesh(&m); continue; }
        run(names[q], &m, 0, 0);
        run(names[q], &m, 1, 1);
        run(names[q], &m, 1, 2);
        orc_free_mesh(&m);
    }
    orc_mesh bad;
    memset(&bad, 0, sizeof bad);
    snprintf(path, sizeof path, "%s/no_such_file", argv[1]);
    if (orc_parse_mesh(path, 1.0f, &bad) == 0) fails++;
    printf("oracle driver: %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
