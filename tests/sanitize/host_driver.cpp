// Host-code driver for the AddressSanitizer / UndefinedBehaviorSanitizer run
// (tests/sanitize/run.sh): exercises the host half of libdsim without a GPU
// (formats, host substrate builds of every grid kind, PGSE schedule, tensor
// fit, .m writer) on the reference's sample inputs.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dsim.h"
#include "dsim_io.h"

static int fails = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                                     \
        }                                                                \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: host_driver <reference_inputs dir> <tmp dir>\n");
        return 2;
    }
    const std::string in = argv[1], tmp = argv[2];
    dsim_config cfg;
    dsim_config_defaults(&cfg);
    EXPECT(dsim_config_read((in + "/sim.cfg").c_str(), &cfg) == DSIM_OK);
    const char *args[] = {"dsim", "-numSpins=1234", "-perm=0.25", "-w"};
    dsim_config_apply_args(&cfg, 4, args);
    EXPECT(cfg.numSpins == 1234);
    std::vector<float> rows(6 * 64);
    uint32_t ng = 0;
    EXPECT(dsim_grads_read((in + "/sim.grads").c_str(), rows.data(), 64, &ng) == DSIM_OK);
    EXPECT(ng == 13);
    for (const char *name : {"sim.fibers_1_r004_f09", "sim.fibers_1_r0064_f05", "sim.triangles_curved",
                             "sim.triangles_8f_1g"}) {
        dsim_mesh_file *mf = nullptr;
        if (dsim_mesh_read((in + "/" + name).c_str(), 6.0f, &mf) != DSIM_OK) {
            fprintf(stderr, "%s: %s\n", name, dsim_io_last_error());
            continue;
        }
        dsim_mesh mesh;
        dsim_mesh_info info;
        EXPECT(dsim_mesh_view(mf, &mesh, &info) == DSIM_OK);
        for (uint32_t kind : {DSIM_GRID_REFERENCE, DSIM_GRID_AABB, DSIM_GRID_TRIBOX}) {
            dsim_grid_params g{};
            g.kind = kind;
            if (kind == DSIM_GRID_REFERENCE) {
                g.dims[0] = g.dims[1] = g.dims[2] = 20;
            } else {
                for (int k = 0; k < 3; ++k) {
                    g.dims[k] = 17 + 3 * k;
                    g.lo[k] = -1.2f;
                    g.cell[k] = 2.4f / (float)g.dims[k];
                }
                g.eps = 1e-5f;
            }
            dsim_substrate *s = nullptr;
            EXPECT(dsim_substrate_create(&mesh, &g, &s) == DSIM_OK);
            uint32_t nc = 0, ne = 0, mx = 0, axis = 9;
            EXPECT(dsim_substrate_grid_info(s, &nc, &ne, &mx) == DSIM_OK);
            std::vector<uint32_t> start(nc + 1), tri(ne ? ne : 1);
            EXPECT(dsim_substrate_get_grid(s, start.data(), tri.data()) == DSIM_OK);
            EXPECT(start[nc] == ne);
            std::vector<float> h((size_t)info.n_triangles * 12 + 1);
            EXPECT(dsim_substrate_get_helpers(s, h.data()) == DSIM_OK);
            if (kind != DSIM_GRID_REFERENCE) {
                std::vector<float> cl(nc);
                EXPECT(dsim_substrate_get_clearance(s, cl.data()) == DSIM_OK);
            }
            EXPECT(dsim_substrate_label_axis(s, &axis) == DSIM_OK && axis < 3);
            printf("  %-24s grid %u: %u triangles, %u cells, %u entries\n", name, kind,
                   info.n_triangles, nc, ne);
            dsim_substrate_destroy(s);
        }
        dsim_mesh_free(mf);
    }
    // PGSE schedule and the tensor fit on synthetic signals
    int32_t L, B, R, ro;
    std::vector<int32_t> sign(1024);
    EXPECT(dsim_pgse_schedule(10.0f, 12.0669f, 1.0f, 0.005f, 10, sign.data(), 1024, &L, &B, &R, &ro) ==
           DSIM_OK);
    std::vector<float> sig(ng);
    for (uint32_t j = 0; j < ng; ++j) sig[j] = j == 0 ? 0.8f : 0.2f + 0.01f * (float)j;
    dsim_tensor_fit fits[5] = {};
    EXPECT(dsim_tensor_fit_compute(ng, rows.data(), sig.data(), &fits[0]) == DSIM_OK);
    std::vector<float> sig5(5 * ng, 0.3f);
    EXPECT(dsim_write_m((tmp + "/san.m").c_str(), &cfg, ng, rows.data(), sig5.data(), fits) == DSIM_OK);
    // rejected inputs return errors, never abort
    dsim_mesh_file *bad = nullptr;
    EXPECT(dsim_mesh_read((in + "/no_such_file").c_str(), 1.0f, &bad) != DSIM_OK);
    EXPECT(dsim_substrate_create(nullptr, nullptr, nullptr) != DSIM_OK);
    printf("host driver: %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
