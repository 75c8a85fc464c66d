// dsim: command-line driver, the drop-in for the reference's `diffusion`
// program (dSim3/src/diffusion.cpp:1068-1184 with update() at :522-737).
//
//   dsim [-config=sim.cfg] [-numSpins=N] [-perm=p] [-startBox=b] [-fiberFile=f]
//        [-stepsPerUpdate=S] [-w|-nw] [-seed=K] [-sharedWalk]
//
// Reads sim.cfg, the gradient file and the mesh; for every gradient line it
// resets the spins, plays the PGSE on/off sequence of update() and reads the
// per-compartment signal; then fits the tensors and appends the MATLAB .m
// output.  -sharedWalk walks the spins once per distinct (delta, Delta,
// readout) timing and derives every gradient line of that timing from the
// gradient moments (same expectation, one walk instead of one per line).
// Display flags are accepted and ignored (no renderer); -cpu is accepted but
// the simulate path always runs on the GPU.
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/dsim.h"
#include "../../include/dsim_io.h"
#include "spin_system.h"

namespace {

int die(const std::string &msg) {
    fprintf(stderr, "dsim: %s\n", msg.c_str());
    return 1;
}

bool has_flag(int argc, char **argv, const char *name) {
    for (int a = 1; a < argc; ++a) {
        const char *s = argv[a];
        while (*s == '-') ++s;
        if (strcmp(s, name) == 0) return true;
    }
    return false;
}

}  // namespace

int main(int argc, char **argv) {
    dsim_config cfg;
    dsim_config_defaults(&cfg);
    dsim_config_apply_args(&cfg, argc, argv);  // picks up -config= first
    std::string cfg_path = cfg.configFile;
    if (dsim_config_read(cfg_path.c_str(), &cfg) != DSIM_OK)
        fprintf(stderr, "dsim: %s (continuing with defaults)\n", dsim_io_last_error());
    dsim_config_apply_args(&cfg, argc, argv);  // command line overrides the file
    const bool shared_walk = has_flag(argc, argv, "sharedWalk");
    if (cfg.seed == 0) cfg.seed = (uint64_t)time(nullptr);

    std::vector<float> rows(6 * 4096);
    uint32_t ng = 0;
    if (cfg.gradsFile[0] && dsim_grads_read(cfg.gradsFile, rows.data(), 4096, &ng) != DSIM_OK)
        return die(dsim_io_last_error());
    rows.resize(6 * (size_t)std::min<uint32_t>(ng, 4096));
    ng = (uint32_t)(rows.size() / 6);
    printf("Loading gradient sequence from %s: %u lines\n", cfg.gradsFile, ng);

    dsim::SpinSystem ps((int)cfg.numSpins, cfg.useGpu != 0, cfg.spaceScale, cfg.gyroMagneticRatio,
                        cfg.useDisplay != 0, cfg.searchMethod, 1u, cfg.extraAdc, cfg.myelinAdc,
                        cfg.intraAdc, cfg.permeability, cfg.extraT2, cfg.myelinT2, cfg.intraT2,
                        cfg.timeStep, cfg.startBox, cfg.seed);
    if (!ps.initFibers(cfg.fiberFile)) return die(ps.error());
    ps.populateCubes();
    ps.constructAllRTrees();
    if (!ps.build()) return die(ps.error());

    const int S = cfg.stepsPerUpdate > 0 ? cfg.stepsPerUpdate : 1;
    std::vector<float> sig(5 * (size_t)ng, 0.0f);  // gradientDataResults columns 6..10
    if (!shared_walk) {
        for (uint32_t j = 0; j < ng; ++j) {
            const float *r = &rows[6 * (size_t)j];
            int32_t L, B, R, ro;
            dsim_pgse_schedule(r[0], r[1], r[2], cfg.timeStep, S, nullptr, 0, &L, &B, &R, &ro);
            std::vector<int32_t> sign((size_t)(L + B + R) + 1);
            dsim_pgse_schedule(r[0], r[1], r[2], cfg.timeStep, S, sign.data(), (uint32_t)sign.size(),
                               &L, &B, &R, &ro);
            printf("Run no. %u\nRunning: d=%gms, D=%gms, readout=%gms, G=[%g,%g,%g]mT/m.\n", j + 1,
                   r[0], r[1], r[2], r[3], r[4], r[5]);
            ps.resetSpins();
            int n = L + B + R, k = 0;
            while (k < n) {  // runs of equal lobe sign become one updateSpins call
                int k1 = k;
                while (k1 < n && sign[k1] == sign[k]) ++k1;
                float g = (float)sign[k];
                ps.setGradient({r[3] * g, r[4] * g, r[5] * g});
                ps.updateSpins(cfg.timeStep, (unsigned)((k1 - k) * S));
                k = k1;
            }
            if (ro) {
                double s5[5];
                ps.getMrSignal(s5);
                for (int c = 0; c < 5; ++c) sig[5 * j + c] = (float)s5[c];
            }
            if (!ps.ok()) return die(ps.error());
            printf("MR Signal: %0.4f\n", sig[5 * j]);
        }
    } else {
        std::map<std::tuple<float, float, float>, std::vector<uint32_t>> groups;
        for (uint32_t j = 0; j < ng; ++j)
            groups[{rows[6 * j], rows[6 * j + 1], rows[6 * j + 2]}].push_back(j);
        for (auto &grp : groups) {
            auto [ld, bd, rdo] = grp.first;
            int32_t L, B, R, ro;
            dsim_pgse_schedule(ld, bd, rdo, cfg.timeStep, S, nullptr, 0, &L, &B, &R, &ro);
            std::vector<int32_t> sign((size_t)(L + B + R) + 1);
            dsim_pgse_schedule(ld, bd, rdo, cfg.timeStep, S, sign.data(), (uint32_t)sign.size(),
                               &L, &B, &R, &ro);
            ps.resetSpins();
            ps.beginAcquisition();
            int n = L + B + R, k = 0;
            while (k < n) {
                int k1 = k;
                while (k1 < n && sign[k1] == sign[k]) ++k1;
                ps.walkLobe((float)sign[k], (unsigned)((k1 - k) * S));
                k = k1;
            }
            std::vector<dsim::Float3> g;
            for (uint32_t j : grp.second) g.push_back({rows[6 * j + 3], rows[6 * j + 4], rows[6 * j + 5]});
            std::vector<double> out;
            ps.signalsForGradients(g, out);
            if (!ps.ok()) return die(ps.error());
            for (size_t q = 0; q < grp.second.size(); ++q)
                if (ro)
                    for (int c = 0; c < 5; ++c) sig[5 * grp.second[q] + c] = (float)out[5 * q + c];
        }
    }

    printf("All runs finished!\n");
    dsim_tensor_fit fits[5];
    memset(fits, 0, sizeof fits);
    for (int c = 0; c < 5; ++c) {
        std::vector<float> sc(ng);
        for (uint32_t j = 0; j < ng; ++j) sc[j] = sig[5 * j + c];
        if (dsim_tensor_fit_compute(ng, rows.data(), sc.data(), &fits[c]) == DSIM_OK)
            printf("Compartment %d: md %g fa %g rd %g\n", c - 1, fits[c].md, fits[c].fa, fits[c].rd);
    }
    for (uint32_t j = 0; j < ng; ++j) printf("Total MR signal, run %u: %g\n", j + 1, sig[5 * j]);
    if (cfg.writeOut && cfg.outFile[0]) {
        if (dsim_write_m(cfg.outFile, &cfg, ng, rows.data(), sig.data(), fits) != DSIM_OK)
            return die(dsim_io_last_error());
        printf("Results appended to %s\n", cfg.outFile);
    }
    return 0;
}
