// dsim_example: the reference host's call sequence on libdsim, through the
// SpinSystem mirror (spin_system.h):
// initFibers, populateCubes, build, setArray, findSpinsInFibers, getArray,
// setGradient, updateSpins, getArray, getMrSignal
// (dSim3/src/spinSystem.cpp:290-673, 1094-1352, 178-281, 1584-1606,
// 1615-1852, 1544-1571, 1926-1971, 1452-1500).
//
//   dsim_example <mesh> <spaceScale> <numSpins> <iterations> <out.bin>
//
// The spins start at positions drawn from a fixed xorshift64 stream, uniform
// in [-0.9, 0.9]^3, with MWC seeds from the same stream; physics as the
// reference's sample sim.cfg (ADC 2.1/2.1/0.1 um^2/ms, T2 80/80/7.5 ms,
// dt 0.005 ms) with permeability 0.2 and a gradient of (80, 80, 0) mT/m.
// out.bin (little-endian): int32 n, then the labelled start state and the
// final state, each as positions f32[n*3], spinData[n] (12 bytes), seeds
// u32[n*2], then getMrSignal's five doubles.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/dsim.h"
#include "spin_system.h"

namespace {

struct XorShift {
    uint64_t s;
    uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    float unit() { return (float)((next() >> 40) * (1.0 / 16777216.0)); }  // [0, 1)
};

int die(const dsim::SpinSystem &ps) {
    fprintf(stderr, "dsim_example: %s\n", ps.error().c_str());
    return 1;
}

void write_state(FILE *f, int n, dsim::SpinSystem &ps) {
    std::vector<float> pos(3 * (size_t)n);
    std::vector<dsim_spin_data> spins(n);
    std::vector<uint32_t> seeds(2 * (size_t)n);
    ps.getArray(pos.data(), spins.data(), seeds.data());
    fwrite(pos.data(), 4, pos.size(), f);
    fwrite(spins.data(), sizeof(dsim_spin_data), spins.size(), f);
    fwrite(seeds.data(), 4, seeds.size(), f);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: dsim_example <mesh> <spaceScale> <numSpins> <iterations> <out.bin>\n");
        return 2;
    }
    const float space = (float)atof(argv[2]);
    const int n = atoi(argv[3]);
    const unsigned iters = (unsigned)atoi(argv[4]);
    const float dt = 0.005f;
    // SpinSystem(numSpins, useGpu, spaceScale, gyroMagneticRatio, useDisplay,
    //            triSearchMethod, reflectionType, extraAdc, myelinAdc, intraAdc,
    //            permeability, extraT2, myelinT2, intraT2, deltaTime, startBoxSize)
    dsim::SpinSystem ps(n, true, space, 42576.0f, false, 0, 1, 2.1f, 0.1f, 2.1f, 0.2f, 80.0f,
                        7.5f, 80.0f, dt, 0.9f);
    if (!ps.initFibers(argv[1])) return die(ps);
    ps.populateCubes();
    ps.constructAllRTrees();
    if (!ps.build()) return die(ps);

    XorShift rng{0x2545F4914F6CDD1Dull};
    std::vector<float> pos(3 * (size_t)n);
    std::vector<dsim_spin_data> spins(n);
    std::vector<uint32_t> seeds(2 * (size_t)n);
    for (auto &p : pos) p = 1.8f * rng.unit() - 0.9f;
    for (auto &s : seeds) s = (uint32_t)(rng.next() >> 32) | 1u;
    for (auto &s : spins) {
        s.signalMagnitude = 1.0f;
        s.signalPhase = 0.0f;
        s.compartmentType = 0;
        s.insideFiber = DSIM_NO_FIBER;
    }
    if (!ps.setArray(pos.data(), spins.data(), seeds.data())) return die(ps);
    ps.findSpinsInFibers();
    if (!ps.ok()) return die(ps);

    FILE *f = fopen(argv[5], "wb");
    if (!f) {
        fprintf(stderr, "dsim_example: cannot write %s\n", argv[5]);
        return 1;
    }
    const int32_t n32 = n;
    fwrite(&n32, 4, 1, f);
    write_state(f, n, ps);
    ps.setGradient({80.0f, 80.0f, 0.0f});
    ps.updateSpins(dt, iters);
    write_state(f, n, ps);
    double sig[5];
    ps.getMrSignal(sig);
    fwrite(sig, 8, 5, f);
    fclose(f);
    if (!ps.ok()) return die(ps);
    printf("dsim_example: %d spins, %u steps, signal %.6f\n", n, iters, sig[0]);
    return 0;
}
