// SpinSystem over the dsim C ABI (see spin_system.h).
#include "spin_system.h"

#include <cmath>
#include <cstdio>
#include <random>

namespace dsim {

SpinSystem::SpinSystem(int numSpins, bool /*useGpu*/, float spaceScale, float gyroMagneticRatio,
                       bool /*useDisplay*/, unsigned /*triSearchMethod*/, unsigned reflectionType,
                       float extraAdc, float myelinAdc, float intraAdc, float permeability,
                       float extraT2, float myelinT2, float intraT2, float deltaTime,
                       float startBoxSize, uint64_t seed)
    : m_numSpins(numSpins), m_spaceScale(spaceScale), m_gamma(gyroMagneticRatio),
      m_permeability(permeability), m_deltaTime(deltaTime), m_startBox(startBoxSize),
      m_reflectionType(reflectionType), m_seed(seed) {
    // spinSystem.cpp:47-56: per-compartment T2 and step deviation, compartments
    // {extra, intra-axonal, myelin, glia interior (intra values)}
    m_t2[0] = extraT2; m_t2[1] = intraT2; m_t2[2] = myelinT2; m_t2[3] = intraT2;
    float s2 = spaceScale * spaceScale;
    m_std[0] = std::sqrt(2 * extraAdc / s2 * deltaTime);
    m_std[1] = std::sqrt(2 * intraAdc / s2 * deltaTime);
    m_std[2] = std::sqrt(2 * myelinAdc / s2 * deltaTime);
    m_std[3] = std::sqrt(2 * intraAdc / s2 * deltaTime);
}

SpinSystem::~SpinSystem() {
    if (m_sim) dsim_sim_destroy(m_sim);
    if (m_sub) dsim_substrate_destroy(m_sub);
    if (m_mesh) dsim_mesh_free(m_mesh);
}

bool SpinSystem::check(int rc, const char *what) {
    if (rc == DSIM_OK) return true;
    const char *msg = dsim_last_error();
    const char *io = dsim_io_last_error();
    m_error = std::string(what) + ": " + ((msg && *msg) ? msg : (io ? io : "error"));
    return false;
}

bool SpinSystem::initFibers(const char *fiberFile) {
    if (m_mesh) dsim_mesh_free(m_mesh);
    m_mesh = nullptr;
    if (!check(dsim_mesh_read(fiberFile, m_spaceScale, &m_mesh), "initFibers")) return false;
    return check(dsim_mesh_view(m_mesh, nullptr, &m_info), "initFibers");
}

void SpinSystem::populateCubes() {
    if (!m_mesh || m_sub) return;
    dsim_mesh view;
    dsim_mesh_view(m_mesh, &view, nullptr);
    dsim_grid_params g{};
    g.kind = DSIM_GRID_REFERENCE;
    g.dims[0] = g.dims[1] = g.dims[2] = 50;  // spinSystem.cpp:1113
    check(dsim_substrate_create(&view, &g, &m_sub), "populateCubes");
}

bool SpinSystem::build() {
    if (!m_sub) populateCubes();
    if (!m_sub) return false;
    dsim_params p{};
    p.dt = m_deltaTime;
    for (int c = 0; c < 4; ++c) { p.std_dev[c] = m_std[c]; p.t2[c] = m_t2[c]; }
    p.permeability = m_permeability;
    p.reflection_type = m_reflectionType;
    p.mode = DSIM_MODE_REFERENCE;
    p.boundary = DSIM_BOUNDARY_CLAMP;
    p.rng = DSIM_RNG_MWC;
    p.seed = m_seed;
    p.spin_radius = 0.005f;
    return check(dsim_sim_create(m_sub, &p, (uint32_t)m_numSpins, &m_sim), "build");
}

void SpinSystem::resetSpins() {
    if (!m_sim) return;
    // new spins for every reset: the position draw takes the handle's step
    // counter, set to the reset count
    dsim_set_step(m_sim, m_resets);
    if (!check(dsim_init_spins_uniform(m_sim, m_startBox), "resetSpins")) return;
    dsim_set_step(m_sim, 0);
    findSpinsInFibers();
    assignSeeds();
    m_resets++;
}

void SpinSystem::findSpinsInFibers() {
    if (!m_sim) return;
    uint32_t amb = 0;
    if (!check(dsim_label_spins(m_sim, m_info.xmax, m_info.ymax, m_info.zmax, 20, &amb),
               "findSpinsInFibers"))
        return;
    if (amb) fprintf(stderr, "findSpinsInFibers: %u spins stayed on membrane edges\n", amb);
}

bool SpinSystem::setArray(const float *pos, const dsim_spin_data *spins, const uint32_t *seeds) {
    return m_sim && check(dsim_set_state_aos(m_sim, pos, spins, seeds), "setArray");
}

bool SpinSystem::getArray(float *pos, dsim_spin_data *spins, uint32_t *seeds) {
    return m_sim && check(dsim_get_state_aos(m_sim, pos, spins, seeds), "getArray");
}

void SpinSystem::assignSeeds() {
    // spinSystem.cpp:1532-1537 draws rand()*(UINT_MAX/RAND_MAX); a seeded
    // Mersenne twister stands in for the process-wide rand()
    std::mt19937_64 gen(m_seed * 0x9E3779B97F4A7C15ull + m_resets);
    std::vector<uint32_t> s(2 * (size_t)m_numSpins);
    for (auto &v : s) v = (uint32_t)(gen() >> 32) | 1u;
    check(dsim_set_state_aos(m_sim, nullptr, nullptr, s.data()), "assignSeeds");
}

float SpinSystem::updateSpins(float deltaTime, unsigned iterations) {
    if (!m_sim || iterations == 0) return 0;
    // spinSystem.cpp:1931-1938
    float g[3] = {m_gradient.x * 1e-9f * m_spaceScale, m_gradient.y * 1e-9f * m_spaceScale,
                  m_gradient.z * 1e-9f * m_spaceScale};
    float pc = 6.28318530717959f * m_gamma * deltaTime;
    check(dsim_integrate(m_sim, g, pc, iterations), "updateSpins");
    return 0;
}

double SpinSystem::getMrSignal() {
    double s[5] = {0, 0, 0, 0, 0};
    getMrSignal(s);
    return s[0];
}

void SpinSystem::getMrSignal(double mrSignal[]) {
    if (!m_sim) return;
    check(dsim_signal(m_sim, mrSignal), "getMrSignal");
}

void SpinSystem::beginAcquisition() {
    if (m_sim) check(dsim_moment_reset(m_sim), "beginAcquisition");
}

void SpinSystem::walkLobe(float weight, unsigned steps) {
    if (m_sim && steps) check(dsim_integrate_moment(m_sim, weight, steps), "walkLobe");
}

void SpinSystem::signalsForGradients(const std::vector<Float3> &g, std::vector<double> &out5) {
    out5.assign(g.size() * 5, 0.0);
    if (!m_sim || g.empty()) return;
    std::vector<float> gs(g.size() * 3);
    for (size_t j = 0; j < g.size(); ++j) {
        gs[3 * j] = g[j].x * 1e-9f * m_spaceScale;
        gs[3 * j + 1] = g[j].y * 1e-9f * m_spaceScale;
        gs[3 * j + 2] = g[j].z * 1e-9f * m_spaceScale;
    }
    float pc = 6.28318530717959f * m_gamma * m_deltaTime;
    check(dsim_signal_moment(m_sim, (uint32_t)g.size(), gs.data(), pc, out5.data()),
          "signalsForGradients");
}

}  // namespace dsim
