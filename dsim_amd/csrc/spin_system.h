// SpinSystem: host-side mirror of the reference's simulation object
// (dSim3/src/spinSystem.h:143-310) built on the dsim C ABI.  Same method
// names, argument meaning and call order as the reference driver uses them;
// the device work runs in libdsim's HIP kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dsim.h"
#include "../../include/dsim_io.h"

namespace dsim {

struct Float3 {
    float x, y, z;
};

class SpinSystem {
   public:
    // spinSystem.cpp:38-86 (argument order of the definition)
    SpinSystem(int numSpins, bool useGpu, float spaceScale, float gyroMagneticRatio,
               bool useDisplay, unsigned triSearchMethod, unsigned reflectionType, float extraAdc,
               float myelinAdc, float intraAdc, float permeability, float extraT2,
               float myelinT2, float intraT2, float deltaTime, float startBoxSize,
               uint64_t seed = 1);
    ~SpinSystem();
    SpinSystem(const SpinSystem &) = delete;
    SpinSystem &operator=(const SpinSystem &) = delete;

    bool initFibers(const char *fiberFile);   // spinSystem.cpp:290-673
    void populateCubes();                     // spinSystem.cpp:1094-1352 (50^3 grid)
    void constructAllRTrees() {}              // the uniform grid replaces the R-trees
    bool build();                             // spinSystem.cpp:178-281
    void resetSpins();                        // spinSystem.cpp:1386-1407
    void assignSeeds();                       // spinSystem.cpp:1532-1537
    void findSpinsInFibers();                 // spinSystem.cpp:1615-1852
    // setArray + setSpinArray (spinSystem.cpp:1584-1606) and getArray +
    // getSpinArray (spinSystem.cpp:1544-1571): host AoS positions [n*3],
    // spinData [n], MWC seeds [n*2]; any pointer may be null.
    bool setArray(const float *pos, const dsim_spin_data *spins, const uint32_t *seeds);
    bool getArray(float *pos, dsim_spin_data *spins, uint32_t *seeds);
    float updateSpins(float deltaTime, unsigned iterations);  // spinSystem.cpp:1926-1971
    void setGradient(Float3 g) { m_gradient = g; }            // spinSystem.h:163
    double getMrSignal();                                      // spinSystem.cpp:1418-1450
    void getMrSignal(double mrSignal[]);                       // spinSystem.cpp:1452-1500
    int getNumSpins() const { return m_numSpins; }
    unsigned getNumCompartments() const { return 4; }
    unsigned getNumFibers() const { return m_info.n_fibers; }

    // Shared-walk PGSE (one walk serves every gradient line of one timing).
    void beginAcquisition();
    void walkLobe(float weight, unsigned steps);
    void signalsForGradients(const std::vector<Float3> &g, std::vector<double> &out5);

    const std::string &error() const { return m_error; }
    bool ok() const { return m_error.empty(); }

   private:
    bool check(int rc, const char *what);

    int m_numSpins;
    float m_spaceScale, m_gamma, m_permeability, m_deltaTime, m_startBox;
    unsigned m_reflectionType;
    float m_std[4], m_t2[4];
    uint64_t m_seed;
    uint64_t m_resets = 0;
    Float3 m_gradient{0, 0, 0};
    dsim_mesh_file *m_mesh = nullptr;
    dsim_mesh_info m_info{};
    dsim_substrate *m_sub = nullptr;
    dsim_sim *m_sim = nullptr;
    std::string m_error;
};

}  // namespace dsim
