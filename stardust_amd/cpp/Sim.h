// Sim.h -- the reference's `class Sim` interface (src/sim/Sim.h of
// chdnl0420-svg/stardust) implemented over the C ABI in
// include/stardust/sd_api.h, for C++ callers that drop this library in for the
// particle-mesh gravity path.
//
// Names, argument meaning and error behaviour follow the reference: no
// exceptions, a process-wide sticky failure flag (Sim::failed / failMessage),
// methods return early after a failure. The plain-PM path is the only physics
// provided: SimConfig keeps every reference field so configurations compile
// and map unchanged, but enabling an off-path feature (pressure, cooling,
// stars, black holes, strong / EM / weak forces, collapse) or the Filament
// preset makes init / reconfigure fail with a message.
#pragma once

#include <cstddef>
#include <string>

#include "../../include/stardust/sd_api.h"

enum class Preset { Filament, Empty };       // Sim.h:14-17
enum class Boundary { Isolated, Periodic };  // Sim.h:22

// SimConfig, Sim.h:157-796: field names and defaults of the reference.
struct SimConfig {
    int particleCount = 2000000;
    int gridSize = 128;
    float lengthScale = 1.0f;
    float timeUnitScale = 1.0f;
    float hubble = 0.0f;
    float bigBangShrink = 0.0f;
    float softBoundR = 0.0f;
    float darkEnergy = 0.0f;
    float velDispersion = 0.0f;
    float diskThickness = 0.001f;
    Preset preset = Preset::Filament;
    Boundary boundary = Boundary::Isolated;
    float gravity = 0.6f;
    float bulgeFraction = 0.11f;
    float bulgeRadius = 0.045f;
    float softeningCells = 3.0f;
    float timeScale = 1.0f;
    int sortInterval = 40;
    bool pressureEnabled = true;
    float pressureK = 0.45f;
    bool coolingEnabled = true;
    float coolingRate = 0.08f;
    bool starFormationEnabled = false;
    float starFormEfficiency = 0.005f;
    float starJeansK = 20000.0f;
    float starSunMass = 3.0f;
    float starSunLifeSim = 1000.0f;
    float starExplodeSim = 0.288f;
    float starKickSpeed = 0.1f;
    float novaEnergyK = 60.0f;
    float bhFrictionK = 1.0f;
    float starAshYield = 1.0f;
    float ashCoolK = 0.05f;
    float starIonizeK = 1.0f;
    float starEmbedTime = 0.065f;
    float ashDiffuseK = 1.0f;
    float darkMatterFraction = 0.0f;
    float haloGasFraction = 0.30f;
    float sphereStart = 1.0f;
    float diskDispersion = 0.0f;
    float diskSpinLag = 0.0f;
    float starWindRate = 0.06f;
    float starBHRatio = 20.0f;
    bool starCollapseToBH = false;
    bool contactEnabled = false;
    float contactStiffness = 1.0e6f;
    float contactDamping = 0.35f;
    bool blackHoleEnabled = false;
    float centralBHFraction = 0.02f;
    float blackHoleMassCap = 0.02f;
    float blackHoleMassScale = 0.02f;
    bool strongForceEnabled = false;
    float strongForceK = 8000.0f;
    bool emForceEnabled = false;
    float emForceK = 0.004f;
    bool weakForceEnabled = false;
    float weakForceRate = 0.05f;
    float newForceDamping = 1.0f;
    float spin = 0.0f;
    bool collapseEnabled = false;
    float collapseDensity = 150.0f;
    // extension: contact radius in cells (the reference fixes 0.5, Sim.cu:3694)
    float contactRadiusCells = 0.5f;
};

// SimTimings, Sim.h:844-852.
struct SimTimings {
    float scatterMs = 0.0f;
    float poissonMs = 0.0f;
    float gatherMs = 0.0f;
    float totalMs = 0.0f;
    float dtUsed = 0.0f;
    float maxSpeed = 0.0f;
};

class Sim {
public:
    Sim();
    ~Sim();
    Sim(const Sim&) = delete;
    Sim& operator=(const Sim&) = delete;

    static bool deviceAvailable();
    static std::string deviceName();
    static std::string deviceDriver();
    static size_t deviceFreeBytes();
    static int deviceMultiProcessors();
    static int maxGridSize(Boundary boundary);
    static size_t estimateBytes(int particleCount, int gridSize, Boundary boundary);
    static int maxParticlesFor(int gridSize, Boundary boundary, size_t freeBytes);
    static double deviceBandwidthGBs();
    static double estimateStepMs(int particleCount, int gridSize, Boundary boundary);

    void init(const SimConfig& cfg);
    void reconfigure(const SimConfig& cfg);
    void reset();
    void step();

    const SimConfig& config() const;
    SimTimings timings() const;
    double simTime() const;

    static bool failed();
    static std::string failMessage();
    int gridSize() const;
    int particleCount() const;
    int activeCount() const;
    int peakCellCount() const;

    double measureTotalGridMass();
    double measureMaxDensity();
    int measureOccupiedCells();
    void measureCentroid(double& cx, double& cy);
    double measureKineticEnergy();
    bool saveState(const std::string& path);
    bool loadState(const std::string& path);

    // Sim::Field (Sim.h:1078); only Density is on this path
    enum class Field { Density, Dispersion, Light, Ash };
    const float* fieldDevicePtr(Field field);  // G^2 floats: density summed along z
    const float* densityDevicePtr() const;     // G^3 floats, row pitch G
    const float* particlePosDevicePtr() const;  // float4 array as float*
    const float* particleVelDevicePtr() const;

    // extensions (no reference counterpart): host particle I/O, fixed dt, the
    // execution modes and the underlying C handle
    bool setParticles(const float* pos4, const float* vel4, int n);
    bool getParticles(float* pos4, float* vel4, int n);
    void setFixedDt(float dt);
    bool setModes(const sd_modes& m);
    sd_sim* handle() { return h_; }

private:
    sd_sim* h_;
    SimConfig cfg_;
};
