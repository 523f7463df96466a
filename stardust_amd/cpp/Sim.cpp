// Sim.cpp -- `class Sim` over the C ABI (see Sim.h).
#include "Sim.h"

#include <cstdio>

namespace {
// Refusals of the shim itself (configurations this path does not provide) are
// sticky like the reference's g_failed (src/sim/Sim.cu:52-75).
bool g_shimFailed = false;
std::string g_shimMsg;

void shim_fail(const std::string& what) {
    if (!g_shimFailed) {
        g_shimFailed = true;
        g_shimMsg = what;
    }
    std::fprintf(stderr, "[sim] %s\n", what.c_str());
}

bool any_failure() { return g_shimFailed || sd_failed(); }

sd_config to_c(const SimConfig& c) {
    sd_config o;
    sd_config_default(&o);
    o.particleCount = c.particleCount;
    o.gridSize = c.gridSize;
    o.boundary = c.boundary == Boundary::Periodic ? SD_BOUNDARY_PERIODIC : SD_BOUNDARY_ISOLATED;
    o.preset = c.preset == Preset::Empty ? SD_PRESET_EMPTY : SD_PRESET_FILAMENT;
    o.gravity = c.gravity;
    o.softeningCells = c.softeningCells;
    o.timeScale = c.timeScale;
    o.sortInterval = c.sortInterval;
    o.lengthScale = c.lengthScale;
    o.timeUnitScale = c.timeUnitScale;
    o.softBoundR = c.softBoundR;
    o.darkEnergy = c.darkEnergy;
    o.contactEnabled = c.contactEnabled ? 1 : 0;
    o.contactStiffness = c.contactStiffness;
    o.contactDamping = c.contactDamping;
    o.pressureEnabled = c.pressureEnabled ? 1 : 0;
    o.coolingEnabled = c.coolingEnabled ? 1 : 0;
    o.starFormationEnabled = c.starFormationEnabled ? 1 : 0;
    o.blackHoleEnabled = c.blackHoleEnabled ? 1 : 0;
    o.strongForceEnabled = c.strongForceEnabled ? 1 : 0;
    o.emForceEnabled = c.emForceEnabled ? 1 : 0;
    o.weakForceEnabled = c.weakForceEnabled ? 1 : 0;
    o.collapseEnabled = c.collapseEnabled ? 1 : 0;
    o.contactRadiusCells = c.contactRadiusCells;
    o.pressureK = c.pressureK;
    o.coolingRate = c.coolingRate;
    return o;
}

// Pressure and cooling run in the reference kernel order with kick-drift: a
// configuration that enables them selects those modes first (the reference's
// own SimConfig defaults enable both).
void gas_modes(sd_sim* h, const SimConfig& c) {
    if (!c.pressureEnabled && !c.coolingEnabled) return;
    sd_modes m;
    sd_get_modes(h, &m);
    if (m.pipeline == SD_PIPELINE_REFERENCE && m.integrator == SD_INTEGRATOR_KD) return;
    m.pipeline = SD_PIPELINE_REFERENCE;
    m.integrator = SD_INTEGRATOR_KD;
    sd_set_modes(h, &m);
}
}  // namespace

Sim::Sim() : h_(sd_create()) {}
Sim::~Sim() { sd_destroy(h_); }

bool Sim::deviceAvailable() { return sd_device_available() != 0; }
std::string Sim::deviceName() {
    char b[256];
    return sd_device_name(b, sizeof(b)) == SD_OK ? std::string(b) : std::string();
}
std::string Sim::deviceDriver() {
    char b[128];
    return sd_device_driver(b, sizeof(b)) == SD_OK ? std::string(b) : std::string();
}
size_t Sim::deviceFreeBytes() { return sd_device_free_bytes(); }
int Sim::deviceMultiProcessors() { return sd_device_multiprocessors(); }
int Sim::maxGridSize(Boundary b) {
    return sd_max_grid_size(b == Boundary::Periodic ? SD_BOUNDARY_PERIODIC : SD_BOUNDARY_ISOLATED);
}
size_t Sim::estimateBytes(int n, int g, Boundary b) {
    return sd_estimate_bytes(n, g, b == Boundary::Periodic ? SD_BOUNDARY_PERIODIC : SD_BOUNDARY_ISOLATED);
}
int Sim::maxParticlesFor(int g, Boundary b, size_t freeBytes) {
    return sd_max_particles_for(g, b == Boundary::Periodic ? SD_BOUNDARY_PERIODIC : SD_BOUNDARY_ISOLATED, freeBytes);
}
double Sim::deviceBandwidthGBs() { return sd_device_bandwidth_gbs(); }
double Sim::estimateStepMs(int n, int g, Boundary b) {
    return sd_estimate_step_ms(n, g, b == Boundary::Periodic ? SD_BOUNDARY_PERIODIC : SD_BOUNDARY_ISOLATED);
}

void Sim::init(const SimConfig& cfg) {
    if (any_failure()) return;
    cfg_ = cfg;
    const sd_config c = to_c(cfg);
    gas_modes(h_, cfg);
    const int rc = sd_init(h_, &c);
    if (rc != SD_OK) shim_fail(std::string("init refused: ") + sd_fail_message());
    else cfg_.gridSize = sd_grid_size(h_);
}

void Sim::reconfigure(const SimConfig& cfg) {
    if (any_failure()) return;
    const sd_config c = to_c(cfg);
    gas_modes(h_, cfg);
    const int rc = sd_reconfigure(h_, &c);
    if (rc != SD_OK) {
        shim_fail(std::string("reconfigure refused: ") + sd_fail_message());
        return;
    }
    cfg_ = cfg;
    cfg_.gridSize = sd_grid_size(h_);
}

void Sim::reset() {
    if (any_failure()) return;
    sd_reset(h_);
}

void Sim::step() {
    if (any_failure()) return;
    sd_step(h_);
}

const SimConfig& Sim::config() const { return cfg_; }

SimTimings Sim::timings() const {
    sd_timings t{};
    sd_timings_get(h_, &t);
    SimTimings o;
    o.scatterMs = t.scatterMs;
    o.poissonMs = t.poissonMs;
    o.gatherMs = t.gatherMs;
    o.totalMs = t.totalMs;
    o.dtUsed = t.dtUsed;
    o.maxSpeed = t.maxSpeed;
    return o;
}

double Sim::simTime() const { return sd_sim_time(h_); }
bool Sim::failed() { return any_failure(); }
std::string Sim::failMessage() { return g_shimFailed ? g_shimMsg : std::string(sd_fail_message()); }
int Sim::gridSize() const { return sd_grid_size(h_); }
int Sim::particleCount() const { return sd_particle_count(h_); }
int Sim::activeCount() const { return sd_active_count(h_); }
int Sim::peakCellCount() const { return sd_peak_cell_count(h_); }
double Sim::measureTotalGridMass() { return any_failure() ? 0.0 : sd_measure_total_grid_mass(h_); }
double Sim::measureMaxDensity() { return any_failure() ? 0.0 : sd_measure_max_density(h_); }
int Sim::measureOccupiedCells() { return any_failure() ? 0 : sd_measure_occupied_cells(h_); }
void Sim::measureCentroid(double& cx, double& cy) {
    cx = cy = 0.5;
    if (!any_failure()) sd_measure_centroid(h_, &cx, &cy);
}
double Sim::measureKineticEnergy() { return any_failure() ? 0.0 : sd_measure_kinetic_energy(h_); }
bool Sim::saveState(const std::string& path) { return !any_failure() && sd_save_state(h_, path.c_str()) == SD_OK; }
bool Sim::loadState(const std::string& path) {
    if (any_failure()) return false;
    if (sd_load_state(h_, path.c_str()) != SD_OK) return false;
    sd_config c;
    sd_get_config(h_, &c);
    cfg_.particleCount = c.particleCount;
    cfg_.gridSize = c.gridSize;
    cfg_.boundary = c.boundary == SD_BOUNDARY_PERIODIC ? Boundary::Periodic : Boundary::Isolated;
    cfg_.gravity = c.gravity;
    cfg_.preset = Preset::Empty;
    return true;
}
const float* Sim::fieldDevicePtr(Field field) { return sd_field_device_ptr(h_, (int)field); }
const float* Sim::densityDevicePtr() const { return sd_density_device_ptr(h_); }
const float* Sim::particlePosDevicePtr() const { return sd_particle_pos_device_ptr(h_); }
const float* Sim::particleVelDevicePtr() const { return sd_particle_vel_device_ptr(h_); }

bool Sim::setParticles(const float* pos4, const float* vel4, int n) {
    return !any_failure() && sd_set_particles(h_, pos4, vel4, n) == SD_OK;
}
bool Sim::getParticles(float* pos4, float* vel4, int n) {
    return !any_failure() && sd_get_particles(h_, pos4, vel4, n) == SD_OK;
}
void Sim::setFixedDt(float dt) { sd_set_fixed_dt(h_, dt); }
bool Sim::setModes(const sd_modes& m) { return sd_set_modes(h_, &m) == SD_OK; }
