// sd_sim.h -- the simulation handle behind the C ABI and the library-internal
// helpers the batch driver shares with it. Not installed; include/stardust/sd_api.h
// is the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/stardust/sd_api.h"
#include "sd_internal.h"

// Sticky failure state (Sim.cu:52-75): the first device error is kept until
// sd_clear_failure; argument errors are reported but not sticky.
#define SD_INTERNAL __attribute__((visibility("hidden")))
SD_INTERNAL void sd_mark_failure(const char* fmt, ...);
SD_INTERNAL bool sd_failure_set();
SD_INTERNAL int sd_arg_error(const char* what);

#define HK(x)                                                                              \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            sd_mark_failure("%s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__);      \
        }                                                                                  \
    } while (0)

struct KernelTimer {
    std::string name;
    std::vector<hipEvent_t> pool;  // start/stop pairs
    int used = 0;
    double totalMs = 0.0;
    long long launches = 0;
    int every = 1;        // time one launch in `every` (each bracket is two queue markers)
    long long seen = 0;   // launches of the named kernel so far
    bool armed = false;   // the current launch is bracketed
};

struct sd_sim {
    sd_config cfg{};
    sd_modes modes{};
    bool ready = false;
    int G = 0, S = 0, P = 0, periodic = 0;
    int allocN = 0;  // slots (mass per particle = gravity / allocN)
    int requestedN = -1;  // the configured particle count (allocN may be cut to fit the device)
    int active = 0;  // slots in use (prefix), for state files
    float fixedDt = 0.f;
    int device = -1;  // HIP device current when the handle was first configured
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;  // mesh stream of overlapped steps (joins st within the step)
    hipEvent_t evFork = nullptr, evJoin = nullptr;

    // particles (borrowed: the four arrays are slices of a batch's storage,
    // which owns them; their addresses never change)
    float4 *pos = nullptr, *vel = nullptr, *posTmp = nullptr, *velTmp = nullptr;
    bool borrowed = false;
    uint32_t *keysA = nullptr, *keysB = nullptr;
    int32_t *idxA = nullptr, *idxB = nullptr;
    uint32_t *hist = nullptr, *chunk = nullptr;
    sd::SortBuffers sb{};
    bool keysSorted = false;
    bool keysAscending = false;   // sb.keys holds an ascending array of allocN keys (the last sort's)
    sd::MoverBuffers mb{};        // lists of the sort of nearly sorted keys (allocated on first use)

    // mesh
    long long* rhoFix = nullptr;
    float* rhoF = nullptr;   // float density (float deposit mode; the density view)
    float* rhoX = nullptr;   // float density of a gather deposit with float nodes (the next x pass's input)
    sd::GatherFixBuffers gfix{};  // its far-mover tile bookkeeping
    float* proj = nullptr;   // G^2 projection of the density along z (the density field view)
    float2 *A = nullptr, *B = nullptr;
    float* green = nullptr;
    float* pot = nullptr;
    float4* accG = nullptr;
    float2* tw = nullptr;
    float4* accContact = nullptr;
    float* disp = nullptr;         // gas dispersion sums x, y, z and counts: 4 x float[G^3]
    float4* accPress = nullptr;    // gas pressure acceleration, float4[G^3]
    int2* cellRange = nullptr;  // contact cell table: (first, one past last) sorted index per cell
    int2* blockRange = nullptr;  // contact (Morton keys): sorted index range per 4^3-cell block
    long long blockCells = 0;
    long long cellTableCells = 0;
    sd::MeshPlan plan{};

    // reductions
    sd::ReducePartial* partials = nullptr;
    int nparts = 0;     // one-particle-per-thread reductions
    int npartsInt = 0;  // k_integrate partials
    sd::StepScalars* sc = nullptr;
    // a finalize whose launch rides on the next mesh solve's x pass (only
    // between steps of one sd_run; outside a run every finalize has been enqueued)
    bool finDeferred = false;
    sd::FinalizeJob finJob{};
    sd::StepScalars* scHost = nullptr;  // pinned mirror for reads

    // state
    bool rhoValid = false;   // rhoFix holds the deposit of the current positions
    bool rhoDirty = false;   // rhoFix not zero although !rhoValid (an x pass that left it for a gather deposit)
    bool rhoFloatOnly = false;  // with rhoValid: the deposit is in rhoX only (rhoFix is zero or stale)
    bool meshValid = false;  // pot matches the current positions
    bool accGValid = false;  // accG holds pot's gradient (with meshValid)
    bool primed = false;     // sc->dt / sc->kick are set for the next integrate
    long long stepCount = 0;

    // learned correction
    struct sd_mlp* mlp = nullptr;
    float mlpOutScale = 0.f;
    float4* accCorr = nullptr;
    uint16_t* logRho = nullptr;  // learned-correction density features (G^3 inference elements)
    int mlpInterval = 1;     // steps between recomputations of the correction
    bool corrValid = false;  // accCorr matches the current particle order
    int corrAge = 0;         // steps since it was computed

    // timings
    hipEvent_t evA = nullptr, ev1 = nullptr, evB = nullptr;
    sd_timings tm{};
    KernelTimer timer;
};


// Shared with the batch driver.
SD_INTERNAL int sd_check_ready(const sd_sim* s);
SD_INTERNAL sd::DtPolicy sd_dt_policy(const sd_sim* s);
SD_INTERNAL float sd_light_speed_sq(float lengthScale, float timeUnitScale);
// Brings a simulation to the state a fused step starts from (density of the
// current positions deposited, dt and kick of the next step set); enqueued on
// the simulation's stream.
SD_INTERNAL void sd_prepare_fused(sd_sim* s);
// Host bookkeeping after `steps` fused steps were run for s by other code.
SD_INTERNAL void sd_note_fused_steps(sd_sim* s, int steps);
