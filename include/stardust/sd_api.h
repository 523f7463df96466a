/*
 * sd_api.h -- C ABI of the MI355X particle-mesh gravity step.
 *
 * Drop-in replacement for the PM path behind Stardust's `class Sim`
 * (src/sim/Sim.h:854-1135 of chdnl0420-svg/stardust). Plain C: an opaque
 * handle, POD structs, plain pointers and sizes, int status codes (0 = ok,
 * negative = error) and a process-wide sticky failure flag with a message, as
 * the reference's g_failed / g_failMsg (Sim.cu:52-75). After a failure every
 * call that launches GPU work returns SD_ERR_FAILED until sd_clear_failure().
 *
 * Every entry point names the reference interface it replaces. Entry points
 * marked "no reference counterpart" extend the API for the parity harness,
 * the force-error metric, the learned correction and batched scenes.
 *
 * Threading: one host thread per handle; every handle owns one HIP stream
 * (default priority). sd_step() is synchronous like Sim::step (Sim.cu:4506);
 * sd_run() enqueues several steps and synchronises once.
 */
#ifndef STARDUST_SD_API_H
#define STARDUST_SD_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SD_OK 0
#define SD_ERR_FAILED -1     /* sticky failure set (see sd_fail_message) */
#define SD_ERR_ARG -2        /* invalid argument / unsupported configuration */
#define SD_ERR_DEVICE -3     /* HIP runtime error */
#define SD_ERR_STATE -4      /* call not valid in the current state */
#define SD_ERR_IO -5         /* file error */

typedef struct sd_sim sd_sim;

/* Boundary, Sim.h:22. */
#define SD_BOUNDARY_ISOLATED 0
#define SD_BOUNDARY_PERIODIC 1

/* Preset, Sim.h:14-17. Filament is the reference's Big-Bang scene generator,
 * which this library does not provide (sd_init fails with SD_ERR_ARG); Empty
 * leaves every slot inactive until sd_set_particles / sd_load_state. */
#define SD_PRESET_FILAMENT 0
#define SD_PRESET_EMPTY 1

/* Subset of SimConfig (Sim.h:157-796) that the PM step reads, with the same
 * names and defaults (sd_config_default), except that pressure and cooling
 * default to off here (the plain-PM path). Pressure and cooling (the gas
 * velocity-dispersion grids, Sim.cu:1212-1415, 3605-3633, 3732-3830) run in the
 * reference kernel order only: they need SD_PIPELINE_REFERENCE and
 * SD_INTEGRATOR_KD (else SD_ERR_ARG), and treat particles with pos.w <= 0 as
 * gas, pos.w > 0 as stars (no cooling, no pressure) and vel.w < -50 as dark
 * matter (neither), as the reference does. The other physics switches are
 * kept so a configuration maps field for field; any of them set makes
 * sd_init / sd_reconfigure fail with SD_ERR_ARG. */
typedef struct sd_config {
    int particleCount;       /* 2000000 */
    int gridSize;            /* 128 */
    int boundary;            /* SD_BOUNDARY_ISOLATED */
    int preset;              /* SD_PRESET_FILAMENT in the reference; */
    float gravity;           /* 0.6 */
    float softeningCells;    /* 3.0 */
    float timeScale;         /* 1.0 */
    int sortInterval;        /* 40 */
    float lengthScale;       /* 1.0 (light-speed clamp, Sim.h lightSpeedSqFor) */
    float timeUnitScale;     /* 1.0 */
    float softBoundR;        /* 0 */
    float darkEnergy;        /* 0 */
    int contactEnabled;      /* 0 */
    float contactStiffness;  /* 1e6 */
    float contactDamping;    /* 0.35 */
    /* gas physics (reference kernel order only, see above) */
    int pressureEnabled;     /* 0 here (the reference's default is 1) */
    int coolingEnabled;      /* 0 here (the reference's default is 1) */
    /* other physics (must be 0 here) */
    int starFormationEnabled;
    int blackHoleEnabled;
    int strongForceEnabled;
    int emForceEnabled;
    int weakForceEnabled;
    int collapseEnabled;
    /* extension: contact radius in cells (the reference hard-codes 0.5,
     * Sim.cu:3694) */
    float contactRadiusCells;
    float pressureK;         /* 0.45 (Sim.h:320) */
    float coolingRate;       /* 0.08 (Sim.h:345) */
} sd_config;

/* Per-step timings, SimTimings (Sim.h:844-852). scatterMs = gravity phase
 * (sort, deposit, Poisson, gradient, contact), poissonMs = 0 (cooling / stars
 * are not on this path), gatherMs = integrate phase; all from HIP events. */
typedef struct sd_timings {
    float scatterMs;
    float poissonMs;
    float gatherMs;
    float totalMs;
    float dtUsed;
    float maxSpeed;
} sd_timings;

/* Execution modes (no reference counterpart). Defaults are the product's. */
#define SD_KEY_ROWMAJOR 0     /* kCellKey order, Sim.cu:2476-2486 */
#define SD_KEY_MORTON 1
#define SD_DEPOSIT_FIXED 0    /* int64 fixed point, deterministic */
#define SD_DEPOSIT_FLOAT 1    /* float atomics as kScatter (order-dependent) */
#define SD_INTEGRATOR_KD 0    /* v += a dt; x += v dt (Sim.cu:822,842) */
#define SD_INTEGRATOR_KDK 1   /* leapfrog; velocities synchronised on read */
#define SD_PIPELINE_REFERENCE 0 /* reference kernel order, separate reductions */
#define SD_PIPELINE_FUSED 1     /* deposit + reductions fused into integrate */

typedef struct sd_modes {
    int keyMode;        /* default SD_KEY_MORTON */
    int depositMode;    /* default SD_DEPOSIT_FIXED */
    int integrator;     /* default SD_INTEGRATOR_KDK */
    int pipeline;       /* default SD_PIPELINE_FUSED */
    int momentumZero;   /* default 1 (Sim.cu:4496-4501) */
    int exactGrid;      /* 0: round G down to a power of two and cap it as
                           Sim::reconfigure (Sim.cu:4081-4086, 4158);
                           1: accept G = 2^k or 3*2^k (<= 512 / 256) as given */
    int sortMerge;      /* re-sorts that start from the last sort's keys (only
                           the keys that changed are radix-sorted, then merged;
                           same permutation as a full sort): 0 (default) for
                           sorts of more than 2^23 keys, 1 always, -1 never */
} sd_modes;

/* Grids readable with sd_get_grid. */
#define SD_GRID_DENSITY_FIXED 0  /* int64[G^3], 2^24 per particle */
#define SD_GRID_DENSITY 1        /* float[G^3], particles per cell */
#define SD_GRID_POTENTIAL 2      /* float[G^3], Sim's pot on the G corner */
#define SD_GRID_ACCEL 3          /* float4[G^3], kGridAccel output */

/* ---- defaults ------------------------------------------------------------ */
void sd_config_default(sd_config* cfg);  /* SimConfig{} field defaults */
void sd_modes_default(sd_modes* m);

/* ---- device statics: Sim::deviceAvailable ... estimateStepMs (Sim.h:861-889) */
int sd_device_available(void);                          /* Sim::deviceAvailable */
int sd_device_name(char* buf, size_t len);              /* Sim::deviceName */
int sd_device_driver(char* buf, size_t len);            /* Sim::deviceDriver */
size_t sd_device_free_bytes(void);                      /* Sim::deviceFreeBytes */
int sd_device_multiprocessors(void);                    /* Sim::deviceMultiProcessors */
int sd_max_grid_size(int boundary);                     /* Sim::maxGridSize */
size_t sd_estimate_bytes(int particleCount, int gridSize, int boundary); /* Sim::estimateBytes */
int sd_max_particles_for(int gridSize, int boundary, size_t freeBytes);  /* Sim::maxParticlesFor */
double sd_device_bandwidth_gbs(void);                   /* Sim::deviceBandwidthGBs */
double sd_estimate_step_ms(int particleCount, int gridSize, int boundary); /* Sim::estimateStepMs */

/* ---- failure state: Sim::failed / Sim::failMessage (Sim.h:994-995) ------- */
int sd_failed(void);
const char* sd_fail_message(void);
void sd_clear_failure(void);  /* no reference counterpart */

/* ---- lifecycle: Sim::Sim/~Sim/init/reconfigure/reset/step (Sim.h:856-895) */
sd_sim* sd_create(void);
void sd_destroy(sd_sim* s);
int sd_set_modes(sd_sim* s, const sd_modes* m);  /* no reference counterpart */
int sd_get_modes(const sd_sim* s, sd_modes* m);  /* no reference counterpart */
int sd_init(sd_sim* s, const sd_config* cfg);
int sd_reconfigure(sd_sim* s, const sd_config* cfg);
int sd_reset(sd_sim* s);
int sd_step(sd_sim* s);
int sd_run(sd_sim* s, int steps);  /* no reference counterpart: k steps, one sync */
int sd_synchronize(sd_sim* s);     /* no reference counterpart */

/* ---- reads: Sim::config/timings/simTime/gridSize/particleCount/activeCount */
int sd_get_config(const sd_sim* s, sd_config* cfg);
int sd_timings_get(sd_sim* s, sd_timings* t);  /* phases of the last step (of the last sd_run) */
double sd_sim_time(sd_sim* s);
long long sd_step_count(sd_sim* s);  /* no reference counterpart */
int sd_grid_size(const sd_sim* s);
int sd_particle_count(const sd_sim* s);
int sd_active_count(sd_sim* s);

/* ---- measures used by the reference tests (Sim.h:1000-1029) ------------- */
double sd_measure_total_grid_mass(sd_sim* s);      /* measureTotalGridMass */
double sd_measure_max_density(sd_sim* s);          /* measureMaxDensity */
int sd_measure_occupied_cells(sd_sim* s);          /* measureOccupiedCells */
int sd_measure_centroid(sd_sim* s, double* cx, double* cy); /* measureCentroid */
double sd_measure_kinetic_energy(sd_sim* s);       /* measureKineticEnergy */
int sd_peak_cell_count(sd_sim* s);                 /* peakCellCount (Sim.h:912-915) */

/* ---- state files: Sim::saveState / loadState (Sim.cu:4985-5120), STARDUST v2 */
int sd_save_state(sd_sim* s, const char* path);
int sd_load_state(sd_sim* s, const char* path);

/* ---- host particle I/O (no reference counterpart; replaces loadState as the
 * way to inject initial conditions). pos4/vel4: float[4n], x < 0 = inactive.
 * n must not exceed the configured particle count; slots n.. are made inactive. */
int sd_set_particles(sd_sim* s, const float* pos4, const float* vel4, int n);
int sd_get_particles(sd_sim* s, float* pos4, float* vel4, int n);
int sd_set_fixed_dt(sd_sim* s, float dt);  /* dt > 0 replaces the CFL rule; 0 restores it */

/* ---- device views: particlePosDevicePtr/particleVelDevicePtr (Sim.h:1124-1128)
 * Invalidated by the next sort, reconfigure or step. Velocities are the raw
 * integrator state (see sd_get_particles for synchronised values). */
const float* sd_particle_pos_device_ptr(sd_sim* s);
const float* sd_particle_vel_device_ptr(sd_sim* s);

/* Density views: Sim::densityDevicePtr / Sim::fieldDevicePtr (Sim.h:1083-1084).
 * sd_density_device_ptr: float[G^3] ([z][y][x], row pitch G), particles per
 * cell of the current positions (the reference's rho is its S-pitched padded
 * grid with the density in the G corner). sd_field_device_ptr(SD_FIELD_DENSITY):
 * float[G^2] ([y][x]), the density summed along z (kProjectXY without a view
 * rotation); the other fields of Sim::Field come from off-path physics and
 * return NULL. Both recompute on each call, synchronise the handle's stream
 * before returning and are invalidated by the next step or call. */
#define SD_FIELD_DENSITY 0
#define SD_FIELD_DISPERSION 1
#define SD_FIELD_LIGHT 2
#define SD_FIELD_ASH 3
const float* sd_density_device_ptr(sd_sim* s);
const float* sd_field_device_ptr(sd_sim* s, int field);

/* ---- parity dumps (no reference counterpart) ----------------------------- */
int sd_get_grid(sd_sim* s, int which, void* out, size_t bytes);
int sd_compute_forces(sd_sim* s);  /* deposit + solve + gradient for current positions */
/* Sort now: keys + radix sort + reorder. keysOut / orderOut (optional, n
 * entries) receive the sorted keys (reference key + 1) and the permutation. */
int sd_sort_now(sd_sim* s, uint32_t* keysOut, int32_t* orderOut);
/* Mesh acceleration at each particle (CIC gather of the current grid). */
int sd_particle_accel(sd_sim* s, float* acc4, int n);

/* ---- force error (no reference counterpart; method of proto/cuda/x3.cu) -- */
/* Plummer direct sum onto `targets` (nt indices) from all live particles:
 * acc3[3*nt] = gm * sum_j (x_j - x_i)/(r^2 + eps^2)^(3/2), box units. */
int sd_direct_sum(sd_sim* s, const int32_t* targets, int nt, double gm, double eps, double* acc3);

/* Allocation-free variant of sd_direct_sum. No reference counterpart: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). targetsDev
 * (nt indices) and acc3Dev (3*nt doubles) are DEVICE pointers; the call is
 * enqueued on the handle's stream and returns without synchronising. The
 * FP64 partial sums live in a fixed scratch array of SD_DIRECT_SCRATCH_DOUBLES
 * doubles (16 MiB) that the library holds on each device it runs on until the
 * process ends; targets are processed in groups that fit it, so any nt is
 * accepted and the results equal sd_direct_sum's bit for bit. Sizes beyond the
 * scratch (more than 11 billion particles) return SD_ERR_ARG; a target index
 * outside [0, particleCount) yields NaN for that target. Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (other handles'
 * streams, other host threads) race. */
#define SD_DIRECT_SCRATCH_DOUBLES 2097152
int sd_direct_sum_noalloc(sd_sim* s, const int32_t* targetsDev, int nt, double gm, double eps, double* acc3Dev);

/* ---- learned sub-cell correction (no reference counterpart) --------------
 * MLP 30 -> 128 -> 128 -> 128 -> 3 (GELU, tanh form), BF16 weights, FP32
 * accumulation on the matrix cores. Inputs per particle (sd_features): log1p
 * of the particle counts at the 27 mesh nodes around its nearest node, then
 * 2 * (p G - 0.5 - node) per axis. Output: the direct-sum minus PM
 * acceleration divided by (a_unit * outScale), a_unit = gravity * G / N.
 * Parameters use the layout W1 [128 x 30] | W2 | W3 [128 x 128] | W4 [3 x 128]
 * (row-major out x in), b1 | b2 | b3 | b4. */
typedef struct sd_mlp sd_mlp;
int sd_mlp_param_counts(int* nW, int* nB);
sd_mlp* sd_mlp_create(unsigned seed);
void sd_mlp_destroy(sd_mlp* m);
int sd_mlp_set_params(sd_mlp* m, const float* W, const float* b);
int sd_mlp_get_params(sd_mlp* m, float* W, float* b);
int sd_mlp_forward(sd_mlp* m, const float* X, int n, float* Y);           /* X [n x 30], Y [n x 3] */
int sd_mlp_grad(sd_mlp* m, const float* X, const float* T, int n, double* loss, float* gW,
                float* gb);                                               /* loss 0.5/n sum |Y - T|^2 */
int sd_mlp_train(sd_mlp* m, const float* X, const float* T, int n, int epochs, int batch, float lr,
                 unsigned seed, float* lossPerEpoch);                     /* Adam, shuffled batches */
int sd_features(sd_sim* s, const int32_t* targets, int nt, float* X);    /* X [nt x 30] */
int sd_attach_mlp(sd_sim* s, sd_mlp* m, float outScale);  /* m = NULL detaches; steps add the correction */
/* Recompute the correction every `steps` steps (default 1); it is always
 * recomputed after a sort or sd_set_particles (the per-particle array follows
 * the particle order). */
int sd_set_mlp_interval(sd_sim* s, int steps);
int sd_mlp_correction(sd_sim* s, float* acc4, int n);      /* correction at current positions, box units */

/* ---- batched scenes (no reference counterpart: one scene per Sim) --------
 * Many independent simulations stepped together: every kernel of the fused
 * step covers all scenes in one launch (meshes grouped by size), and a run of
 * `steps` steps is recorded into one HIP graph and launched once. Each scene
 * is an ordinary sd_sim (sd_batch_scene) whose particles live in the batch's
 * arrays: set its particles / fixed dt through the sd_* calls, read it the
 * same way. A batched scene evolves bit-identically to the same scene stepped
 * alone with exactGrid = 1. Scenes must share the sort interval and step
 * count, use the fused pipeline and fixed-point deposit, and have no contact
 * or learned correction. The batch owns its scenes (do not sd_destroy them). */
typedef struct sd_batch sd_batch;
sd_batch* sd_batch_create(void);
void sd_batch_destroy(sd_batch* b);
int sd_batch_add_scene(sd_batch* b, const sd_config* cfg);  /* scene index, or < 0 */
int sd_batch_init(sd_batch* b);                             /* allocates every scene */
int sd_batch_scene_count(const sd_batch* b);
sd_sim* sd_batch_scene(sd_batch* b, int i);
long long sd_batch_total_particles(const sd_batch* b);
/* Enqueue `steps` steps of every scene as one graph launch (recorded on first
 * use, replayed while steps, snapshotEvery and the scenes' configuration are
 * unchanged); positions are copied to snapshot k after step (k+1)*snapshotEvery
 * (0 = none). Asynchronous; scene calls are ordered after it. */
int sd_batch_run(sd_batch* b, int steps, int snapshotEvery);
int sd_batch_synchronize(sd_batch* b);
float sd_batch_last_run_ms(sd_batch* b);                    /* device time of the last run */
long long sd_batch_graph_nodes(const sd_batch* b);          /* nodes of the recorded graph */
int sd_batch_snapshot_count(const sd_batch* b);
int sd_batch_get_snapshot(sd_batch* b, int snap, int scene, float* pos4, int n);

/* ---- kernel timing (no reference counterpart) ----------------------------
 * Brackets every launch of one named kernel with HIP events on the handle's
 * stream; sd_kernel_time_ms returns total milliseconds and launch count. */
int sd_kernel_timer_enable(sd_sim* s, const char* kernelName);
/* Bracket only one launch in `every` (default 1): each bracket is a pair of
 * queue markers, so sampling keeps the timed run close to an untimed one. */
int sd_kernel_timer_every(sd_sim* s, int every);
int sd_kernel_time_ms(sd_sim* s, double* totalMs, long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* STARDUST_SD_API_H */
