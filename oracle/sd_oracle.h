/*
 * sd_oracle.h -- CPU oracle for the Stardust particle-mesh gravity step.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference algorithm
 * (src/sim/Sim.cu of chdnl0420-svg/stardust) on the CPU so that tests, the
 * smoke check and the CPU-baseline leg of bench.py can check the HIP path
 * against it. Nothing in the product links, loads or calls it.
 *
 * Conventions shared with the product:
 *   - particles are float4 arrays (x, y, z, w) in box units [0,1); x < 0 marks
 *     an inactive slot (Sim.cu:507, 692);
 *   - cell keys are unsigned "reference key + 1": inactive -> 0, active cell
 *     c -> c + 1, so an unsigned ascending sort reproduces CUB's signed order
 *     with the -1 keys first (Sim.cu:2482-2485, 3660-3667);
 *   - the fixed-point density holds 2^24 per particle (SD_FIX_ONE).
 *
 * Parity status of every function is listed in its comment in sd_oracle.c.
 */
#ifndef SD_ORACLE_H
#define SD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SD_FIX_SHIFT 24
#define SD_FIX_ONE (1LL << SD_FIX_SHIFT)

/* key modes */
#define SDO_KEY_ROWMAJOR 0
#define SDO_KEY_MORTON 1
#define SDO_KEY_MORTON_NODE 2

/* --- keys, sort, reorder ------------------------------------------------ */
uint32_t sdo_morton3(uint32_t x, uint32_t y, uint32_t z);
void sdo_cell_keys(const float* pos4, int n, int G, int mode, uint32_t* keys, int32_t* idx);
void sdo_stable_sort_pairs(uint32_t* keys, int32_t* vals, int n);
void sdo_reorder4(const float* src4, const int32_t* order, int n, float* dst4);

/* --- deposit ------------------------------------------------------------- */
void sdo_deposit_float(const float* pos4, int n, int G, int S, int periodic, float* rho);
void sdo_deposit_fixed(const float* pos4, int n, int G, int periodic, int64_t* rho);
void sdo_cic_weights_fixed(float px, float py, float pz, int G, int* base3, int64_t* w8);

/* --- FFT / Poisson ------------------------------------------------------- */
void sdo_fft1d(double* re, double* im, int n, int inverse);
void sdo_fft3d(double* re, double* im, int nx, int ny, int nz, int inverse);
void sdo_green_real(int S, float soft, double* g);
void sdo_poisson_iso(const double* rho, int G, float softCells, double* pot);
void sdo_poisson_periodic(const double* rho, int G, float softCells, double* pot);
void sdo_fft3d_f32(float* re, float* im, int nx, int ny, int nz, int inverse);
void sdo_poisson_iso_f32(const float* rho, int G, float softCells, float* pot);
void sdo_poisson_periodic_f32(const float* rho, int G, float softCells, float* pot);
void sdo_grid_accel_f32(const float* pot, int G, float potScale, int periodic, float* acc4);
void sdo_grid_accel(const double* pot, int G, float potScale, int periodic, float* acc4);

/* --- particles ----------------------------------------------------------- */
void sdo_sample_acc(const float* acc4, float px, float py, float pz, int G, int periodic,
                    float* out3);
void sdo_sample_acc_all(const float* acc4, const float* pos4, int n, int G, int periodic, float* out4);
float sdo_max_speed(const float* pos4, const float* vel4, int n);
void sdo_integrate(const float* acc4, float* pos4, float* vel4, int n, int G, float kick,
                   float drift, int periodic, float c2, const float* accContact4,
                   float softBoundR, float darkEnergy, const float* meanSub3, const float* accPress4);
void sdo_momentum(const float* pos4, const float* vel4, int n, double* out4);
void sdo_subtract_mean(const float* pos4, float* vel4, int n, const double* acc4);

/* --- contact hash grid --------------------------------------------------- */
void sdo_cell_range(const uint32_t* keys, int n, int ncells, int32_t* start, int32_t* end);
void sdo_contact(const float* pos4, const float* vel4, int n, int G, int periodic,
                 int keyMode, const int32_t* cellStart, const int32_t* cellEnd, float radius,
                 float stiffness, float damping, float* accOut4);

/* --- whole reference step ------------------------------------------------ */
typedef struct sdo_params {
    int G;
    int periodic;
    float gravity;
    float softeningCells;
    float timeScale;
    int sortInterval;
    int keyMode;
    int fixedDeposit;   /* 1: int64 fixed-point deposit, 0: float adds in particle order */
    float fixedDt;      /* > 0 overrides the CFL policy */
    int momentumZero;   /* per-step momentum removal (Sim.cu:4496-4501) */
    int contact;
    float contactRadiusCells;
    float contactStiffness;
    float contactDamping;
    float c2;           /* speed-of-light squared clamp */
    float softBoundR;
    float darkEnergy;
    int allocN;         /* mass per particle = gravity / allocN */
    int fp32Mesh;       /* 1: single-precision mesh (the CPU baseline's timing path) */
    int pressureEnabled; /* gas pressure (Sim.cu:1368-1399, 3617-3631) */
    int coolingEnabled;  /* gas cooling (Sim.cu:1212-1366, 3732-3830) */
    float pressureK;
    float coolingRate;
} sdo_params;

typedef struct sdo_state {
    long long stepCount;
    double simTime;
    float lastDt;
    float lastVmax;
} sdo_state;

float sdo_dt_policy(const sdo_params* prm, float vmax);
void sdo_compute_accel(const sdo_params* prm, const float* pos4, int n, float* accG4,
                       int64_t* rhoFixOut, double* potOut);
int sdo_step(const sdo_params* prm, sdo_state* st, float* pos4, float* vel4, int n,
             float* scratchPos4, float* scratchVel4, float* gas4);
void sdo_cool_cells(const float* pos4, float* vel4, const uint32_t* keys, int n, int G, float rate, float dt,
                    float* disp4);
void sdo_pressure(const float* disp4, int G, int periodic, float k, float maxAcc, float* accP4);

/* --- direct sum and error metric ---------------------------------------- */
void sdo_direct_sum(const float* pos4, int n, const int32_t* targets, int nt, double gm,
                    double eps, double* acc3);
double sdo_rel_rms(const double* approx3, const double* truth3, int n, double* lsqScale);

/* --- MLP 30 -> H -> H -> H -> 3, GELU (tanh form) ------------------------ */
void sdo_mlp_forward(const double* W, const double* b, int H, const double* x, int nb,
                     double* y);
double sdo_mlp_loss_grad(const double* W, const double* b, int H, const double* x,
                         const double* target, int nb, double* gW, double* gb);
size_t sdo_mlp_wcount(int H);
size_t sdo_mlp_bcount(int H);

#ifdef __cplusplus
}
#endif
#endif
