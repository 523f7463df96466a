// sd_api.cpp -- host side of the C ABI (include/stardust/sd_api.h).
//
// Plays the role of Sim::Impl in src/sim/Sim.cu: owns every device buffer of
// one simulation, sequences the kernels of a step on the handle's stream and
// keeps the per-step scalars (dt, kick, momentum, time) on the device so a
// step has no host round trip. Error handling follows the reference's sticky
// g_failed / g_failMsg (Sim.cu:52-75).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/stardust/sd_api.h"
#include "sd_internal.h"
#include "sd_mlp.h"
#include "sd_sim.h"

using namespace sd;

// ---------------------------------------------------------------------------
// sticky failure state (Sim.cu:52-75)
// ---------------------------------------------------------------------------
static bool g_failed = false;
static char g_failMsg[256] = {0};

void sd_mark_failure(const char* fmt, ...) {
    if (g_failed) return;
    g_failed = true;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_failMsg, sizeof(g_failMsg), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[stardust_amd] failure: %s\n", g_failMsg);
}

bool sd_failure_set() { return g_failed; }

int sd_arg_error(const char* what) {
    // argument errors are reported but are not sticky
    snprintf(g_failMsg, sizeof(g_failMsg), "%s", what);
    fprintf(stderr, "[stardust_amd] invalid argument: %s\n", what);
    return SD_ERR_ARG;
}

#define mark_failure sd_mark_failure
#define arg_error sd_arg_error

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

static float light_speed_sq(float lengthScale, float timeUnitScale) {
    // lightSpeedFor, Sim.h:131-148: c = (1e7 T) / (1e5 L)
    const double L = 1.0e5 * (double)(lengthScale > 1e-6f ? lengthScale : 1e-6f);
    const double T = 1.0e7 * (double)(timeUnitScale > 1e-6f ? timeUnitScale : 1e-6f);
    const float c = (float)(T / (L > 1e-30 ? L : 1e-30));
    return c * c;
}

template <typename T>
static void dfree(T*& p) {
    if (p) {
        (void)hipFree((void*)p);
        p = nullptr;
    }
}

template <typename T>
static bool dalloc(T*& p, size_t count) {
    dfree(p);
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * count);
    if (e != hipSuccess) {
        mark_failure("hipMalloc(%zu bytes): %s", sizeof(T) * count, hipGetErrorString(e));
        p = nullptr;
        return false;
    }
    return true;
}

static void timer_begin(sd_sim* s, const char* name, hipStream_t st) {
    KernelTimer& t = s->timer;
    t.armed = false;
    if (t.name.empty() || t.name != name) return;
    if ((t.seen++ % t.every) != 0) return;
    t.armed = true;
    if (t.used + 2 > (int)t.pool.size()) {
        // harvest what is pending to recycle events
        HK(hipStreamSynchronize(s->st));
        for (int k = 0; k + 1 < t.used; k += 2) {
            float ms = 0.f;
            HK(hipEventElapsedTime(&ms, t.pool[k], t.pool[k + 1]));
            t.totalMs += ms;
            t.launches++;
        }
        t.used = 0;
    }
    HK(hipEventRecord(t.pool[t.used], st));
}
static void timer_end(sd_sim* s, const char* name, hipStream_t st) {
    KernelTimer& t = s->timer;
    if (!t.armed || t.name.empty() || t.name != name) return;
    t.armed = false;
    HK(hipEventRecord(t.pool[t.used + 1], st));
    t.used += 2;
}
static void timer_harvest(sd_sim* s) {
    KernelTimer& t = s->timer;
    if (t.used == 0) return;
    HK(hipStreamSynchronize(s->st));
    for (int k = 0; k + 1 < t.used; k += 2) {
        float ms = 0.f;
        HK(hipEventElapsedTime(&ms, t.pool[k], t.pool[k + 1]));
        t.totalMs += ms;
        t.launches++;
    }
    t.used = 0;
}

// A timed launch that records its own start / stop events (hipExtLaunchKernel):
// hands out the event pair when this launch of `name` is sampled.
static bool timer_arm(sd_sim* s, const char* name, hipEvent_t& e0, hipEvent_t& e1) {
    KernelTimer& t = s->timer;
    t.armed = false;
    if (t.name.empty() || t.name != name) return false;
    if ((t.seen++ % t.every) != 0) return false;
    if (t.used + 2 > (int)t.pool.size()) timer_harvest(s);
    e0 = t.pool[t.used];
    e1 = t.pool[t.used + 1];
    t.used += 2;
    return true;
}

// Work of one timed phase on stream st (the library's stream, or the mesh
// stream of an overlapped step, which always joins the library's stream).
#define TIMED_ON(s, st, name, stmt)  \
    do {                             \
        timer_begin(s, name, st);    \
        stmt;                        \
        timer_end(s, name, st);      \
    } while (0)
#define TIMED(s, name, stmt) TIMED_ON(s, (s)->st, name, stmt)

// Zero the fixed-point density grid (and note it is clean).
static void fill_rho(sd_sim* s) {
    launch_fill_i64(s->rhoFix, (long long)s->G * s->G * s->G, s->st);
    s->rhoDirty = false;
}
// Atomic deposits add into rhoFix and need it zero: a grid an x pass left
// unconsumed for a gather deposit is cleared first.
static void clean_rho(sd_sim* s) {
    if (s->rhoDirty) fill_rho(s);
}

static void free_buffers(sd_sim* s) {
    if (!s->borrowed) {
        dfree(s->pos); dfree(s->vel); dfree(s->posTmp); dfree(s->velTmp);
    }
    dfree(s->keysA); dfree(s->keysB); dfree(s->idxA); dfree(s->idxB);
    dfree(s->mb.mKeys); dfree(s->mb.mKeysAlt); dfree(s->mb.sKeys); dfree(s->mb.mIdx); dfree(s->mb.mIdxAlt);
    dfree(s->mb.sIdx); dfree(s->mb.tileMovers); dfree(s->mb.counts);
    s->keysAscending = false;
    dfree(s->hist); dfree(s->chunk);
    dfree(s->rhoFix); dfree(s->rhoF); dfree(s->rhoX); dfree(s->gfix.fixCount); dfree(s->gfix.tileFlag);
    dfree(s->gfix.tileList); dfree(s->proj); dfree(s->A); dfree(s->B); dfree(s->green);
    dfree(s->pot); dfree(s->accG); dfree(s->tw); dfree(s->accContact); dfree(s->disp); dfree(s->accPress);
    dfree(s->cellRange); dfree(s->blockRange); dfree(s->partials); dfree(s->sc); dfree(s->accCorr); dfree(s->logRho);
    s->cellTableCells = 0;
    s->blockCells = 0;
    if (s->scHost) {
        (void)hipHostFree(s->scHost);
        s->scHost = nullptr;
    }
    s->ready = false;
}

static int valid_grid(int G, int boundary, int exact) {
    if (!exact) return 1;
    if (boundary == SD_BOUNDARY_PERIODIC) return is_pow2(G) && fft_length_supported(G);
    return fft_length_supported(2 * G);
}

// Reference grid policy (Sim.cu:4081-4086, 4158), or exact sizes.
static int choose_grid(const sd_config* c, const sd_modes* m) {
    if (m->exactGrid) return c->gridSize;
    const int cap = sd_max_grid_size(c->boundary);
    int g = c->gridSize > 0 ? c->gridSize : 64;
    if (g > cap) g = cap;
    if (g < 16) g = 16;
    int p = 16;
    while (p * 2 <= g) p *= 2;
    return p;
}

// Pressure and cooling (the gas dispersion grids) run in the reference kernel
// order with kick-drift, where the stored velocities are the physical ones.
static bool gas_physics(const sd_config* c) { return c->pressureEnabled || c->coolingEnabled; }
static int check_gas_modes(const sd_config* c, const sd_modes* m) {
    if (gas_physics(c) && (m->pipeline != SD_PIPELINE_REFERENCE || m->integrator != SD_INTEGRATOR_KD))
        return arg_error("pressure and cooling run in the reference kernel order: they need "
                         "SD_PIPELINE_REFERENCE and SD_INTEGRATOR_KD");
    return SD_OK;
}

static int check_config(const sd_config* c) {
    if (!c) return arg_error("null config");
    if (c->boundary != SD_BOUNDARY_ISOLATED && c->boundary != SD_BOUNDARY_PERIODIC)
        return arg_error("boundary must be isolated (0) or periodic (1)");
    if (c->starFormationEnabled || c->blackHoleEnabled || c->strongForceEnabled || c->emForceEnabled ||
        c->weakForceEnabled || c->collapseEnabled)
        return arg_error("star formation, black holes, strong/EM/weak forces and collapse are not provided "
                         "(the PM path with contact, pressure and cooling is)");
    if (c->preset != SD_PRESET_EMPTY)
        return arg_error("preset Filament is not provided; use SD_PRESET_EMPTY and sd_set_particles");
    if (c->particleCount < 0) return arg_error("particleCount < 0");
    return SD_OK;
}

static void upload_twiddles(sd_sim* s) {
    std::vector<float2> h(s->S);
    for (int m = 0; m < s->S; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)s->S;
        h[m] = make_float2((float)cos(a), (float)sin(a));
    }
    HK(hipMemcpy(s->tw, h.data(), sizeof(float2) * s->S, hipMemcpyHostToDevice));
}

static void reset_scalars(sd_sim* s) {
    StepScalars z{};
    z.kdkOpen = 1;
    HK(hipMemcpyAsync(s->sc, &z, sizeof(z), hipMemcpyHostToDevice, s->st));
    HK(hipStreamSynchronize(s->st));
    s->primed = false;
    s->stepCount = 0;
}

static void build_green(sd_sim* s) {
    if (s->periodic) return;
    const int S = s->S;
    float2* ta = nullptr;
    float2* tb = nullptr;
    const size_t n = (size_t)S * S * s->P;
    if (!dalloc(ta, n) || !dalloc(tb, n)) {
        dfree(ta);
        dfree(tb);
        return;
    }
    // kGreen uses soft = max(softeningCells, 0.5), Sim.cu:3572
    const float soft = fmaxf(s->cfg.softeningCells, 0.5f);
    mesh_build_green(s->plan, soft, ta, tb, s->green, s->st);
    HK(hipGetLastError());
    HK(hipStreamSynchronize(s->st));
    dfree(ta);
    dfree(tb);
}

static int allocate(sd_sim* s) {
    free_buffers(s);
    const int G = s->G, S = s->S, P = s->P;
    const size_t N = (size_t)(s->allocN > 0 ? s->allocN : 1);
    const size_t cells = (size_t)G * G * G;
    bool ok = (s->borrowed || (dalloc(s->pos, N) && dalloc(s->vel, N) && dalloc(s->posTmp, N) &&
                               dalloc(s->velTmp, N))) &&
              dalloc(s->keysA, N) && dalloc(s->keysB, N) && dalloc(s->idxA, N) && dalloc(s->idxB, N);
    const int ntiles = sort_tiles_for((int)N);
    ok = ok && dalloc(s->hist, (size_t)256 * ntiles) && dalloc(s->chunk, (size_t)sort_chunks_for((int)N) + 1);
    ok = ok && dalloc(s->rhoFix, cells) && dalloc(s->A, (size_t)G * G * P) && dalloc(s->B, (size_t)G * S * P) &&
         dalloc(s->pot, cells) && dalloc(s->accG, cells) && dalloc(s->tw, (size_t)S) &&
         dalloc(s->proj, (size_t)G * G);
    if (!s->periodic) ok = ok && dalloc(s->green, (size_t)(S / 2 + 1) * (S / 2 + 1) * P);
    s->nparts = particle_blocks((int)N);
    s->npartsInt = integrate_blocks((int)N);
    ok = ok && dalloc(s->partials, (size_t)s->nparts) && dalloc(s->sc, 1);
    if (ok && hipHostMalloc((void**)&s->scHost, sizeof(StepScalars), hipHostMallocDefault) != hipSuccess) {
        mark_failure("hipHostMalloc failed");
        ok = false;
    }
    if (!ok) {
        free_buffers(s);
        return SD_ERR_DEVICE;
    }
    s->sb.keys = s->keysA;
    s->sb.idx = s->idxA;
    s->sb.keysAlt = s->keysB;
    s->sb.idxAlt = s->idxB;
    s->sb.hist = s->hist;
    s->sb.chunk = s->chunk;
    s->sb.maxTiles = ntiles;
    upload_twiddles(s);
    s->plan.G = G;
    s->plan.S = S;
    s->plan.P = P;
    s->plan.periodic = s->periodic;
    s->plan.tw = s->tw;
    // kPoissonPeriodic scale (Sim.cu:3594): norm / (4 pi^2) with norm = 1/S^3
    const float norm = 1.0f / (float)((double)S * S * S);
    s->plan.perScale = -(norm / (4.0f * 3.14159265f * 3.14159265f));
    {
        const float sg = 6.2831853f * s->cfg.softeningCells / (float)G;
        s->plan.perSoft = s->cfg.softeningCells > 0.f ? sg * sg : 0.f;
    }
    fill_rho(s);
    build_green(s);
    if (g_failed) return SD_ERR_DEVICE;
    // all slots inactive
    std::vector<float4> hp(N, make_float4(-1.f, -1.f, -1.f, 0.f)), hv(N, make_float4(0.f, 0.f, 0.f, 0.f));
    HK(hipMemcpy(s->pos, hp.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
    HK(hipMemcpy(s->vel, hv.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
    s->active = 0;
    s->rhoValid = false;
    s->meshValid = false;
    s->keysSorted = false;
    reset_scalars(s);
    s->ready = !g_failed;
    return g_failed ? SD_ERR_DEVICE : SD_OK;
}

// Cell-table size: G^3 for row-major keys; Morton keys of a 3*2^k mesh run
// up to the next power of two cubed.
static long long cell_table_size(const sd_sim* s) {
    long long g = s->G;
    if (s->modes.keyMode == SD_KEY_MORTON) {
        g = 1;
        while (g < s->G) g <<= 1;
    }
    return g * g * g;
}

// Row-major keys: one table entry per cell. Morton keys: one per 4^3-cell
// block (a cell's particles are found inside its block's run of sorted keys).
static bool ensure_contact_buffers(sd_sim* s) {
    const size_t cells = (size_t)cell_table_size(s);
    if (!s->accContact && !dalloc(s->accContact, (size_t)s->allocN)) return false;
    if (s->modes.keyMode == SD_KEY_MORTON) {
        // Morton cell codes run over (2^k)^3 >= G^3; a block is 64 consecutive codes
        const long long blocks = (long long)(cells + 63) / 64;
        if (s->blockCells < blocks) dfree(s->blockRange);
        s->blockCells = blocks;
        // the block runs, then the blocks' cell-occupancy words
        if (!s->blockRange && !dalloc(s->blockRange, 2 * (size_t)blocks)) return false;
        return true;
    }
    if (s->cellTableCells < (long long)cells) dfree(s->cellRange);
    s->cellTableCells = (long long)cells;
    if (!s->cellRange && !dalloc(s->cellRange, cells)) return false;
    return true;
}

// Gas dispersion grids (zero until the first cooling pass) and the pressure grid.
static bool ensure_gas_buffers(sd_sim* s) {
    if (s->disp && s->accPress) return true;
    const long long cells = (long long)s->G * s->G * s->G;
    if (!dalloc(s->disp, (size_t)cells * 4) || !dalloc(s->accPress, (size_t)cells)) return false;
    launch_fill_f32(s->disp, cells * 4, 0.f, s->st);
    return true;
}

static bool ensure_float_rho(sd_sim* s) {
    if (s->rhoF) return true;
    return dalloc(s->rhoF, (size_t)s->G * s->G * s->G);
}

static int key_bits(const sd_sim* s) {
    // keys are cell + 1 <= G^3 (row-major) or morton + 1 < 2^(3 ceil(log2 G)) + 1
    unsigned long long maxKey;
    if (s->modes.keyMode == SD_KEY_MORTON) {
        int b = 0;
        while ((1 << b) < s->G) ++b;
        maxKey = (1ull << (3 * b));
    } else {
        maxKey = (unsigned long long)s->G * s->G * s->G;
    }
    int bits = 0;
    while ((1ull << bits) <= maxKey) ++bits;
    return bits;
}

#ifndef SD_FUSE_REORDER
#define SD_FUSE_REORDER 1
#endif

// Keys and the stable radix sort only: the permutation is left in s->sb.idx
// for an integrate launch that reads through it (the fused reorder).
// The sort of nearly sorted keys, when the last sort left its ascending keys
// in sb.keys: between step sorts few particles change cell, and only those
// keys are radix-sorted before the merge (exact for any ascending sb.keys).
#ifndef SD_SORT_MERGE
#define SD_SORT_MERGE 1
#endif
#ifndef SD_SORT_MERGE_MIN_N
#define SD_SORT_MERGE_MIN_N (1 << 23)
#endif
// (An allocation that fails here leaves the full sort in use; it is not a
// device failure.)
template <typename T>
static bool try_alloc(T*& p, size_t count) {
    return hipMalloc((void**)&p, sizeof(T) * count) == hipSuccess;
}
static bool ensure_mover_buffers(sd_sim* s) {
    if (s->mb.counts) return true;
    const size_t N = (size_t)(s->allocN > 0 ? s->allocN : 1);
    MoverBuffers& m = s->mb;
    const bool ok = try_alloc(m.mKeys, N) && try_alloc(m.mKeysAlt, N) && try_alloc(m.sKeys, N) &&
                    try_alloc(m.mIdx, N) && try_alloc(m.mIdxAlt, N) && try_alloc(m.sIdx, N) &&
                    try_alloc(m.tileMovers, (size_t)mover_tiles_for((int)N)) && try_alloc(m.counts, 2);
    if (!ok) {
        (void)hipGetLastError();
        dfree(m.mKeys); dfree(m.mKeysAlt); dfree(m.sKeys); dfree(m.mIdx); dfree(m.mIdxAlt); dfree(m.sIdx);
        dfree(m.tileMovers); dfree(m.counts);
        return false;
    }
    return true;
}

// Cell keys of the current positions, stably sorted into sb.keys / sb.idx.
static void sort_cell_keys(sd_sim* s, int kmode, hipStream_t st) {
    const int n = s->allocN;
    // automatic: large sorts only (below, the three full radix passes take
    // about as long as the changed keys' radix passes plus the merge)
    const int pol = s->modes.sortMerge;
    const bool want = pol > 0 || (pol == 0 && n > SD_SORT_MERGE_MIN_N);
    if (SD_SORT_MERGE && want && s->keysAscending && n > 1 && ensure_mover_buffers(s)) {
        sort_nearly_sorted(s->sb, s->mb, s->pos, n, s->G, kmode, key_bits(s), st);
    } else {
        launch_cell_keys_hist(s->pos, n, s->G, kmode, s->sb, st);
        radix_sort_pairs(s->sb, n, key_bits(s), st, true);
    }
    s->keysAscending = true;
}

static void sort_keys(sd_sim* s, bool forContact = false) {
    const int kmode = s->modes.keyMode == SD_KEY_MORTON ? (forContact ? 1 : 2) : 0;
    TIMED(s, "sort", sort_cell_keys(s, kmode, s->st));
    s->keysSorted = true;
    s->corrValid = false;
}

// Impl::sortParticles, Sim.cu:3656-3671. forContact selects the contact
// cell key (Morton of the cell) instead of the stencil-node key.
static void sort_particles(sd_sim* s, bool forContact = false) {
    const int n = s->allocN;
    const int kmode = s->modes.keyMode == SD_KEY_MORTON ? (forContact ? 1 : 2) : 0;
    TIMED(s, "sort", {
        sort_cell_keys(s, kmode, s->st);
        launch_reorder(s->pos, s->vel, s->sb.idx, s->posTmp, s->velTmp, n, s->st);
    });
    if (s->borrowed) {
        // a batch owns the arrays: keep their addresses, copy the result back
        launch_copy_f4(s->posTmp, s->pos, n, s->st);
        launch_copy_f4(s->velTmp, s->vel, n, s->st);
    } else {
        std::swap(s->pos, s->posTmp);
        std::swap(s->vel, s->velTmp);
    }
    s->keysSorted = true;
    s->corrValid = false;  // per-particle correction no longer in particle order
}

// The reference's doCooling (Sim.cu:3732-3830), deterministic: every 1 (contact
// on), 2 (pressure on) or 8 steps, with dt times that interval, particles
// sorted by cell (the contact pass already did; else a cell-key sort here),
// dispersion sums cleared, then one pass per cell.
static void cooling_pass(sd_sim* s, bool contactSorted) {
    const int every = s->cfg.contactEnabled ? 1 : (s->cfg.pressureEnabled ? 2 : 8);
    if ((s->stepCount % every) != 0) return;
    if (!ensure_gas_buffers(s)) return;
    if (!contactSorted) sort_particles(s, true);
    const long long cells = (long long)s->G * s->G * s->G;
    launch_fill_f32(s->disp, cells * 4, 0.f, s->st);
    const float rate = s->cfg.coolingEnabled && s->cfg.coolingRate > 0.f ? s->cfg.coolingRate : 0.f;
    TIMED(s, "cooling", launch_cool_cells(s->pos, s->vel, s->sb.keys, s->allocN, s->G, rate, (float)every, s->sc,
                                          s->disp, s->st));
}

static DtPolicy dt_policy(const sd_sim* s) {
    DtPolicy p{};
    p.fixedDt = s->fixedDt;
    const float ts = s->cfg.timeScale;
    // dtScaleFor, Sim.h:89-93 (kMaxStepReps = 3)
    const float scale = ts <= 1.0f ? ts : (ts <= 3.0f ? 1.0f : ts / 3.0f);
    p.baseDt = 0.0016f * scale;
    p.cell = 1.0f / (float)s->G;
    p.dtStiff = s->cfg.contactEnabled ? 0.5f / sqrtf(fmaxf(s->cfg.contactStiffness, 1.0f)) : 0.f;
    p.leapfrog = s->modes.integrator == SD_INTEGRATOR_KDK ? 1 : 0;
    return p;
}

// deposit of the current positions into rhoFix (or rhoF in float mode)
static void deposit_now(sd_sim* s) {
    const long long cells = (long long)s->G * s->G * s->G;
    if (s->modes.depositMode == SD_DEPOSIT_FLOAT) {
        if (!ensure_float_rho(s)) return;
        launch_fill_f32(s->rhoF, cells, 0.f, s->st);
        TIMED(s, "deposit", launch_deposit_float(s->pos, s->allocN, s->G, s->periodic, s->rhoF, s->st));
    } else {
        // rhoFix is zero here (allocated zero, or cleared by the mesh pass)
        clean_rho(s);
        TIMED(s, "deposit", launch_deposit_fixed(s->pos, s->allocN, s->G, s->periodic, s->rhoFix, s->st));
        s->rhoValid = true;
        s->rhoFloatOnly = false;
    }
}

// rhoFix holds the deposit of the current positions
static bool rho_fix_ready(const sd_sim* s) {
    return s->rhoValid && !s->rhoFloatOnly && s->modes.depositMode == SD_DEPOSIT_FIXED;
}

// X pass for the float deposit: convert rhoF into the int64 staging format is
// not needed -- the float grid is read through the fixed-point path by scaling.
// (Float mode deposits into rhoF; rows_fwd reads rhoFix, so float mode copies.)
__global__ void k_float_to_fix(const float* __restrict__ f, long long* __restrict__ q, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = (long long)llrintf(f[i] * (float)SD_FIX_ONE);
}

// solvePoisson + kGridAccel (Sim.cu:3586-3611): consumes (and clears) rhoFix.
// Density -> potential -> acceleration grid (and a deferred finalize) on st:
// the library's stream, or the mesh stream of an overlapped step.
static float pot_scale(const sd_sim* s) { return s->cfg.gravity / (float)(s->allocN > 0 ? s->allocN : 1); }

// kGridAccel: the acceleration grid from the potential
static void grad_now(sd_sim* s, hipStream_t st) {
    TIMED_ON(s, st, "grad", mesh_grid_accel(s->pot, s->accG, s->G, pot_scale(s), s->periodic, st));
    s->accGValid = true;
}

static bool gather_deposit(const sd_sim* s);

// The fused step's integrate differentiates the potential in its gather, so
// it solves without the gradient (grad = false); the acceleration grid is
// then made only when something reads it (ensure_mesh).
// keepRho: leave the density in place (a learned correction reads it beside
// the solve); the next atomic deposit clears it first.
static void solve_mesh(sd_sim* s, hipStream_t st, bool grad, bool keepRho = false) {
    const long long cells = (long long)s->G * s->G * s->G;
    if (s->modes.depositMode == SD_DEPOSIT_FLOAT) {
        hipLaunchKernelGGL(k_float_to_fix, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, s->rhoF,
                           s->rhoFix, cells);
    }
    const FinalizeJob* fin = s->finDeferred ? &s->finJob : nullptr;
    if (s->rhoValid && s->rhoFloatOnly) {
        // a gather deposit's float nodes (rhoFix is left as it is)
        TIMED_ON(s, st, "fft_x", mesh_rows_fwd_f32(s->plan, s->rhoX, s->A, fin, st));
    } else {
        // The x pass clears the density it reads for the atomic deposit of the
        // next integrate; a gather deposit writes every node, so it needs no clear.
        const bool keep = keepRho || gather_deposit(s);
        TIMED_ON(s, st, "fft_x", mesh_rows_fwd(s->plan, s->rhoFix, s->A, keep ? 0 : 1, fin, st));
        s->rhoDirty = keep;
    }
    s->rhoFloatOnly = false;
    TIMED_ON(s, st, "fft_yzy", mesh_solve_rest(s->plan, s->A, s->B, s->green, s->pot, fin, st));
    s->finDeferred = false;
    if (grad) grad_now(s, st);
    s->accGValid = grad;
    s->rhoValid = false;
    s->meshValid = true;
}

// Invariant (fixed-point mode): rhoFix is all zero whenever !rhoValid.
static void ensure_rho(sd_sim* s) {
    if (s->modes.depositMode == SD_DEPOSIT_FLOAT || !s->rhoValid) deposit_now(s);
}
// The same, in rhoFix (also when the valid deposit is a gather deposit's float grid)
static void ensure_rho_fix(sd_sim* s) {
    if (!rho_fix_ready(s)) deposit_now(s);
}

static void ensure_mesh(sd_sim* s, bool grad = true) {
    if (s->meshValid) {
        if (grad && !s->accGValid) grad_now(s, s->st);
        return;
    }
    ensure_rho(s);
    solve_mesh(s, s->st, grad);
}

// doContact, Sim.cu:3673-3721 (contact term; gate N <= 0.764 G^3, Sim.cu:3685),
// in two parts: the cell-key sort, then the forces. keysOnly: sort the keys
// only and read the particles through the permutation (the integrate that
// follows does the reorder itself); the forces are the same.
static bool contact_due(const sd_sim* s) {
    const int G = s->G;
    return s->cfg.contactEnabled && (double)s->allocN <= 0.764 * (double)G * G * G;
}

// Contact steps at under SD_GATHER_DENSITY particles per cell deposit in
// gather form (launch_deposit_gather): the contact sort's Morton block runs
// locate each tile's particles, and every node is written once.
#ifndef SD_GATHER_DEPOSIT
#define SD_GATHER_DEPOSIT 1
#endif
#ifndef SD_GATHER_DENSITY
#define SD_GATHER_DENSITY 0.5
#endif
static bool gather_deposit(const sd_sim* s) {
    return SD_GATHER_DEPOSIT && contact_due(s) && s->modes.keyMode == SD_KEY_MORTON &&
           s->modes.pipeline == SD_PIPELINE_FUSED && s->modes.depositMode == SD_DEPOSIT_FIXED && !s->borrowed &&
           gather_deposit_supported(s->G, s->periodic) && (double)s->allocN < SD_GATHER_DENSITY * (double)s->G * s->G * s->G;
}

// A gather deposit writes float nodes (and the x pass reads 4 bytes a node
// instead of 8) wherever the x pass is the staged one (S != 256).
#ifndef SD_GATHER_FLOAT
#define SD_GATHER_FLOAT 1
#endif
static bool gather_float(const sd_sim* s) { return SD_GATHER_FLOAT && gather_deposit(s) && s->plan.S != 256; }

static bool ensure_gather_float_buffers(sd_sim* s) {
    if (s->rhoX) return true;
    const int tiles = gather_deposit_tiles(s->G);
    if (!dalloc(s->gfix.fixCount, 2) || !dalloc(s->gfix.tileFlag, (size_t)tiles) ||
        !dalloc(s->gfix.tileList, (size_t)tiles) || !dalloc(s->rhoX, (size_t)s->G * s->G * s->G)) {
        dfree(s->gfix.fixCount); dfree(s->gfix.tileFlag); dfree(s->gfix.tileList); dfree(s->rhoX);
        return false;
    }
    launch_fill_i32(s->gfix.fixCount, 2, 0, s->st);
    launch_fill_i32(s->gfix.tileFlag, tiles, 0, s->st);
    s->gfix.parity = 0;
    return true;
}

static bool contact_sort(sd_sim* s, bool keysOnly) {
    if (!contact_due(s)) return false;
    if (!ensure_contact_buffers(s)) return false;
    if (keysOnly)
        sort_keys(s, true);
    else
        sort_particles(s, true);
    return true;
}

static void contact_forces(sd_sim* s, bool keysOnly) {
    TIMED(s, "contact", {
        const bool morton = s->modes.keyMode == SD_KEY_MORTON;
        int2* blocks = morton ? s->blockRange : nullptr;
        launch_cell_range(s->sb.keys, s->allocN, morton ? nullptr : s->cellRange, blocks, s->blockCells, s->st);
        launch_contact(s->pos, s->vel, keysOnly ? s->sb.idx : nullptr, s->allocN, s->G, s->periodic,
                       s->modes.keyMode, s->cellRange, s->sb.keys, blocks, s->blockCells,
                       s->cfg.contactRadiusCells / (float)s->G,
                       s->cfg.contactStiffness, s->cfg.contactDamping, s->accContact, s->st);
    });
}

// ---- learned correction -------------------------------------------------------
struct sd_mlp {
    MlpDev d{};
    hipStream_t st = nullptr;
    float* dX = nullptr;  // staging for host-provided inputs
    float* dT = nullptr;
    int32_t* dIdx = nullptr;
    int cap = 0;
    int adamStep = 0;
};

static bool mlp_alloc_batch(sd_mlp* m, int rows) {
    MlpDev& d = m->d;
    const int nPad = mlp_pad_rows(rows);
    if (nPad <= d.maxBatch) return true;
    auto F = [](auto*& p) { dfree(p); };
    F(d.Xb); F(d.Xt); F(d.Y); F(d.dYb); F(d.dYt); F(d.dYf); F(d.lossPart);
    for (int l = 0; l < 3; ++l) { F(d.U[l]); F(d.Hb[l]); F(d.Ht[l]); F(d.Df[l]); F(d.Db[l]); F(d.Dt[l]); }
    const size_t R = (size_t)nPad;
    bool ok = dalloc(d.Xb, R * MLP_INP) && dalloc(d.Xt, R * MLP_INP) && dalloc(d.Y, R * MLP_OUTP) &&
              dalloc(d.dYb, R * MLP_OUTP) && dalloc(d.dYt, R * MLP_OUTP) && dalloc(d.dYf, R * MLP_OUTP) &&
              dalloc(d.lossPart, (R * MLP_OUTP + 255) / 256);
    for (int l = 0; l < 3 && ok; ++l)
        ok = dalloc(d.U[l], R * MLP_H) && dalloc(d.Hb[l], R * MLP_H) && dalloc(d.Ht[l], R * MLP_H) &&
             dalloc(d.Df[l], R * MLP_H) && dalloc(d.Db[l], R * MLP_H) && dalloc(d.Dt[l], R * MLP_H);
    d.maxBatch = ok ? nPad : 0;
    return ok;
}

static bool mlp_stage(sd_mlp* m, int rows) {
    if (rows <= m->cap) return true;
    dfree(m->dX); dfree(m->dT); dfree(m->dIdx);
    const bool ok = dalloc(m->dX, (size_t)rows * MLP_IN) && dalloc(m->dT, (size_t)rows * MLP_OUT) &&
                    dalloc(m->dIdx, (size_t)rows);
    m->cap = ok ? rows : 0;
    return ok;
}

// Runs the network on every particle of the current positions (needs the
// fixed-point density of those positions) into s->accCorr.
// order: nullptr, or the permutation of a sort whose reorder the next
// integrate does (the correction comes out in sorted order either way)
static bool mlp_correction(sd_sim* s, const int32_t* order = nullptr) {
    if (!s->mlp) return false;
    if (!s->accCorr && !dalloc(s->accCorr, (size_t)s->allocN)) return false;
    if (!s->logRho && !dalloc(s->logRho, (size_t)s->G * s->G * s->G)) return false;
    const bool hadRho = rho_fix_ready(s);
    if (!hadRho) {
        // rhoFix is zero whenever it is not valid (fixed-point invariant)
        if (s->modes.depositMode == SD_DEPOSIT_FLOAT)
            fill_rho(s);
        else
            clean_rho(s);
        TIMED(s, "deposit", launch_deposit_fixed(s->pos, s->allocN, s->G, s->periodic, s->rhoFix, s->st));
    }
    const long long cells = (long long)s->G * s->G * s->G;
    TIMED(s, "logrho", launch_log_density(s->rhoFix, cells, s->logRho, s->st));
    MlpInferArgs a{};
    a.pos = s->pos;
    a.order = order;
    a.n = s->allocN;
    a.lrho = s->logRho;
    a.G = s->G;
    a.periodic = s->periodic;
    const MlpDev& d = s->mlp->d;
    a.W1p = d.W1p; a.W2q = d.W2q; a.W3q = d.W3q; a.W4q = d.W4q; a.b = d.b; a.bq = d.bq;
    a.scale = s->mlpOutScale * s->cfg.gravity * (float)s->G / (float)(s->allocN > 0 ? s->allocN : 1);
    a.out = s->accCorr;
    int cus = sd_device_multiprocessors();
    TIMED(s, "mlp", launch_mlp_infer(a, cus > 0 ? cus : 256, s->st));
    if (!hadRho) {
        if (s->modes.depositMode == SD_DEPOSIT_FIXED && !s->meshValid) {
            if (s->rhoValid && s->rhoFloatOnly) {
                // The mesh still reads the gather deposit's float grid (an
                // overlapped solve may already have been forked before this
                // deposit); rhoFix is cleared before it next takes far movers.
                s->rhoDirty = true;
            }
            s->rhoValid = true;
        } else {
            fill_rho(s);
        }
    }
    return true;
}

static void prime_scalars(sd_sim* s) {
    launch_vmax(s->pos, s->vel, s->allocN, s->partials, s->st);
    launch_finalize(s->partials, s->nparts, 2 | 4, dt_policy(s), s->sc, s->st);
    s->primed = true;
}

// One step; no host synchronisation inside.
// timed: record the three phase events of sd_timings (each is a queue marker
// that drains the pipeline, so runs of steps record them on their last step).
// Enqueue a finalize that no mesh solve has carried.
static void flush_finalize(sd_sim* s) {
    if (!s->finDeferred) return;
    const FinalizeJob& j = s->finJob;
    TIMED(s, "finalize", launch_finalize(j.partials, j.nparts, j.mode, j.pol, j.sc, s->st));
    s->finDeferred = false;
}

#ifndef SD_OVERLAP_MESH
#define SD_OVERLAP_MESH 1
#endif

// last: the final step of a run, whose finalize is launched at once.
static void enqueue_step(sd_sim* s, bool timed, bool last) {
    const bool fused = s->modes.pipeline == SD_PIPELINE_FUSED;
    if (timed) HK(hipEventRecord(s->evA, s->st));
    const bool sortDue = s->cfg.sortInterval > 0 && (s->stepCount % s->cfg.sortInterval) == 0 &&
                         !s->cfg.contactEnabled;
    // On a sort step the integrate launch does the reorder itself: it reads
    // through the permutation and writes the sorted arrays, which saves the
    // separate gather pass; the learned correction before it reads the
    // particles through the permutation too.
    // (Only when the density of the current positions is already deposited: a
    // first deposit from unsorted positions would contend on the same cells.)
    const bool fuseReorder = SD_FUSE_REORDER && sortDue && fused && !s->borrowed && s->rhoValid &&
                             s->modes.depositMode == SD_DEPOSIT_FIXED;
    const bool contactDue = contact_due(s);
    // Contact steps reorder inside the integrate too (same conditions): the
    // contact kernel reads the particles through the permutation.
    const bool fuseContact = SD_FUSE_REORDER && contactDue && fused && !s->borrowed && s->rhoValid &&
                             s->modes.depositMode == SD_DEPOSIT_FIXED;
    // The mesh reads only the deposited density and the sort / contact chain
    // only the particles, so on such steps the mesh runs on a second stream
    // beside them and joins before the integrate. (Not with a learned
    // correction, which reads the density that the mesh's first pass clears,
    // nor while the density still has to be deposited from the particles.)
    const bool overlap = SD_OVERLAP_MESH && fused && (sortDue || contactDue) && !s->mlp && !s->borrowed &&
                         s->rhoValid && !s->meshValid && s->modes.depositMode == SD_DEPOSIT_FIXED && s->st2;
    // A step that recomputes the learned correction runs the mesh beside the
    // network instead: both only read the deposited density (the x pass leaves
    // it in place and the integrate's deposit clears it first). The fork is
    // taken here, before the sort, so the mesh waits for neither.
    const bool corrDue = s->mlp && (sortDue || contactDue || !s->corrValid || s->corrAge >= s->mlpInterval);
    const bool mlpOverlap = SD_OVERLAP_MESH && fused && corrDue && !s->borrowed && s->rhoValid && !s->meshValid &&
                            s->modes.depositMode == SD_DEPOSIT_FIXED && s->st2;
    if (overlap || mlpOverlap) HK(hipEventRecord(s->evFork, s->st));
    if (overlap) {
        HK(hipStreamWaitEvent(s->st2, s->evFork, 0));
        solve_mesh(s, s->st2, !fused);
        HK(hipEventRecord(s->evJoin, s->st2));
    }
    if (sortDue) {
        if (fuseReorder)
            sort_keys(s);
        else
            sort_particles(s);
    }
    // The contact sort runs before the learned correction: the correction is
    // stored per slot in the order the integrate writes, so it must see the
    // permutation of this step's last sort.
    const bool contact = contact_sort(s, fuseContact);
    const bool reorderInIntegrate = fuseReorder || (contact && fuseContact);
    bool corr = false;
    if (s->mlp) {
        if (!s->corrValid || s->corrAge >= s->mlpInterval) {
            corr = mlp_correction(s, reorderInIntegrate ? s->sb.idx : nullptr);
            s->corrValid = corr;
            s->corrAge = 0;
        } else {
            corr = true;
        }
    }
    if (mlpOverlap) {
        // enqueued after the network (which reads rhoValid's density) on the host
        // side; on the device it waits only for the fork
        HK(hipStreamWaitEvent(s->st2, s->evFork, 0));
        solve_mesh(s, s->st2, false, true);
        HK(hipEventRecord(s->evJoin, s->st2));
    }
    if (fused) {
        ensure_mesh(s, false);
    } else {
        ensure_rho(s);
        solve_mesh(s, s->st, true);
    }
    // the pressure grid from the last cooling pass's dispersion (computeAccel,
    // Sim.cu:3617-3631); gas physics runs in the reference order only
    const bool gasPressure = !fused && s->cfg.pressureEnabled && ensure_gas_buffers(s);
    if (gasPressure)
        TIMED(s, "pressure", launch_pressure(s->disp, s->accPress, s->G, s->periodic, s->cfg.pressureK, 5.0f, s->st));
    if (contact) contact_forces(s, fuseContact);
    if (overlap || mlpOverlap) HK(hipStreamWaitEvent(s->st, s->evJoin, 0));
    if (timed) HK(hipEventRecord(s->ev1, s->st));
    flush_finalize(s);  // the scalars must be final before this integrate reads them

    IntegrateArgs a{};
    a.accG = s->accG;
    if (fused) {
        // kGridAccel's differences taken in the gather (bit-identical accelerations)
        a.pot = s->pot;
        a.gradH = 0.5f * (float)s->G * pot_scale(s);
        // under half a particle per cell, each block gathers from an LDS copy of its potential box
        a.boxGather = (double)s->allocN < 0.5 * (double)s->G * s->G * s->G ? 1 : 0;
    }
    a.pos = s->pos;
    a.vel = s->vel;
    a.n = s->allocN;
    a.G = s->G;
    a.periodic = s->periodic;
    a.c2 = light_speed_sq(s->cfg.lengthScale, s->cfg.timeUnitScale);
    a.accContact = contact ? s->accContact : nullptr;
    a.accCorr = corr ? s->accCorr : nullptr;
    a.softBoundR = s->periodic ? 0.f : fmaxf(s->cfg.softBoundR, 0.f);
    a.darkEnergy = s->cfg.darkEnergy;
    a.sc = s->sc;
    const DtPolicy pol = dt_policy(s);
    if (fused) {
        if (!s->primed) prime_scalars(s);
        a.applyMeanSub = s->modes.momentumZero ? 1 : 0;
        const bool fusedDeposit = s->modes.depositMode == SD_DEPOSIT_FIXED;
        const bool gatherDep = fusedDeposit && contact && gather_deposit(s);
        a.depositTo = fusedDeposit && !gatherDep ? s->rhoFix : nullptr;
        if (a.depositTo) clean_rho(s);
        a.partials = s->partials;
        if (reorderInIntegrate) {
            a.order = s->sb.idx;
            a.srcPos = s->pos;
            a.srcVel = s->vel;
            a.pos = s->posTmp;
            a.vel = s->velTmp;
        }
        {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer_arm(s, "integrate", e0, e1))
                launch_integrate(a, s->st, e0, e1);
            else
                launch_integrate(a, s->st);
        }
        if (reorderInIntegrate) {
            std::swap(s->pos, s->posTmp);
            std::swap(s->vel, s->velTmp);
        }
        bool floatOnly = false;
        if (gatherDep && gather_float(s) && ensure_gather_float_buffers(s)) {
            // float nodes into rhoX; far movers add into rhoFix, which must be
            // all zero (and is again afterwards)
            clean_rho(s);
            TIMED(s, "deposit", launch_deposit_gather_f32(s->pos, s->blockRange, s->G, s->periodic, s->rhoX,
                                                          s->rhoFix, s->gfix, s->st));
            floatOnly = true;
        } else if (gatherDep) {
            // the new positions are in the contact sort's order (sb.keysAlt is
            // the sort's free buffer: the far-mover list)
            TIMED(s, "deposit", launch_deposit_gather(s->pos, s->blockRange, s->G, s->periodic, s->rhoFix,
                                                      reinterpret_cast<int32_t*>(s->sb.keysAlt), s->sc, s->st));
            s->rhoDirty = false;
        }
        const int mode = (s->modes.momentumZero ? 1 : 0) | 2 | 8 | 4;
        if (last) {
            TIMED(s, "finalize", launch_finalize(s->partials, s->npartsInt, mode, pol, s->sc, s->st));
        } else {
            // Within a run no mesh pass (nor the sort) reads the scalars: the
            // next solve's x pass carries the reduction as one more block.
            s->finJob = FinalizeJob{s->partials, s->npartsInt, mode, pol, s->sc};
            s->finDeferred = true;
        }
        s->primed = true;
        s->rhoValid = fusedDeposit;
        s->rhoFloatOnly = floatOnly;
    } else {
        // kMaxSpeed + dt policy, then kIntegrate, then momentum removal
        launch_vmax(s->pos, s->vel, s->allocN, s->partials, s->st);
        launch_finalize(s->partials, s->nparts, 2 | 4, pol, s->sc, s->st);
        if (gas_physics(&s->cfg)) {
            // cooling after the dt policy and before the kick (Sim.cu:4290); its
            // cell sort may move the particles
            cooling_pass(s, contact);
            a.pos = s->pos;
            a.vel = s->vel;
        }
        a.accPress = gasPressure ? s->accPress : nullptr;
        a.applyMeanSub = 0;
        a.depositTo = nullptr;
        a.partials = nullptr;
        {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer_arm(s, "integrate", e0, e1))
                launch_integrate(a, s->st, e0, e1);
            else
                launch_integrate(a, s->st);
        }
        if (s->modes.momentumZero) {
            launch_momentum(s->pos, s->vel, s->allocN, s->partials, s->st);
            launch_finalize(s->partials, s->nparts, 1 | 8, pol, s->sc, s->st);
            launch_subtract_mean(s->pos, s->vel, s->allocN, s->sc, s->st);
        } else {
            launch_finalize(s->partials, 0, 8, pol, s->sc, s->st);
        }
        s->primed = false;
        s->rhoValid = false;
    }
    s->meshValid = false;
    s->keysSorted = false;
    s->corrAge++;
    s->stepCount++;
    if (timed) HK(hipEventRecord(s->evB, s->st));
}

static void collect_timings(sd_sim* s) {
    HK(hipEventSynchronize(s->evB));
    float a = 0.f, c = 0.f, tot = 0.f;
    HK(hipEventElapsedTime(&a, s->evA, s->ev1));
    HK(hipEventElapsedTime(&c, s->ev1, s->evB));
    HK(hipEventElapsedTime(&tot, s->evA, s->evB));
    HK(hipMemcpy(s->scHost, s->sc, sizeof(StepScalars), hipMemcpyDeviceToHost));
    s->tm.scatterMs = a;
    s->tm.poissonMs = 0.f;
    s->tm.gatherMs = c;
    s->tm.totalMs = tot;
    s->tm.dtUsed = s->scHost->dtPrev;
    s->tm.maxSpeed = s->scHost->vmax;
}

__global__ void k_restart_integrator(StepScalars* sc) {
    sc->meanSub[0] = sc->meanSub[1] = sc->meanSub[2] = 0.f;
    sc->kdkOpen = 1;
}

static void synced_velocity(sd_sim* s);
static bool query_density(sd_sim* s);

static int check_ready(const sd_sim* s) {
    if (g_failed) return SD_ERR_FAILED;
    if (!s || !s->ready) return SD_ERR_STATE;
    return SD_OK;
}

int sd_check_ready(const sd_sim* s) { return check_ready(s); }
DtPolicy sd_dt_policy(const sd_sim* s) { return dt_policy(s); }
float sd_light_speed_sq(float lengthScale, float timeUnitScale) {
    return light_speed_sq(lengthScale, timeUnitScale);
}

void sd_prepare_fused(sd_sim* s) {
    ensure_rho_fix(s);
    if (!s->primed) prime_scalars(s);
}

void sd_note_fused_steps(sd_sim* s, int steps) {
    s->rhoValid = true;
    s->rhoFloatOnly = false;
    s->meshValid = false;
    s->primed = true;
    s->keysSorted = false;
    s->corrValid = false;
    s->corrAge += steps;
    s->stepCount += steps;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free direct sum (sd_direct_sum_noalloc): the FP64
// chunk partials live in this file-scope device array instead of memory
// allocated per call. Each device holds its own copy, at its own address,
// until the process ends; the address is looked up once per device when a
// handle is first configured there. Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol), only the looked-up
// address.
// ---------------------------------------------------------------------------
#define SD_MAX_DEVICES 64
static __device__ __align__(256) double direct_scratch[SD_DIRECT_SCRATCH_DOUBLES];
static double* g_directScratch[SD_MAX_DEVICES] = {};
static std::mutex g_directScratchMu;

// With `device` current: look up this device's copy of the scratch array.
static bool resolve_direct_scratch(int device) {
    if (device < 0 || device >= SD_MAX_DEVICES) {
        mark_failure("device index %d beyond the %d devices the library tracks", device, SD_MAX_DEVICES);
        return false;
    }
    std::lock_guard<std::mutex> lock(g_directScratchMu);
    if (g_directScratch[device]) return true;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(direct_scratch));
    if (e != hipSuccess || !ptr) {
        mark_failure("hipGetSymbolAddress(direct scratch) on device %d: %s", device, hipGetErrorString(e));
        return false;
    }
    g_directScratch[device] = (double*)ptr;
    return true;
}

// ---------------------------------------------------------------------------
// defaults / statics
// ---------------------------------------------------------------------------
extern "C" {

void sd_config_default(sd_config* c) {
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->particleCount = 2000000;
    c->gridSize = 128;
    c->boundary = SD_BOUNDARY_ISOLATED;
    c->preset = SD_PRESET_FILAMENT;
    c->gravity = 0.6f;
    c->softeningCells = 3.0f;
    c->timeScale = 1.0f;
    c->sortInterval = 40;
    c->lengthScale = 1.0f;
    c->timeUnitScale = 1.0f;
    c->softBoundR = 0.f;
    c->darkEnergy = 0.f;
    c->contactEnabled = 0;
    c->contactStiffness = 1.0e6f;
    c->contactDamping = 0.35f;
    c->pressureEnabled = 1;  // SimConfig defaults (Sim.h); plain PM turns them off
    c->coolingEnabled = 1;
    c->contactRadiusCells = 0.5f;
    c->pressureK = 0.45f;
    c->coolingRate = 0.08f;
}

void sd_modes_default(sd_modes* m) {
    if (!m) return;
    m->keyMode = SD_KEY_MORTON;
    m->depositMode = SD_DEPOSIT_FIXED;
    m->integrator = SD_INTEGRATOR_KDK;
    m->pipeline = SD_PIPELINE_FUSED;
    m->momentumZero = 1;
    m->exactGrid = 0;
    m->sortMerge = 0;
}

int sd_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

int sd_device_name(char* buf, size_t len) {
    if (!buf || len == 0) return SD_ERR_ARG;
    buf[0] = 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return SD_ERR_DEVICE;
    snprintf(buf, len, "%s (%s)", p.name, p.gcnArchName);
    return SD_OK;
}

int sd_device_driver(char* buf, size_t len) {
    if (!buf || len == 0) return SD_ERR_ARG;
    int v = 0;
    buf[0] = 0;
    if (hipDriverGetVersion(&v) != hipSuccess) return SD_ERR_DEVICE;
    snprintf(buf, len, "HIP %d.%d", v / 10000000, (v / 100000) % 100);
    return SD_OK;
}

size_t sd_device_free_bytes(void) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    return f;
}

int sd_device_multiprocessors(void) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 0;
    return v;
}

int sd_max_grid_size(int boundary) { return boundary == SD_BOUNDARY_ISOLATED ? 256 : 512; }

size_t sd_estimate_bytes(int particleCount, int gridSize, int boundary) {
    const size_t N = (size_t)(particleCount > 0 ? particleCount : 1);
    const size_t G = (size_t)(gridSize > 0 ? gridSize : 1);
    const size_t S = boundary == SD_BOUNDARY_ISOLATED ? 2 * G : G;
    const size_t P = (size_t)fft_pitch((int)S);
    const size_t cells = G * G * G;
    size_t b = 0;
    b += 16 * N * 4;                    // pos, vel, posTmp, velTmp
    b += 4 * N * 4;                     // keys x2, idx x2
    b += 16 * N;                        // accContact (when contact is on)
    b += 8 * cells;                     // fixed-point density
    b += 8 * G * G * P + 8 * G * S * P; // half-spectrum work arrays
    b += 4 * cells + 16 * cells;        // pot, accG
    b += 8 * cells;                     // contact cell ranges (row-major keys; Morton keys: cells / 8)
    if (boundary == SD_BOUNDARY_ISOLATED) {
        b += 4 * (S / 2 + 1) * (S / 2 + 1) * P;  // Green octant
        b += 2 * 8 * S * S * P;                  // transient Green build arrays
    }
    b += 64ull * 1024 * 1024;           // partials, sort tables, slack
    return b;
}

int sd_max_particles_for(int gridSize, int boundary, size_t freeBytes) {
    const size_t gridPart = sd_estimate_bytes(1, gridSize, boundary);
    if (freeBytes <= gridPart) return 0;
    const size_t per = 16 * 5 + 4 * 4;
    const size_t n = (size_t)((double)(freeBytes - gridPart) * 0.8) / per;
    return n > 1000000000ull ? 1000000000 : (int)n;
}

double sd_device_bandwidth_gbs(void) {
    int kHz = 0, bits = 0;
    if (hipDeviceGetAttribute(&kHz, hipDeviceAttributeMemoryClockRate, 0) != hipSuccess) return 0.0;
    if (hipDeviceGetAttribute(&bits, hipDeviceAttributeMemoryBusWidth, 0) != hipSuccess) return 0.0;
    if (kHz <= 0 || bits <= 0) return 0.0;
    return (double)kHz * 1000.0 * ((double)bits / 8.0) * 2.0 / 1.0e9;
}

double sd_estimate_step_ms(int particleCount, int gridSize, int boundary) {
    const double bw = sd_device_bandwidth_gbs();
    if (bw <= 0.0) return 0.0;
    const double G = (double)(gridSize > 0 ? gridSize : 1);
    // compulsory bytes per step: 64 B per particle (fused gather / kick / drift /
    // deposit) + 140 (isolated) or 60 (periodic) B per mesh cell
    const double mesh = (boundary == SD_BOUNDARY_ISOLATED ? 140.0 : 60.0) * G * G * G;
    const double part = 64.0 * (double)particleCount;
    return (mesh + part) / (bw * 0.5e9) * 1000.0;
}

int sd_failed(void) { return g_failed ? 1 : 0; }
const char* sd_fail_message(void) { return g_failMsg; }
void sd_clear_failure(void) {
    g_failed = false;
    g_failMsg[0] = 0;
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------
sd_sim* sd_create(void) {
    sd_sim* s = new sd_sim();
    sd_config_default(&s->cfg);
    sd_modes_default(&s->modes);
    return s;
}

void sd_destroy(sd_sim* s) {
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    free_buffers(s);
    for (hipEvent_t e : s->timer.pool) (void)hipEventDestroy(e);
    if (s->evA) (void)hipEventDestroy(s->evA);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->evB) (void)hipEventDestroy(s->evB);
    if (s->evFork) (void)hipEventDestroy(s->evFork);
    if (s->evJoin) (void)hipEventDestroy(s->evJoin);
    if (s->st2) (void)hipStreamDestroy(s->st2);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

int sd_set_modes(sd_sim* s, const sd_modes* m) {
    if (!s || !m) return SD_ERR_ARG;
    if (m->keyMode < 0 || m->keyMode > 1 || m->depositMode < 0 || m->depositMode > 1 || m->integrator < 0 ||
        m->integrator > 1 || m->pipeline < 0 || m->pipeline > 1 || m->sortMerge < -1 || m->sortMerge > 1)
        return arg_error("invalid mode value");
    if (s->ready) {
        if (g_failed) return SD_ERR_FAILED;
        if (m->exactGrid != s->modes.exactGrid) return arg_error("exactGrid can change only before sd_init");
        const int rc = check_gas_modes(&s->cfg, m);
        if (rc) return rc;
        // hand the state over synchronised: pending momentum correction and the
        // leapfrog half kick applied, integrator restarted
        synced_velocity(s);
        if (s->borrowed)
            launch_copy_f4(s->velTmp, s->vel, s->allocN, s->st);
        else
            std::swap(s->vel, s->velTmp);
        hipLaunchKernelGGL(k_restart_integrator, dim3(1), dim3(1), 0, s->st, s->sc);
        fill_rho(s);
        s->rhoValid = false;
        s->meshValid = false;
        s->primed = false;
        HK(hipStreamSynchronize(s->st));
    }
    s->modes = *m;
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_get_modes(const sd_sim* s, sd_modes* m) {
    if (!s || !m) return SD_ERR_ARG;
    *m = s->modes;
    return SD_OK;
}

int sd_reconfigure(sd_sim* s, const sd_config* c) {
    if (!s) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    int rc = check_config(c);
    if (rc) return rc;
    rc = check_gas_modes(c, &s->modes);
    if (rc) return rc;
    const int G = choose_grid(c, &s->modes);
    if (!valid_grid(G, c->boundary, s->modes.exactGrid))
        return arg_error("grid size not supported (G = 2^k, or 3*2^k isolated; FFT side <= 512)");
    if (c->boundary == SD_BOUNDARY_PERIODIC && !is_pow2(G)) return arg_error("periodic grid must be 2^k");
    if (!s->st) {
        HK(hipGetDevice(&s->device));
        if (!resolve_direct_scratch(s->device)) return SD_ERR_DEVICE;
        HK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        HK(hipEventCreate(&s->evA));
        HK(hipEventCreate(&s->ev1));
        HK(hipEventCreate(&s->evB));
        // the mesh stream of overlapped steps (default priority, like the library's stream)
        HK(hipStreamCreateWithFlags(&s->st2, hipStreamNonBlocking));
        HK(hipEventCreateWithFlags(&s->evFork, hipEventDisableTiming));
        HK(hipEventCreateWithFlags(&s->evJoin, hipEventDisableTiming));
        if (g_failed) return SD_ERR_DEVICE;
    }
    const bool structural = !s->ready || G != s->G || c->boundary != (s->periodic ? 1 : 0) ||
                            c->particleCount != s->requestedN ||
                            (!s->periodic && c->softeningCells != s->cfg.softeningCells) ||
                            (s->periodic && c->softeningCells != s->cfg.softeningCells);
    if (structural && s->ready && s->borrowed)
        return arg_error("a scene of a batch keeps its particle count, grid and boundary");
    s->cfg = *c;
    s->cfg.gridSize = G;
    if (!structural) return SD_OK;
    // Sim::reconfigure's memory policy (Sim.cu:4120-4135): the particle count is
    // cut to what the free device memory holds; a grid that alone does not fit
    // is a failure. (A batch scene's slice is fixed: it is not cut.)
    int n = c->particleCount;
    const size_t freeB = sd_device_free_bytes();
    if (freeB > 0 && !s->borrowed) {
        const int cap = sd_max_particles_for(G, c->boundary, freeB);
        if (cap <= 0) {
            mark_failure("grid %d does not fit the device (%.1f GB free)", G, freeB / 1e9);
            return SD_ERR_FAILED;
        }
        if (n > cap) {
            fprintf(stderr, "[stardust_amd] particle count %d cut to %d (%.1f GB free)\n", n, cap, freeB / 1e9);
            n = cap;
        }
    }
    s->requestedN = c->particleCount;
    s->G = G;
    s->periodic = c->boundary == SD_BOUNDARY_PERIODIC ? 1 : 0;
    s->S = s->periodic ? G : 2 * G;
    s->P = fft_pitch(s->S);
    s->allocN = n;
    return allocate(s);
}

int sd_init(sd_sim* s, const sd_config* c) {
    if (!s) return SD_ERR_ARG;
    s->ready = false;
    return sd_reconfigure(s, c);
}

int sd_reset(sd_sim* s) {
    int rc = check_ready(s);
    if (rc) return rc;
    // Empty preset: every slot inactive, time zero (Sim::reset, Sim.cu:4160-4238)
    std::vector<float4> hp((size_t)s->allocN, make_float4(-1.f, -1.f, -1.f, 0.f));
    std::vector<float4> hv((size_t)s->allocN, make_float4(0.f, 0.f, 0.f, 0.f));
    HK(hipMemcpy(s->pos, hp.data(), sizeof(float4) * hp.size(), hipMemcpyHostToDevice));
    HK(hipMemcpy(s->vel, hv.data(), sizeof(float4) * hv.size(), hipMemcpyHostToDevice));
    fill_rho(s);
    s->active = 0;
    s->rhoValid = s->meshValid = s->keysSorted = false;
    reset_scalars(s);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_step(sd_sim* s) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (s->allocN <= 0) return SD_OK;
    enqueue_step(s, true, true);
    HK(hipGetLastError());
    collect_timings(s);
    timer_harvest(s);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_run(sd_sim* s, int steps) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (steps < 0) return SD_ERR_ARG;
    if (s->allocN <= 0) return SD_OK;
    for (int k = 0; k < steps; ++k) enqueue_step(s, k == steps - 1, k == steps - 1);
    HK(hipGetLastError());
    if (steps > 0) {
        HK(hipStreamSynchronize(s->st));
        collect_timings(s);
    }
    timer_harvest(s);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_synchronize(sd_sim* s) {
    int rc = check_ready(s);
    if (rc) return rc;
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_get_config(const sd_sim* s, sd_config* c) {
    if (!s || !c) return SD_ERR_ARG;
    *c = s->cfg;
    return SD_OK;
}

int sd_timings_get(sd_sim* s, sd_timings* t) {
    if (!s || !t) return SD_ERR_ARG;
    *t = s->tm;
    return SD_OK;
}

static void read_scalars(sd_sim* s) {
    HK(hipStreamSynchronize(s->st));
    HK(hipMemcpy(s->scHost, s->sc, sizeof(StepScalars), hipMemcpyDeviceToHost));
}

double sd_sim_time(sd_sim* s) {
    if (check_ready(s)) return 0.0;
    read_scalars(s);
    return s->scHost->simTime;
}

long long sd_step_count(sd_sim* s) { return s ? s->stepCount : 0; }
int sd_grid_size(const sd_sim* s) { return s ? s->G : 0; }
int sd_particle_count(const sd_sim* s) { return s ? s->allocN : 0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// reductions used by the measures (device side)
// ---------------------------------------------------------------------------
struct GridStats {
    double sum;
    long long occupied;
    float maxv;
    int maxCount;
};

__global__ void k_grid_stats(const long long* __restrict__ q, long long n, double* sums, long long* occ,
                             unsigned long long* maxBits) {
    __shared__ double s_sum[4];
    __shared__ long long s_occ[4];
    __shared__ long long s_max[4];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long v = i < n ? q[i] : 0;
    double x = (double)v;
    long long o = ((double)v / (double)SD_FIX_ONE > 1e-6) ? 1 : 0;
    long long m = v;
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        o += __shfl_xor(o, off, 64);
        const long long mo = __shfl_xor(m, off, 64);
        m = mo > m ? mo : m;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_sum[w] = x;
        s_occ[w] = o;
        s_max[w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = 0;
        long long to = 0, tm = 0;
        for (int k = 0; k < 4; ++k) {
            ts += s_sum[k];
            to += s_occ[k];
            tm = s_max[k] > tm ? s_max[k] : tm;
        }
        sums[blockIdx.x] = ts;
        occ[blockIdx.x] = to;
        atomicMax(maxBits, (unsigned long long)tm);
    }
}

__global__ void k_particle_stats(const float4* __restrict__ pos, const float4* __restrict__ vel, int n,
                                 double* __restrict__ out5 /* per block: sx, sy, cnt, ke, pad */) {
    __shared__ double s[4][4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sx = 0, sy = 0, c = 0, ke = 0;
    if (i < n) {
        const float4 p = pos[i];
        if (p.x >= 0.f) {
            const float4 v = vel[i];
            sx = p.x;
            sy = p.y;
            c = 1;
            ke = (double)(v.x * v.x + v.y * v.y + v.z * v.z) * 0.5;
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    c = wave_sum(c);
    ke = wave_sum(ke);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s[w][0] = sx;
        s[w][1] = sy;
        s[w][2] = c;
        s[w][3] = ke;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 4; ++q) {
            double t = 0;
            for (int k = 0; k < 4; ++k) t += s[k][q];
            out5[(size_t)blockIdx.x * 4 + q] = t;
        }
    }
}

__global__ void k_cell_count(const float4* __restrict__ pos, int n, int G, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.x < 0.f) return;
    const float fG = (float)G;
    const int cx = min(max((int)(p.x * fG), 0), G - 1);
    const int cy = min(max((int)(p.y * fG), 0), G - 1);
    const int cz = min(max((int)(p.z * fG), 0), G - 1);
    atomicAdd(cnt + ((size_t)cz * G + cy) * G + cx, 1);
}

__global__ void k_max_i32(const int* __restrict__ a, long long n, int* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int v = i < n ? a[i] : 0;
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

__global__ void k_gather_acc(const float4* __restrict__ accG, const float4* __restrict__ pos, int n, int G,
                             int periodic, float4* __restrict__ out);

// Density grid of the current positions into rhoFix for a query; returns true
// if the caller must clear rhoFix afterwards (the mesh already consumed it and
// the fused pipeline expects it empty).
static bool query_density(sd_sim* s) {
    if (rho_fix_ready(s)) return false;
    const bool clearAfter = s->meshValid && s->modes.pipeline == SD_PIPELINE_FUSED;
    fill_rho(s);
    TIMED(s, "deposit", launch_deposit_fixed(s->pos, s->allocN, s->G, s->periodic, s->rhoFix, s->st));
    if (!clearAfter && s->modes.depositMode == SD_DEPOSIT_FIXED) {
        s->rhoValid = true;
        s->rhoFloatOnly = false;
    }
    return clearAfter || s->modes.depositMode != SD_DEPOSIT_FIXED;
}

static GridStats grid_stats(sd_sim* s) {
    GridStats r{};
    const bool clearAfter = query_density(s);
    const long long cells = (long long)s->G * s->G * s->G;
    const int nb = (int)((cells + 255) / 256);
    double* dsum = nullptr;
    long long* docc = nullptr;
    unsigned long long* dmax = nullptr;
    if (!dalloc(dsum, nb) || !dalloc(docc, nb) || !dalloc(dmax, 1)) {
        dfree(dsum); dfree(docc); dfree(dmax);
        return r;
    }
    launch_fill_i64((long long*)dmax, 1, s->st);
    hipLaunchKernelGGL(k_grid_stats, dim3(nb), dim3(256), 0, s->st, s->rhoFix, cells, dsum, docc, dmax);
    std::vector<double> hs(nb);
    std::vector<long long> ho(nb);
    unsigned long long hm = 0;
    HK(hipMemcpyAsync(hs.data(), dsum, sizeof(double) * nb, hipMemcpyDeviceToHost, s->st));
    HK(hipMemcpyAsync(ho.data(), docc, sizeof(long long) * nb, hipMemcpyDeviceToHost, s->st));
    HK(hipMemcpyAsync(&hm, dmax, sizeof(hm), hipMemcpyDeviceToHost, s->st));
    if (clearAfter) fill_rho(s);
    HK(hipStreamSynchronize(s->st));
    for (int b = 0; b < nb; ++b) {
        r.sum += hs[b];
        r.occupied += ho[b];
    }
    r.sum /= (double)SD_FIX_ONE;
    r.maxv = (float)((double)hm / (double)SD_FIX_ONE);
    dfree(dsum); dfree(docc); dfree(dmax);
    return r;
}

// velocities as handed out (pending momentum correction, leapfrog sync) into velTmp
static void synced_velocity(sd_sim* s) {
    const bool kdk = s->modes.integrator == SD_INTEGRATOR_KDK;
    if (kdk) ensure_mesh(s);
    launch_sync_velocity(kdk ? s->accG : nullptr, s->pos, s->vel, s->allocN, s->G, s->periodic, s->sc, s->velTmp,
                         s->st);
}

extern "C" {

int sd_active_count(sd_sim* s) {
    if (check_ready(s)) return 0;
    const int nb = (s->allocN + 255) / 256;
    double* d = nullptr;
    if (!dalloc(d, (size_t)nb * 4)) return 0;
    hipLaunchKernelGGL(k_particle_stats, dim3(nb), dim3(256), 0, s->st, s->pos, s->vel, s->allocN, d);
    std::vector<double> h((size_t)nb * 4);
    HK(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    dfree(d);
    double c = 0;
    for (int b = 0; b < nb; ++b) c += h[(size_t)b * 4 + 2];
    return (int)c;
}

double sd_measure_total_grid_mass(sd_sim* s) {
    if (check_ready(s)) return 0.0;
    return grid_stats(s).sum;
}

double sd_measure_max_density(sd_sim* s) {
    if (check_ready(s)) return 0.0;
    return grid_stats(s).maxv;
}

int sd_measure_occupied_cells(sd_sim* s) {
    if (check_ready(s)) return 0;
    return (int)grid_stats(s).occupied;
}

static int particle_stats(sd_sim* s, bool synced, double out[4]) {
    const int nb = (s->allocN + 255) / 256;
    double* d = nullptr;
    if (!dalloc(d, (size_t)nb * 4)) return SD_ERR_DEVICE;
    const float4* v = s->vel;
    if (synced) {
        synced_velocity(s);
        v = s->velTmp;
    }
    hipLaunchKernelGGL(k_particle_stats, dim3(nb), dim3(256), 0, s->st, s->pos, v, s->allocN, d);
    std::vector<double> h((size_t)nb * 4);
    HK(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    dfree(d);
    for (int q = 0; q < 4; ++q) out[q] = 0;
    for (int b = 0; b < nb; ++b)
        for (int q = 0; q < 4; ++q) out[q] += h[(size_t)b * 4 + q];
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_measure_centroid(sd_sim* s, double* cx, double* cy) {
    if (cx) *cx = 0.5;
    if (cy) *cy = 0.5;
    int rc = check_ready(s);
    if (rc) return rc;
    double o[4];
    rc = particle_stats(s, false, o);
    if (rc) return rc;
    if (o[2] > 0) {
        if (cx) *cx = o[0] / o[2];
        if (cy) *cy = o[1] / o[2];
    }
    return SD_OK;
}

double sd_measure_kinetic_energy(sd_sim* s) {
    if (check_ready(s)) return 0.0;
    double o[4];
    if (particle_stats(s, true, o)) return 0.0;
    return o[3];
}

int sd_peak_cell_count(sd_sim* s) {
    if (check_ready(s)) return -1;
    const long long cells = (long long)s->G * s->G * s->G;
    // The counts borrow the acceleration grid's storage (16 B per cell, 4 used,
    // plus the maximum after them); that grid is marked stale and rebuilt from
    // the potential when something next reads it. Nothing is allocated.
    int* cnt = reinterpret_cast<int*>(s->accG);
    int* mx = cnt + cells;
    s->accGValid = false;
    launch_fill_i32(cnt, cells, 0, s->st);
    launch_fill_i32(mx, 1, 0, s->st);
    if (s->allocN > 0)
        hipLaunchKernelGGL(k_cell_count, dim3((s->allocN + 255) / 256), dim3(256), 0, s->st, s->pos, s->allocN, s->G,
                           cnt);
    hipLaunchKernelGGL(k_max_i32, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->st, cnt, cells, mx);
    int h = 0;
    HK(hipMemcpyAsync(&h, mx, sizeof(int), hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    return g_failed ? -1 : h;
}

// ---------------------------------------------------------------------------
// particles in / out
// ---------------------------------------------------------------------------
int sd_set_particles(sd_sim* s, const float* pos4, const float* vel4, int n) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (n < 0 || n > s->allocN || (n > 0 && (!pos4 || !vel4))) return arg_error("sd_set_particles: bad n or null");
    HK(hipStreamSynchronize(s->st));
    if (n > 0) {
        HK(hipMemcpy(s->pos, pos4, sizeof(float4) * (size_t)n, hipMemcpyHostToDevice));
        HK(hipMemcpy(s->vel, vel4, sizeof(float4) * (size_t)n, hipMemcpyHostToDevice));
    }
    if (n < s->allocN) {
        std::vector<float4> hp((size_t)(s->allocN - n), make_float4(-1.f, -1.f, -1.f, 0.f));
        std::vector<float4> hv((size_t)(s->allocN - n), make_float4(0.f, 0.f, 0.f, 0.f));
        HK(hipMemcpy(s->pos + n, hp.data(), sizeof(float4) * hp.size(), hipMemcpyHostToDevice));
        HK(hipMemcpy(s->vel + n, hv.data(), sizeof(float4) * hv.size(), hipMemcpyHostToDevice));
    }
    s->active = n;
    fill_rho(s);
    s->rhoValid = s->meshValid = s->keysSorted = false;
    s->corrValid = false;
    // keep simTime, restart the integrator state
    read_scalars(s);
    StepScalars z{};
    z.kdkOpen = 1;
    z.simTime = s->scHost->simTime;
    HK(hipMemcpy(s->sc, &z, sizeof(z), hipMemcpyHostToDevice));
    s->primed = false;
    s->stepCount = 0;
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_get_particles(sd_sim* s, float* pos4, float* vel4, int n) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (n < 0 || n > s->allocN) return arg_error("sd_get_particles: bad n");
    if (n == 0) return SD_OK;
    if (vel4) synced_velocity(s);
    if (pos4) HK(hipMemcpyAsync(pos4, s->pos, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, s->st));
    if (vel4) HK(hipMemcpyAsync(vel4, s->velTmp, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_set_fixed_dt(sd_sim* s, float dt) {
    if (!s || dt < 0.f) return SD_ERR_ARG;
    s->fixedDt = dt;
    s->primed = false;
    return SD_OK;
}

// Density views (Sim::densityDevicePtr / fieldDevicePtr, Sim.cu:5151, 5229-5240)
__global__ void k_fix_to_count(const long long* __restrict__ q, long long n, float* __restrict__ f) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = (float)((double)q[i] * (1.0 / (double)SD_FIX_ONE));
}

// kProjectXY without a view rotation: out[y][x] = sum over z of rho[z][y][x],
// added in ascending z (one thread per column, no atomics).
__global__ void k_project_z(const float* __restrict__ rho, int G, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G * G) return;
    const size_t plane = (size_t)G * G;
    float acc = 0.f;
    for (int z = 0; z < G; ++z) acc += rho[(size_t)z * plane + i];
    out[i] = acc;
}

// The float density of the current positions (particles per cell) in rhoF.
static bool density_view(sd_sim* s) {
    if (!ensure_float_rho(s)) return false;
    const long long cells = (long long)s->G * s->G * s->G;
    const bool clearAfter = query_density(s);
    hipLaunchKernelGGL(k_fix_to_count, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->st, s->rhoFix, cells,
                       s->rhoF);
    if (clearAfter) fill_rho(s);
    return true;
}

const float* sd_density_device_ptr(sd_sim* s) {
    if (check_ready(s)) return nullptr;
    if (!density_view(s)) return nullptr;
    HK(hipStreamSynchronize(s->st));
    return g_failed ? nullptr : s->rhoF;
}

const float* sd_field_device_ptr(sd_sim* s, int field) {
    if (check_ready(s)) return nullptr;
    if (field != SD_FIELD_DENSITY) {
        arg_error("sd_field_device_ptr: only the density field is on this path (dispersion, light and ash "
                  "come from the star and gas physics)");
        return nullptr;
    }
    if (!density_view(s)) return nullptr;
    hipLaunchKernelGGL(k_project_z, dim3((unsigned)((s->G * s->G + 255) / 256)), dim3(256), 0, s->st, s->rhoF, s->G,
                       s->proj);
    HK(hipStreamSynchronize(s->st));
    return g_failed ? nullptr : s->proj;
}

const float* sd_particle_pos_device_ptr(sd_sim* s) { return s && s->ready ? (const float*)s->pos : nullptr; }
const float* sd_particle_vel_device_ptr(sd_sim* s) { return s && s->ready ? (const float*)s->vel : nullptr; }

int sd_compute_forces(sd_sim* s) {
    int rc = check_ready(s);
    if (rc) return rc;
    ensure_mesh(s);
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_get_grid(sd_sim* s, int which, void* out, size_t bytes) {
    int rc = check_ready(s);
    if (rc) return rc;
    const size_t cells = (size_t)s->G * s->G * s->G;
    if (!out) return SD_ERR_ARG;
    if (which == SD_GRID_DENSITY_FIXED || which == SD_GRID_DENSITY) {
        const size_t need = which == SD_GRID_DENSITY_FIXED ? cells * 8 : cells * 4;
        if (bytes < need) return arg_error("sd_get_grid: buffer too small");
        const bool clearAfter = query_density(s);
        std::vector<long long> h(cells);
        HK(hipMemcpyAsync(h.data(), s->rhoFix, cells * 8, hipMemcpyDeviceToHost, s->st));
        if (clearAfter) fill_rho(s);
        HK(hipStreamSynchronize(s->st));
        if (which == SD_GRID_DENSITY_FIXED) {
            memcpy(out, h.data(), cells * 8);
        } else {
            float* f = (float*)out;
            for (size_t i = 0; i < cells; ++i) f[i] = (float)((double)h[i] / (double)SD_FIX_ONE);
        }
        return g_failed ? SD_ERR_FAILED : SD_OK;
    }
    if (which == SD_GRID_POTENTIAL || which == SD_GRID_ACCEL) {
        const size_t need = which == SD_GRID_POTENTIAL ? cells * 4 : cells * 16;
        if (bytes < need) return arg_error("sd_get_grid: buffer too small");
        ensure_mesh(s, which == SD_GRID_ACCEL);
        HK(hipMemcpyAsync(out, which == SD_GRID_POTENTIAL ? (const void*)s->pot : (const void*)s->accG, need,
                          hipMemcpyDeviceToHost, s->st));
        HK(hipStreamSynchronize(s->st));
        return g_failed ? SD_ERR_FAILED : SD_OK;
    }
    return arg_error("sd_get_grid: unknown grid");
}

int sd_sort_now(sd_sim* s, uint32_t* keysOut, int32_t* orderOut) {
    int rc = check_ready(s);
    if (rc) return rc;
    sort_particles(s);
    if (keysOut)
        HK(hipMemcpyAsync(keysOut, s->sb.keys, sizeof(uint32_t) * s->allocN, hipMemcpyDeviceToHost, s->st));
    if (orderOut)
        HK(hipMemcpyAsync(orderOut, s->sb.idx, sizeof(int32_t) * s->allocN, hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_particle_accel(sd_sim* s, float* acc4, int n) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (!acc4 || n < 0 || n > s->allocN) return SD_ERR_ARG;
    ensure_mesh(s);
    hipLaunchKernelGGL(k_gather_acc, dim3((s->allocN + 255) / 256), dim3(256), 0, s->st, s->accG, s->pos, s->allocN,
                       s->G, s->periodic, s->velTmp);
    HK(hipMemcpyAsync(acc4, s->velTmp, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_direct_sum(sd_sim* s, const int32_t* targets, int nt, double gm, double eps, double* acc3) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (nt <= 0 || !targets || !acc3) return SD_ERR_ARG;
    for (int t = 0; t < nt; ++t)
        if (targets[t] < 0 || targets[t] >= s->allocN) return arg_error("sd_direct_sum: target out of range");
    const int nch = direct_chunks_for(s->allocN);
    int32_t* dt = nullptr;
    double* part = nullptr;
    double* out = nullptr;
    if (!dalloc(dt, (size_t)nt) || !dalloc(part, (size_t)nch * nt * 3) || !dalloc(out, (size_t)nt * 3)) {
        dfree(dt); dfree(part); dfree(out);
        return SD_ERR_DEVICE;
    }
    HK(hipMemcpyAsync(dt, targets, sizeof(int32_t) * nt, hipMemcpyHostToDevice, s->st));
    TIMED(s, "direct", launch_direct_sum(s->pos, s->allocN, dt, nt, (float)(eps * eps), part, nch, out, gm, s->st));
    HK(hipMemcpyAsync(acc3, out, sizeof(double) * 3 * nt, hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    timer_harvest(s);
    dfree(dt); dfree(part); dfree(out);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_direct_sum_noalloc(sd_sim* s, const int32_t* targetsDev, int nt, double gm, double eps, double* acc3Dev) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (nt <= 0 || !targetsDev || !acc3Dev) return SD_ERR_ARG;
    double* scratch = (s->device >= 0 && s->device < SD_MAX_DEVICES) ? g_directScratch[s->device] : nullptr;
    if (!scratch) return arg_error("sd_direct_sum_noalloc: no scratch address for the handle's device");
    // Targets go in groups whose chunk partials (3 doubles per target and
    // source chunk) fit the scratch; a target's partials and their summation
    // order do not depend on the grouping, so the sums equal sd_direct_sum's.
    const int nch = direct_chunks_for(s->allocN);
    const long long perTarget = 3LL * nch;
    if (perTarget > SD_DIRECT_SCRATCH_DOUBLES)
        return arg_error("sd_direct_sum_noalloc: too many particles for the scratch (3 doubles per 16384 "
                         "particles per target must fit SD_DIRECT_SCRATCH_DOUBLES)");
    const int group = (int)std::min<long long>(nt, SD_DIRECT_SCRATCH_DOUBLES / perTarget);
    const float eps2 = (float)(eps * eps);
    for (int t0 = 0; t0 < nt; t0 += group) {
        const int m = std::min(group, nt - t0);
        TIMED(s, "direct", launch_direct_sum(s->pos, s->allocN, targetsDev + t0, m, eps2, scratch, nch,
                                             acc3Dev + 3 * (size_t)t0, gm, s->st));
    }
    HK(hipGetLastError());
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_kernel_timer_enable(sd_sim* s, const char* name) {
    if (!s) return SD_ERR_ARG;
    KernelTimer& t = s->timer;
    t.name = name ? name : "";
    t.used = 0;
    t.totalMs = 0.0;
    t.launches = 0;
    t.seen = 0;
    t.armed = false;
    if (!t.name.empty() && t.pool.empty()) {
        t.pool.resize(1024);
        for (auto& e : t.pool) HK(hipEventCreate(&e));
    }
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_kernel_timer_every(sd_sim* s, int every) {
    if (!s || every < 1) return SD_ERR_ARG;
    s->timer.every = every;
    s->timer.seen = 0;
    return SD_OK;
}

int sd_kernel_time_ms(sd_sim* s, double* totalMs, long long* launches) {
    if (!s) return SD_ERR_ARG;
    if (s->st) timer_harvest(s);
    if (totalMs) *totalMs = s->timer.totalMs;
    if (launches) *launches = s->timer.launches;
    return SD_OK;
}

// ---------------------------------------------------------------------------
// learned correction
// ---------------------------------------------------------------------------
int sd_mlp_param_counts(int* nW, int* nB) {
    if (nW) *nW = MLP_NW;
    if (nB) *nB = MLP_NB;
    return SD_OK;
}

sd_mlp* sd_mlp_create(unsigned seed) {
    if (g_failed) return nullptr;
    sd_mlp* m = new sd_mlp();
    MlpDev& d = m->d;
    HK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    bool ok = dalloc(d.W, MLP_NW) && dalloc(d.b, MLP_NB) && dalloc(d.gW, MLP_NW) && dalloc(d.gb, MLP_NB) &&
              dalloc(d.mW, MLP_NW) && dalloc(d.vW, MLP_NW) && dalloc(d.mb, MLP_NB) && dalloc(d.vb, MLP_NB) &&
              dalloc(d.b4p, MLP_OUTP) && dalloc(d.gb4p, MLP_OUTP) && dalloc(d.W1p, MLP_H * MLP_INP) &&
              dalloc(d.W2, MLP_H * MLP_H) && dalloc(d.W3, MLP_H * MLP_H) && dalloc(d.W4p, MLP_OUTP * MLP_H) &&
              dalloc(d.W2t, MLP_H * MLP_H) && dalloc(d.W3t, MLP_H * MLP_H) && dalloc(d.W4pt, MLP_OUTP * MLP_H) &&
              dalloc(d.part, 16 * MLP_H * MLP_H) && dalloc(d.W2q, MLP_H * MLP_H) &&
              dalloc(d.W3q, MLP_H * MLP_H) && dalloc(d.W4q, MLP_OUTP * MLP_H) && dalloc(d.bq, 3 * MLP_H);
    if (!ok) {
        sd_mlp_destroy(m);
        return nullptr;
    }
    // deterministic init: uniform(-sqrt(6/fan_in), +) from a 64-bit LCG, zero biases
    std::vector<float> W(MLP_NW), b(MLP_NB, 0.f);
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ (unsigned long long)seed;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    const int fan[4] = {MLP_IN, MLP_H, MLP_H, MLP_H};
    const int cnt[4] = {MLP_H * MLP_IN, MLP_H * MLP_H, MLP_H * MLP_H, MLP_OUT * MLP_H};
    int off = 0;
    for (int l = 0; l < 4; ++l) {
        const double a = std::sqrt(6.0 / fan[l]) * (l == 3 ? 0.1 : 1.0);
        for (int i = 0; i < cnt[l]; ++i) W[off + i] = (float)((2.0 * rnd() - 1.0) * a);
        off += cnt[l];
    }
    sd_mlp_set_params(m, W.data(), b.data());
    return m;
}

void sd_mlp_destroy(sd_mlp* m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    MlpDev& d = m->d;
    auto F = [](auto*& p) { dfree(p); };
    F(d.W); F(d.b); F(d.gW); F(d.gb); F(d.mW); F(d.vW); F(d.mb); F(d.vb); F(d.b4p); F(d.gb4p);
    F(d.W1p); F(d.W2); F(d.W3); F(d.W4p); F(d.W2t); F(d.W3t); F(d.W4pt); F(d.part);
    F(d.W2q); F(d.W3q); F(d.W4q); F(d.bq);
    F(d.Xb); F(d.Xt); F(d.Y); F(d.dYb); F(d.dYt); F(d.dYf); F(d.lossPart);
    for (int l = 0; l < 3; ++l) { F(d.U[l]); F(d.Hb[l]); F(d.Ht[l]); F(d.Df[l]); F(d.Db[l]); F(d.Dt[l]); }
    F(m->dX); F(m->dT); F(m->dIdx);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

int sd_mlp_set_params(sd_mlp* m, const float* W, const float* b) {
    if (!m || !W || !b) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    MlpDev& d = m->d;
    HK(hipMemcpy(d.W, W, sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
    HK(hipMemcpy(d.b, b, sizeof(float) * MLP_NB, hipMemcpyHostToDevice));
    std::vector<float> z(MLP_NW, 0.f);
    HK(hipMemcpy(d.mW, z.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
    HK(hipMemcpy(d.vW, z.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
    HK(hipMemcpy(d.mb, z.data(), sizeof(float) * MLP_NB, hipMemcpyHostToDevice));
    HK(hipMemcpy(d.vb, z.data(), sizeof(float) * MLP_NB, hipMemcpyHostToDevice));
    HK(hipMemcpy(d.b4p, z.data(), sizeof(float) * MLP_OUTP, hipMemcpyHostToDevice));
    m->adamStep = 0;
    mlp_pack(d, m->st);
    HK(hipStreamSynchronize(m->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_mlp_get_params(sd_mlp* m, float* W, float* b) {
    if (!m) return SD_ERR_ARG;
    HK(hipStreamSynchronize(m->st));
    if (W) HK(hipMemcpy(W, m->d.W, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost));
    if (b) HK(hipMemcpy(b, m->d.b, sizeof(float) * MLP_NB, hipMemcpyDeviceToHost));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_mlp_forward(sd_mlp* m, const float* X, int n, float* Y) {
    if (!m || n <= 0 || !X || !Y) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    if (!mlp_alloc_batch(m, n) || !mlp_stage(m, n)) return SD_ERR_DEVICE;
    HK(hipMemcpy(m->dX, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice));
    mlp_forward(m->d, m->dX, n, m->st);
    std::vector<float> y((size_t)mlp_pad_rows(n) * MLP_OUTP);
    HK(hipMemcpyAsync(y.data(), m->d.Y, sizeof(float) * y.size(), hipMemcpyDeviceToHost, m->st));
    HK(hipStreamSynchronize(m->st));
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < MLP_OUT; ++c) Y[(size_t)i * MLP_OUT + c] = y[(size_t)i * MLP_OUTP + c];
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

static double mlp_read_loss(sd_mlp* m) {
    std::vector<double> lp((size_t)m->d.lossBlocks);
    HK(hipMemcpyAsync(lp.data(), m->d.lossPart, sizeof(double) * lp.size(), hipMemcpyDeviceToHost, m->st));
    HK(hipStreamSynchronize(m->st));
    double L = 0.0;
    for (double v : lp) L += v;
    return L;
}

int sd_mlp_grad(sd_mlp* m, const float* X, const float* T, int n, double* loss, float* gW, float* gb) {
    if (!m || n <= 0 || !X || !T) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    if (!mlp_alloc_batch(m, n) || !mlp_stage(m, n)) return SD_ERR_DEVICE;
    HK(hipMemcpy(m->dX, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice));
    HK(hipMemcpy(m->dT, T, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyHostToDevice));
    mlp_batch_grad(m->d, m->dX, m->dT, n, m->st);
    HK(hipMemcpyAsync(m->d.gb + 3 * MLP_H, m->d.gb4p, sizeof(float) * MLP_OUT, hipMemcpyDeviceToDevice, m->st));
    if (loss) *loss = mlp_read_loss(m);
    if (gW) HK(hipMemcpyAsync(gW, m->d.gW, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost, m->st));
    if (gb) HK(hipMemcpyAsync(gb, m->d.gb, sizeof(float) * MLP_NB, hipMemcpyDeviceToHost, m->st));
    HK(hipStreamSynchronize(m->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_mlp_train(sd_mlp* m, const float* X, const float* T, int n, int epochs, int batch, float lr,
                 unsigned seed, float* lossOut) {
    if (!m || n <= 0 || !X || !T || epochs < 0 || batch <= 0) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    batch = batch > n ? n : batch;
    if (!mlp_alloc_batch(m, batch) || !mlp_stage(m, n)) return SD_ERR_DEVICE;
    float *bx = nullptr, *bt = nullptr;
    if (!dalloc(bx, (size_t)batch * MLP_IN) || !dalloc(bt, (size_t)batch * MLP_OUT)) {
        dfree(bx);
        dfree(bt);
        return SD_ERR_DEVICE;
    }
    HK(hipMemcpy(m->dX, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice));
    HK(hipMemcpy(m->dT, T, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyHostToDevice));
    std::vector<int32_t> perm((size_t)n);
    unsigned long long rs = 0xD1B54A32D192ED03ull ^ seed;
    const int nb = n / batch;
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with a 64-bit LCG
            rs = rs * 6364136223846793005ull + 1442695040888963407ull;
            const int j = (int)((rs >> 33) % (unsigned long long)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        HK(hipMemcpyAsync(m->dIdx, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, m->st));
        double epochLoss = 0.0;
        for (int bI = 0; bI < nb; ++bI) {
            launch_gather_rows(m->dX, m->dIdx + (size_t)bI * batch, batch, MLP_IN, bx, m->st);
            launch_gather_rows(m->dT, m->dIdx + (size_t)bI * batch, batch, MLP_OUT, bt, m->st);
            mlp_batch_grad(m->d, bx, bt, batch, m->st);
            mlp_adam(m->d, lr, ++m->adamStep, m->st);
            mlp_pack(m->d, m->st);
            if (bI == nb - 1) epochLoss = mlp_read_loss(m);
        }
        if (lossOut) lossOut[e] = (float)epochLoss;
        HK(hipStreamSynchronize(m->st));
    }
    dfree(bx);
    dfree(bt);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_features(sd_sim* s, const int32_t* targets, int nt, float* X) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (nt <= 0 || !targets || !X) return SD_ERR_ARG;
    for (int t = 0; t < nt; ++t)
        if (targets[t] < 0 || targets[t] >= s->allocN) return arg_error("sd_features: target out of range");
    const bool clearAfter = query_density(s);
    int32_t* dt = nullptr;
    float* dx = nullptr;
    if (!dalloc(dt, (size_t)nt) || !dalloc(dx, (size_t)nt * MLP_IN)) {
        dfree(dt);
        dfree(dx);
        return SD_ERR_DEVICE;
    }
    HK(hipMemcpyAsync(dt, targets, sizeof(int32_t) * nt, hipMemcpyHostToDevice, s->st));
    launch_features(s->pos, dt, nt, s->rhoFix, s->G, s->periodic, dx, s->st);
    HK(hipMemcpyAsync(X, dx, sizeof(float) * MLP_IN * (size_t)nt, hipMemcpyDeviceToHost, s->st));
    if (clearAfter) fill_rho(s);
    HK(hipStreamSynchronize(s->st));
    dfree(dt);
    dfree(dx);
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_attach_mlp(sd_sim* s, sd_mlp* m, float outScale) {
    if (!s) return SD_ERR_ARG;
    if (m) HK(hipStreamSynchronize(m->st));
    s->mlp = m;
    s->mlpOutScale = outScale;
    s->corrValid = false;
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

int sd_set_mlp_interval(sd_sim* s, int steps) {
    if (!s || steps < 1) return SD_ERR_ARG;
    s->mlpInterval = steps;
    return SD_OK;
}

int sd_mlp_correction(sd_sim* s, float* acc4, int n) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (!s->mlp) return SD_ERR_STATE;
    if (!acc4 || n < 0 || n > s->allocN) return SD_ERR_ARG;
    mlp_correction(s);
    s->corrValid = true;
    s->corrAge = 0;
    HK(hipMemcpyAsync(acc4, s->accCorr, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost, s->st));
    HK(hipStreamSynchronize(s->st));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

// ---------------------------------------------------------------------------
// STARDUST v2 state files (Sim.cu:4985-5120), portable stdio
// ---------------------------------------------------------------------------
struct StateHeader {
    char magic[8];
    int version;
    int count;
    int gridSize;
    int preset;
    int boundary;
    float gravity;
    float simTime;
    float bhX, bhY, bhRs, bhMass;
    int bhActive;
    int reserved[5];
};
static_assert(sizeof(StateHeader) == 76, "STARDUST header is 76 bytes");

int sd_save_state(sd_sim* s, const char* path) {
    int rc = check_ready(s);
    if (rc) return rc;
    if (!path) return SD_ERR_ARG;
    const int n = s->active > 0 ? s->active : s->allocN;
    std::vector<float4> hp((size_t)n), hv((size_t)n);
    rc = sd_get_particles(s, (float*)hp.data(), (float*)hv.data(), n);
    if (rc) return rc;
    StateHeader h{};
    memcpy(h.magic, "STARDUST", 8);
    h.version = 2;
    h.count = n;
    h.gridSize = s->G;
    h.preset = s->cfg.preset;
    h.boundary = s->cfg.boundary;
    h.gravity = s->cfg.gravity;
    h.simTime = (float)sd_sim_time(s);
    h.bhX = h.bhY = 0.5f;
    h.reserved[0] = 0;  // no black holes on this path
    FILE* f = fopen(path, "wb");
    if (!f) return SD_ERR_IO;
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(hp.data(), sizeof(float4), (size_t)n, f) == (size_t)n &&
              fwrite(hv.data(), sizeof(float4), (size_t)n, f) == (size_t)n;
    ok = (fclose(f) == 0) && ok;
    return ok ? SD_OK : SD_ERR_IO;
}

int sd_load_state(sd_sim* s, const char* path) {
    if (!s || !path) return SD_ERR_ARG;
    if (g_failed) return SD_ERR_FAILED;
    FILE* f = fopen(path, "rb");
    if (!f) return SD_ERR_IO;
    StateHeader h{};
    if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, "STARDUST", 8) != 0 || (h.version != 1 && h.version != 2) ||
        h.count <= 0 || h.count > 100000000) {
        fclose(f);
        return SD_ERR_IO;
    }
    const int bh = h.version >= 2 ? h.reserved[0] : (h.bhActive ? 1 : 0);
    if (bh != 0) {
        fclose(f);
        return arg_error("state file holds black holes, which this path does not simulate");
    }
    sd_config c = s->cfg;
    c.particleCount = h.count;
    c.gridSize = h.gridSize;
    c.preset = SD_PRESET_EMPTY;
    c.boundary = h.boundary;
    c.gravity = h.gravity;
    int rc = sd_reconfigure(s, &c);
    if (rc) {
        fclose(f);
        return rc;
    }
    const int n = h.count < s->allocN ? h.count : s->allocN;
    std::vector<float4> hp((size_t)n), hv((size_t)n);
    const bool ok = fread(hp.data(), sizeof(float4), (size_t)n, f) == (size_t)n &&
                    fread(hv.data(), sizeof(float4), (size_t)n, f) == (size_t)n;
    fclose(f);
    if (!ok) return SD_ERR_IO;
    rc = sd_set_particles(s, (const float*)hp.data(), (const float*)hv.data(), n);
    if (rc) return rc;
    StepScalars z{};
    z.kdkOpen = 1;
    z.simTime = (double)h.simTime;
    HK(hipMemcpy(s->sc, &z, sizeof(z), hipMemcpyHostToDevice));
    return g_failed ? SD_ERR_FAILED : SD_OK;
}

}  // extern "C"

__global__ void k_gather_acc(const float4* __restrict__ accG, const float4* __restrict__ pos, int n, int G,
                             int periodic, float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.x < 0.f) {
        out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    // sampleAcc, Sim.cu:614-629
    const float fG = (float)G;
    const float gx = fmaf(p.x, fG, -0.5f), gy = fmaf(p.y, fG, -0.5f), gz = fmaf(p.z, fG, -0.5f);
    const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? fx : 1.f - fx) * (oy ? fy : 1.f - fy) * (oz ? fz : 1.f - fz);
        const int x = wrap_or_clamp(ix + ox, G, periodic), y = wrap_or_clamp(iy + oy, G, periodic),
                  z = wrap_or_clamp(iz + oz, G, periodic);
        const float4 g = accG[((size_t)z * G + y) * G + x];
        ax = fmaf(g.x, w, ax);
        ay = fmaf(g.y, w, ay);
        az = fmaf(g.z, w, az);
    }
    out[i] = make_float4(ax, ay, az, 0.f);
}
