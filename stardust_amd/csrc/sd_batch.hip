// sd_batch.hip -- many independent simulations advanced together.
//
// No counterpart in the reference (one scene per Sim, Sim.cu:53). A batch is a
// set of ordinary simulation handles (sd_sim) whose particle arrays are slices
// of four arrays the batch owns, so every per-scene call of the C ABI (set /
// get particles, measures, forces, direct sums) keeps working on a scene. What
// the batch adds is the step: every kernel of the fused PM step runs once for
// ALL scenes --
//   sort     : one key launch with the scene index above each scene's own key,
//              one radix sort and one reorder over the whole array (a stable
//              sort keeps every scene in its slice and orders it as its own
//              sort would);
//   mesh     : scenes grouped by mesh size (G, S, boundary), one launch per FFT
//              pass per group;
//   particles: one integrate launch (blocks mapped to scenes through a prefix
//              table; the gather differentiates each scene's potential, so
//              there is no gradient launch) and one finalize launch (a block
//              per scene);
// and a whole run of K steps (plus position snapshots every M steps) is
// recorded once into a HIP graph and replayed with a single launch. Each
// scene's arithmetic is exactly its own single-scene step, so a batched scene
// is bit-identical to the same scene stepped alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/stardust/sd_api.h"
#include "sd_internal.h"
#include "sd_sim.h"

using namespace sd;

struct sd_batch {
    std::vector<sd_config> cfgs;
    std::vector<sd_sim*> scenes;
    std::vector<int> base;  // first slot of each scene in the shared arrays
    long long totalN = 0;
    bool ready = false;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // shared particle storage (the scenes borrow slices)
    float4 *pos = nullptr, *vel = nullptr, *posTmp = nullptr, *velTmp = nullptr;

    // sort over all scenes
    uint32_t *keysA = nullptr, *keysB = nullptr;
    int32_t *idxA = nullptr, *idxB = nullptr;
    uint32_t *hist = nullptr, *chunk = nullptr;
    SortBuffers sb{};
    int sortBits = 0;

    // job tables (rebuilt when a run is recorded)
    KeyJob* dKeys = nullptr;
    int* dKeyPrefix = nullptr;
    int keyBlocks = 0;
    IntegrateArgs* dInteg = nullptr;
    int* dIntegPrefix = nullptr;
    int integBlocks = 0;
    FinalizeJob* dFin = nullptr;
    std::vector<MeshGroup*> groups;

    // snapshots of positions, [snapshot][slot]
    float4* snaps = nullptr;
    int snapCap = 0;
    int nSnaps = 0;

    // recorded run
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t graphNodes = 0;
    int recSteps = -1, recEvery = -1;
    long long recPhase = -1;
    std::vector<sd_config> recCfg;
    std::vector<sd_modes> recModes;  // the graph bakes in each scene's modes too
    std::vector<float> recDt;
};

template <typename T>
static void bfree(T*& p) {
    if (p) {
        (void)hipFree((void*)p);
        p = nullptr;
    }
}

template <typename T>
static bool balloc(T*& p, size_t count) {
    bfree(p);
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc((void**)&p, sizeof(T) * count);
    if (e != hipSuccess) {
        sd_mark_failure("batch hipMalloc(%zu bytes): %s", sizeof(T) * count, hipGetErrorString(e));
        p = nullptr;
        return false;
    }
    return true;
}

static void drop_graph(sd_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graphNodes = 0;
    b->recSteps = b->recEvery = -1;
    b->recPhase = -1;
}

static void drop_tables(sd_batch* b) {
    for (MeshGroup* g : b->groups) mesh_group_destroy(g);
    b->groups.clear();
    bfree(b->dKeys);
    bfree(b->dKeyPrefix);
    bfree(b->dInteg);
    bfree(b->dIntegPrefix);
    bfree(b->dFin);
}

// Key bits of one scene's step sort (row-major cell + 1 or Morton + 1).
static int scene_key_bits(const sd_sim* s) {
    unsigned long long maxKey;
    if (s->modes.keyMode == SD_KEY_MORTON) {
        int b = 0;
        while ((1 << b) < s->G) ++b;
        maxKey = 1ull << (3 * b);
    } else {
        maxKey = (unsigned long long)s->G * s->G * s->G;
    }
    int bits = 0;
    while ((1ull << bits) <= maxKey) ++bits;
    return bits;
}

// ---------------------------------------------------------------------------
sd_batch* sd_batch_create(void) { return new sd_batch(); }

void sd_batch_destroy(sd_batch* b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    drop_graph(b);
    drop_tables(b);
    for (sd_sim* s : b->scenes) sd_destroy(s);  // borrowed arrays are not freed here
    b->scenes.clear();
    bfree(b->pos); bfree(b->vel); bfree(b->posTmp); bfree(b->velTmp);
    bfree(b->keysA); bfree(b->keysB); bfree(b->idxA); bfree(b->idxB);
    bfree(b->hist); bfree(b->chunk); bfree(b->snaps);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int sd_batch_add_scene(sd_batch* b, const sd_config* cfg) {
    if (!b || !cfg) return SD_ERR_ARG;
    if (b->ready) return sd_arg_error("sd_batch_add_scene: the batch is already initialised");
    if (cfg->particleCount <= 0) return sd_arg_error("sd_batch_add_scene: particleCount must be > 0");
    if (cfg->contactEnabled) return sd_arg_error("sd_batch_add_scene: contact is not batched");
    b->cfgs.push_back(*cfg);
    return (int)b->cfgs.size() - 1;
}

int sd_batch_init(sd_batch* b) {
    if (!b) return SD_ERR_ARG;
    if (sd_failure_set()) return SD_ERR_FAILED;
    if (b->ready) return sd_arg_error("sd_batch_init: already initialised");
    const int ns = (int)b->cfgs.size();
    if (ns == 0) return sd_arg_error("sd_batch_init: no scenes");
    if (!sd_device_available()) {
        sd_arg_error("sd_batch_init: no HIP device");
        return SD_ERR_DEVICE;
    }
    long long total = 0;
    b->base.clear();
    for (const sd_config& c : b->cfgs) {
        b->base.push_back((int)total);
        total += c.particleCount;
        if (c.sortInterval != b->cfgs[0].sortInterval)
            return sd_arg_error("sd_batch_init: every scene needs the same sortInterval");
    }
    if (total > 0x7fffffffLL) return sd_arg_error("sd_batch_init: more than 2^31 - 1 particles");
    b->totalN = total;
    const size_t N = (size_t)total;
    bool ok = balloc(b->pos, N) && balloc(b->vel, N) && balloc(b->posTmp, N) && balloc(b->velTmp, N) &&
              balloc(b->keysA, N) && balloc(b->keysB, N) && balloc(b->idxA, N) && balloc(b->idxB, N);
    const int ntiles = sort_tiles_for((int)N);
    ok = ok && balloc(b->hist, (size_t)256 * ntiles) && balloc(b->chunk, (size_t)sort_chunks_for((int)N) + 1);
    if (!ok) return SD_ERR_DEVICE;
    b->sb.keys = b->keysA;
    b->sb.idx = b->idxA;
    b->sb.keysAlt = b->keysB;
    b->sb.idxAlt = b->idxB;
    b->sb.hist = b->hist;
    b->sb.chunk = b->chunk;
    b->sb.maxTiles = ntiles;
    HK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    HK(hipEventCreate(&b->ev0));
    HK(hipEventCreate(&b->ev1));
    int keyBits = 0;
    for (int i = 0; i < ns; ++i) {
        sd_sim* s = sd_create();
        b->scenes.push_back(s);
        sd_modes m;
        sd_modes_default(&m);
        m.exactGrid = 1;  // scene meshes are taken as given (3 * 2^k included)
        sd_set_modes(s, &m);
        s->borrowed = true;
        s->pos = b->pos + b->base[i];
        s->vel = b->vel + b->base[i];
        s->posTmp = b->posTmp + b->base[i];
        s->velTmp = b->velTmp + b->base[i];
        const int rc = sd_init(s, &b->cfgs[i]);
        if (rc != SD_OK) return rc;
        keyBits = std::max(keyBits, scene_key_bits(s));
    }
    int tagBits = 0;
    while ((1 << tagBits) < ns) ++tagBits;
    if (keyBits + tagBits > 32) return sd_arg_error("sd_batch_init: too many scenes for the sort key");
    b->sortBits = keyBits + tagBits;
    b->ready = !sd_failure_set();
    return b->ready ? SD_OK : SD_ERR_FAILED;
}

int sd_batch_scene_count(const sd_batch* b) { return b ? (int)b->cfgs.size() : 0; }

sd_sim* sd_batch_scene(sd_batch* b, int i) {
    if (!b || !b->ready || i < 0 || i >= (int)b->scenes.size()) return nullptr;
    return b->scenes[i];
}

// Job tables from the scenes' current state (fixed dt, gravity, modes).
static int build_tables(sd_batch* b) {
    drop_tables(b);
    const int ns = (int)b->scenes.size();
    int keyBits = 0;
    for (sd_sim* s : b->scenes) keyBits = std::max(keyBits, scene_key_bits(s));
    std::vector<KeyJob> keys(ns);
    std::vector<int> keyPre(ns + 1, 0), intPre(ns + 1, 0);
    std::vector<IntegrateArgs> integ(ns);
    std::vector<FinalizeJob> fin(ns);
    std::map<std::tuple<int, int, int>, std::vector<MeshGroupJob>> groups;
    for (int i = 0; i < ns; ++i) {
        sd_sim* s = b->scenes[i];
        const int n = s->allocN;
        keys[i] = KeyJob{b->base[i], n, s->G, s->modes.keyMode == SD_KEY_MORTON ? 2 : 0, (uint32_t)i << keyBits};
        keyPre[i + 1] = keyPre[i] + (n + 255) / 256;
        IntegrateArgs a{};
        a.accG = s->accG;
        a.pot = s->pot;  // the gather differentiates the potential: no gradient launches
        a.gradH = 0.5f * (float)s->G * (s->cfg.gravity / (float)(s->allocN > 0 ? s->allocN : 1));
        a.pos = s->pos;
        a.vel = s->vel;
        a.n = n;
        a.G = s->G;
        a.periodic = s->periodic;
        a.c2 = sd_light_speed_sq(s->cfg.lengthScale, s->cfg.timeUnitScale);
        a.softBoundR = s->periodic ? 0.f : fmaxf(s->cfg.softBoundR, 0.f);
        a.darkEnergy = s->cfg.darkEnergy;
        a.sc = s->sc;
        a.applyMeanSub = s->modes.momentumZero ? 1 : 0;
        a.depositTo = s->rhoFix;
        a.partials = s->partials;
        integ[i] = a;
        intPre[i + 1] = intPre[i] + integrate_blocks(n);
        fin[i] = FinalizeJob{s->partials, s->npartsInt, (s->modes.momentumZero ? 1 : 0) | 2 | 8 | 4, sd_dt_policy(s),
                             s->sc};
        MeshGroupJob J{};
        J.plan = s->plan;
        J.rhoFix = s->rhoFix;
        J.A = s->A;
        J.B = s->B;
        J.green = s->green;
        J.pot = s->pot;
        J.accG = s->accG;
        J.potScale = s->cfg.gravity / (float)(s->allocN > 0 ? s->allocN : 1);
        groups[std::make_tuple(s->G, s->S, s->periodic)].push_back(J);
    }
    bool ok = balloc(b->dKeys, ns) && balloc(b->dKeyPrefix, ns + 1) && balloc(b->dInteg, ns) &&
              balloc(b->dIntegPrefix, ns + 1) && balloc(b->dFin, ns);
    if (!ok) return SD_ERR_DEVICE;
    HK(hipMemcpy(b->dKeys, keys.data(), sizeof(KeyJob) * ns, hipMemcpyHostToDevice));
    HK(hipMemcpy(b->dKeyPrefix, keyPre.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
    HK(hipMemcpy(b->dInteg, integ.data(), sizeof(IntegrateArgs) * ns, hipMemcpyHostToDevice));
    HK(hipMemcpy(b->dIntegPrefix, intPre.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
    HK(hipMemcpy(b->dFin, fin.data(), sizeof(FinalizeJob) * ns, hipMemcpyHostToDevice));
    b->keyBlocks = keyPre[ns];
    b->integBlocks = intPre[ns];
    for (auto& kv : groups) {
        MeshGroup* g = mesh_group_create(kv.second.data(), (int)kv.second.size());
        if (!g) {
            sd_mark_failure("batch mesh group (G=%d) could not be created", std::get<0>(kv.first));
            return SD_ERR_DEVICE;
        }
        b->groups.push_back(g);
    }
    return sd_failure_set() ? SD_ERR_FAILED : SD_OK;
}

static void record_step(sd_batch* b, long long stepIndex, int sortInterval) {
    hipStream_t st = b->st;
    const int n = (int)b->totalN;
    if (sortInterval > 0 && stepIndex % sortInterval == 0) {
        launch_cell_keys_batch(b->pos, b->dKeys, b->dKeyPrefix, (int)b->scenes.size(), b->keyBlocks, b->sb.keys,
                               b->sb.idx, st);
        radix_sort_pairs(b->sb, n, b->sortBits, st);
        launch_reorder(b->pos, b->vel, b->sb.idx, b->posTmp, b->velTmp, n, st);
        launch_copy_f4(b->posTmp, b->pos, n, st);
        launch_copy_f4(b->velTmp, b->vel, n, st);
    }
    for (MeshGroup* g : b->groups) mesh_group_solve(g, false, st);
    launch_integrate_batch(b->dInteg, b->dIntegPrefix, (int)b->scenes.size(), b->integBlocks, st);
    launch_finalize_batch(b->dFin, (int)b->scenes.size(), st);
}

static int record(sd_batch* b, int steps, int every) {
    drop_graph(b);
    int rc = build_tables(b);
    if (rc) return rc;
    const int nsnap = every > 0 ? steps / every : 0;
    if (nsnap > b->snapCap) {
        if (!balloc(b->snaps, (size_t)nsnap * (size_t)b->totalN)) return SD_ERR_DEVICE;
        b->snapCap = nsnap;
    }
    const long long phase = b->scenes[0]->stepCount;
    const int sortInterval = b->scenes[0]->cfg.sortInterval;
    HK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
    int si = 0;
    for (int k = 0; k < steps; ++k) {
        record_step(b, phase + k, sortInterval);
        if (every > 0 && (k + 1) % every == 0)
            launch_copy_f4(b->pos, b->snaps + (size_t)si++ * (size_t)b->totalN, b->totalN, b->st);
    }
    HK(hipStreamEndCapture(b->st, &b->graph));
    if (sd_failure_set()) return SD_ERR_FAILED;
    HK(hipGraphGetNodes(b->graph, nullptr, &b->graphNodes));
    HK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    if (sd_failure_set()) return SD_ERR_FAILED;
    b->recSteps = steps;
    b->recEvery = every;
    b->recPhase = phase;
    b->recCfg.clear();
    b->recModes.clear();
    b->recDt.clear();
    for (sd_sim* s : b->scenes) {
        b->recCfg.push_back(s->cfg);
        b->recModes.push_back(s->modes);
        b->recDt.push_back(s->fixedDt);
    }
    return SD_OK;
}

static bool recording_matches(const sd_batch* b, int steps, int every) {
    if (!b->exec || b->recSteps != steps || b->recEvery != every) return false;
    if (b->recPhase != b->scenes[0]->stepCount) {
        // the sort schedule repeats when the run covers whole sort intervals
        const int si = b->scenes[0]->cfg.sortInterval;
        if (si <= 0) return true;
        if ((b->recPhase % si) != (b->scenes[0]->stepCount % si)) return false;
    }
    for (size_t i = 0; i < b->scenes.size(); ++i)
        if (memcmp(&b->recCfg[i], &b->scenes[i]->cfg, sizeof(sd_config)) != 0 ||
            memcmp(&b->recModes[i], &b->scenes[i]->modes, sizeof(sd_modes)) != 0 ||
            b->recDt[i] != b->scenes[i]->fixedDt)
            return false;
    return true;
}

int sd_batch_run(sd_batch* b, int steps, int snapshotEvery) {
    if (!b) return SD_ERR_ARG;
    if (sd_failure_set()) return SD_ERR_FAILED;
    if (!b->ready) return SD_ERR_STATE;
    if (steps <= 0) return sd_arg_error("sd_batch_run: steps must be > 0");
    if (snapshotEvery < 0) return sd_arg_error("sd_batch_run: snapshotEvery must be >= 0");
    const long long phase = b->scenes[0]->stepCount;
    for (sd_sim* s : b->scenes) {
        int rc = sd_check_ready(s);
        if (rc) return rc;
        if (s->mlp) return sd_arg_error("sd_batch_run: a scene has a learned correction attached");
        if (s->cfg.contactEnabled) return sd_arg_error("sd_batch_run: contact is not batched");
        if (s->modes.pipeline != SD_PIPELINE_FUSED || s->modes.depositMode != SD_DEPOSIT_FIXED)
            return sd_arg_error("sd_batch_run: scenes must use the fused pipeline and the fixed-point deposit");
        if (s->stepCount != phase) return sd_arg_error("sd_batch_run: scenes are at different step counts");
    }
    // scene state every step starts from: density deposited, dt and kick set
    for (sd_sim* s : b->scenes) sd_prepare_fused(s);
    for (sd_sim* s : b->scenes) HK(hipStreamSynchronize(s->st));
    if (sd_failure_set()) return SD_ERR_FAILED;
    if (!recording_matches(b, steps, snapshotEvery)) {
        HK(hipStreamSynchronize(b->st));  // a previous run may still read the tables
        int rc = record(b, steps, snapshotEvery);
        if (rc) return rc;
    }
    HK(hipEventRecord(b->ev0, b->st));
    HK(hipGraphLaunch(b->exec, b->st));
    HK(hipEventRecord(b->ev1, b->st));
    b->nSnaps = snapshotEvery > 0 ? steps / snapshotEvery : 0;
    for (sd_sim* s : b->scenes) {
        sd_note_fused_steps(s, steps);
        HK(hipStreamWaitEvent(s->st, b->ev1, 0));
    }
    return sd_failure_set() ? SD_ERR_FAILED : SD_OK;
}

int sd_batch_synchronize(sd_batch* b) {
    if (!b || !b->st) return SD_ERR_ARG;
    HK(hipStreamSynchronize(b->st));
    return sd_failure_set() ? SD_ERR_FAILED : SD_OK;
}

float sd_batch_last_run_ms(sd_batch* b) {
    if (!b || !b->exec) return -1.f;
    HK(hipEventSynchronize(b->ev1));
    float ms = -1.f;
    HK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    return ms;
}

long long sd_batch_graph_nodes(const sd_batch* b) { return b ? (long long)b->graphNodes : 0; }
long long sd_batch_total_particles(const sd_batch* b) { return b ? b->totalN : 0; }
int sd_batch_snapshot_count(const sd_batch* b) { return b ? b->nSnaps : 0; }

int sd_batch_get_snapshot(sd_batch* b, int snap, int scene, float* pos4, int n) {
    if (!b || !pos4) return SD_ERR_ARG;
    if (!b->ready) return SD_ERR_STATE;
    if (snap < 0 || snap >= b->nSnaps) return sd_arg_error("sd_batch_get_snapshot: no such snapshot");
    if (scene < 0 || scene >= (int)b->scenes.size()) return sd_arg_error("sd_batch_get_snapshot: no such scene");
    if (n < 0 || n > b->cfgs[scene].particleCount) return sd_arg_error("sd_batch_get_snapshot: bad n");
    HK(hipStreamSynchronize(b->st));
    const float4* src = b->snaps + (size_t)snap * (size_t)b->totalN + b->base[scene];
    HK(hipMemcpy(pos4, src, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost));
    return sd_failure_set() ? SD_ERR_FAILED : SD_OK;
}
