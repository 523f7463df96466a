// sd_internal.h -- host-side launchers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sd_device.h"

namespace sd {

// ---- mesh -------------------------------------------------------------------
struct MeshPlan {
    int G;          // mesh side holding mass
    int S;          // FFT side (2G isolated, G periodic)
    int P;          // complex row pitch of the half spectrum
    int periodic;
    const float2* tw;   // exp(-2 pi i m / S), m < S
    float perScale;     // periodic multiplier scale (negative)
    float perSoft;      // periodic Gaussian (2 pi soft / G)^2, 0 = off
};

bool fft_length_supported(int n);
int fft_pitch(int S);
// The solve: x pass of the density (mesh_rows_fwd), then the y, z, y' and x'
// passes (mesh_solve_rest). fin (optional, the same job to both): the previous
// integrate's partials reduction, carried as one more block by the x pass or
// the fused z pass where their block size allows, else launched just before
// the x pass.
int mesh_rows_fwd(const MeshPlan& m, long long* rhoFix, float2* A, int clearSrc, const FinalizeJob* fin,
                  hipStream_t st);
// The x pass from a float density grid (the gather deposit's float output);
// -1 at S = 256, whose register x pass reads only the int64 grid.
int mesh_rows_fwd_f32(const MeshPlan& m, const float* rho, float2* A, const FinalizeJob* fin, hipStream_t st);
int mesh_solve_rest(const MeshPlan& m, float2* A, float2* B, const float* green, float* pot,
                    const FinalizeJob* fin, hipStream_t st);
int mesh_build_green(const MeshPlan& m, float softCells, float2* tmpA, float2* tmpB,
                     float* greenOut, hipStream_t st);
void mesh_grid_accel(const float* pot, float4* acc, int G, float potScale, int periodic,
                     hipStream_t st);

// Meshes of one size (G, S, boundary) solved with one launch per pass. Each job
// is one mesh's buffers; the solve consumes (clears) rhoFix and writes accG.
struct MeshGroupJob {
    MeshPlan plan;
    long long* rhoFix;
    float2* A;
    float2* B;
    const float* green;
    float* pot;
    float4* accG;
    float potScale;
};
struct MeshGroup;
MeshGroup* mesh_group_create(const MeshGroupJob* jobs, int n);  // nullptr if the jobs differ in size
void mesh_group_destroy(MeshGroup* g);
// grad: also write each job's acceleration grid (accG) from its potential
int mesh_group_solve(const MeshGroup* g, bool grad, hipStream_t st);

// ---- keys / sort ------------------------------------------------------------
struct SortBuffers {
    uint32_t* keys;     // n
    int32_t* idx;       // n
    uint32_t* keysAlt;  // n
    int32_t* idxAlt;    // n
    uint32_t* hist;     // 256 * maxTiles
    uint32_t* chunk;    // 256 digit totals of the current pass
    int maxTiles;
};
// Keys of many particle sets held in one array: job j covers particles
// base .. base + n - 1 and its keys get `tag` above the cell key bits, so a
// stable sort keeps every set in place and orders each set as its own sort would.
struct KeyJob {
    int base;
    int n;
    int G;
    int mode;
    uint32_t tag;  // job index << key bits
};
void launch_cell_keys_batch(const float4* pos, const KeyJob* jobs, const int* prefix, int njobs,
                            int totalBlocks, uint32_t* keys, int32_t* idx, hipStream_t st);
int sort_tiles_for(int n);
int sort_chunks_for(int n);  // entries of SortBuffers::chunk
// Cell keys (mode 0 row-major, 1 Morton cell, 2 Morton CIC base node; key =
// cell + 1, 0 for an inactive slot) and identity values into buf.keys /
// buf.idx, and the per-tile histogram of the first radix digit into buf.hist.
void launch_cell_keys_hist(const float4* pos, int n, int G, int mode, SortBuffers& buf, hipStream_t st);
// Stable LSD sort of (keys, idx) by the low `bits` bits. The result is in
// buf.keys / buf.idx on return. firstHistReady: buf.hist already holds the
// first pass's histogram (launch_cell_keys_hist).
void radix_sort_pairs(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady = false);
// Lists of the sort of nearly sorted keys (n entries each, tileMovers one per
// 2048-key tile, counts 2).
struct MoverBuffers {
    uint32_t *mKeys, *mKeysAlt, *sKeys, *tileMovers, *counts;
    int32_t *mIdx, *mIdxAlt, *sIdx;
};
int mover_tiles_for(int n);
// Cell keys of pos (as launch_cell_keys_hist) stably sorted into buf.keys /
// buf.idx, given that buf.keys holds any ascending array of n keys on entry
// (the previous sort's): the keys that differ from it are sorted alone and
// merged into the rest. Exact whatever that array was; fast when few differ.
void sort_nearly_sorted(SortBuffers& buf, const MoverBuffers& mb, const float4* pos, int n, int G, int mode,
                        int bits, hipStream_t st);
void launch_reorder(const float4* srcP, const float4* srcV, const int32_t* order, float4* dstP,
                    float4* dstV, int n, hipStream_t st);
// Cell table of sorted keys (key = cell + 1, 0 = inactive), when range is not
// null: range[c] = (first, one past last) sorted index of cell c. Only
// non-empty cells are written; an entry is current exactly when its first
// index f satisfies 0 <= f < n and keys[f] == c + 1 (entries of cells emptied
// since an earlier call keep stale values, so the table is never cleared).
// With blockRange (Morton keys: cell c = Morton(x, y, z), block c >> 6 = Morton
// of the 4^3-cell block), blockRange[b] = (first, one past last) sorted index of
// block b's particles, (0, 0) for an empty block; the blockEntries int2 after
// them hold each block's 64 cell-occupancy bits (uint64, bit c & 63 of block
// c >> 6). The table (2 * blockEntries int2) is zeroed here first. Contact with
// Morton keys needs only the block table.
void launch_cell_range(const uint32_t* keys, int n, int2* range, int2* blockRange, long long blockEntries,
                       hipStream_t st);

// ---- deposit / integrate ----------------------------------------------------
void launch_deposit_fixed(const float4* pos, int n, int G, int periodic, long long* rho,
                          hipStream_t st);
void launch_deposit_float(const float4* pos, int n, int G, int periodic, float* rho,
                          hipStream_t st);
// Gather-form deposit of positions sorted by 4^3-cell Morton blocks (a contact
// step's order, with blockRange the sorted index range of each block): each
// 32 x 16 x 8 node tile adds the exact fixed-point weights of the particles in
// the surrounding blocks and writes every node once (no clearing, no global
// atomics). Particles more than the halo allows from their block are listed
// and added by a one-block atomic pass, which also resets sc->farCount.
// Gas cooling and pressure (Sim.cu:1212-1415). disp holds four G^3 float grids:
// the x, y, z dispersion sums and the particle count of each cell. Particles
// must be sorted by cell (keys: a cell key + 1, 0 inactive).
void launch_cool_cells(const float4* pos, float4* vel, const uint32_t* keys, int n, int G, float rate, float dtMul,
                       const StepScalars* sc, float* disp, hipStream_t st);
void launch_pressure(const float* disp, float4* accP, int G, int periodic, float k, float maxAcc, hipStream_t st);
bool gather_deposit_supported(int G, int periodic);
void launch_deposit_gather(const float4* pos, const int2* blockRange, int G, int periodic, long long* rho,
                           int32_t* farList, StepScalars* sc, hipStream_t st);
// The same deposit with float nodes (each the value the x pass makes of the
// int64 node) into rhoF. Far movers go to farGrid, an int64 grid that must be
// all zero and is all zero again afterwards, and the tiles they land in are
// summed again exactly (k_deposit_fix). b: device bookkeeping, zero at
// allocation; b.parity flips with each call.
struct GatherFixBuffers {
    int* fixCount = nullptr;  // two counts, used by alternate calls
    int* tileFlag = nullptr;  // gather_deposit_tiles(G) flags
    int* tileList = nullptr;  // gather_deposit_tiles(G) entries
    int parity = 0;
};
int gather_deposit_tiles(int G);
void launch_deposit_gather_f32(const float4* pos, const int2* blockRange, int G, int periodic, float* rhoF,
                               long long* farGrid, GatherFixBuffers& b, hipStream_t st);
void launch_fill_i64(long long* p, long long n, hipStream_t st);
void launch_fill_f32(float* p, long long n, float v, hipStream_t st);
void launch_fill_i32(int32_t* p, long long n, int v, hipStream_t st);

struct IntegrateArgs {
    const float4* accG;
    float4* pos;
    float4* vel;
    int n;
    int G;
    int periodic;
    float c2;
    const float4* accContact;
    const float4* accCorr;   // learned correction (nullptr = off)
    float softBoundR;
    float darkEnergy;
    StepScalars* sc;
    int applyMeanSub;     // subtract sc->meanSub on load
    long long* depositTo; // fused deposit of the new positions (nullptr = off)
    ReducePartial* partials;
    // Fused reorder: when order != nullptr, particle i is read from
    // srcPos/srcVel[order[i]] (a sort's permutation) and written to pos/vel[i],
    // which must be other arrays; otherwise it is read from pos/vel[i].
    const int32_t* order;
    const float4* srcPos;
    const float4* srcVel;
    // With pot set, the gather takes kGridAccel's central differences of the
    // potential itself (accG unused): node accelerations -(pot[+1] - pot[-1]) * gradH,
    // gradH = 0.5 * G * potScale, bit-identical to gathering the gradient grid.
    const float* pot;
    float gradH;
    // gas pressure acceleration grid (float4[G^3]); nullptr = off
    const float4* accPress;
    // sparse meshes (under half a particle per cell): each block copies the
    // potential box its particles' stencils cover into LDS once and gathers
    // from there (same operations, bit-identical accelerations)
    int boxGather;
};
int integrate_blocks(int n);  // k_integrate blocks (partials it writes)
int particle_blocks(int n);   // one-particle-per-thread kernels (vmax, momentum)
void launch_integrate(const IntegrateArgs& a, hipStream_t st, hipEvent_t evStart = nullptr, hipEvent_t evStop = nullptr);

// reductions / scalars
void launch_vmax(const float4* pos, const float4* vel, int n, ReducePartial* partials,
                 hipStream_t st);
void launch_momentum(const float4* pos, const float4* vel, int n, ReducePartial* partials,
                     hipStream_t st);
// Reduce `nparts` partials; mode bit0 = momentum (store mean into sc->meanSub and
// sc->mom), bit1 = vmax -> sc->vmax, bit2 = compute next dt/kick from vmax,
// bit3 = advance simTime/stepCount by the current dt.
void launch_finalize(const ReducePartial* partials, int nparts, int mode, DtPolicy pol,
                     StepScalars* sc, hipStream_t st);
void launch_subtract_mean(const float4* pos, float4* vel, int n, StepScalars* sc, hipStream_t st);

// Many simulations in one launch: job j owns integrate blocks
// prefix[j] .. prefix[j + 1] - 1 (prefix[njobs] = totalBlocks).
void launch_integrate_batch(const IntegrateArgs* jobs, const int* prefix, int njobs, int totalBlocks,
                            hipStream_t st);
void launch_finalize_batch(const FinalizeJob* jobs, int njobs, hipStream_t st);
void launch_copy_f4(const float4* src, float4* dst, long long n, hipStream_t st);
void launch_sync_velocity(const float4* accG, const float4* pos, const float4* vel, int n, int G,
                          int periodic, const StepScalars* sc, float4* velOut, hipStream_t st);

// ---- contact ------------------------------------------------------------------
// keyMode 0 (row-major keys) reads the cell table, keyMode 1 (Morton keys) the
// 4^3-block table and the sorted keys only (cellRange unused). order: nullptr
// when pos / vel are in sorted order, else the sort's permutation (sorted
// particle j in slot order[j]); accOut is in sorted order either way.
void launch_contact(const float4* pos, const float4* vel, const int32_t* order, int n, int G, int periodic,
                    int keyMode, const int2* cellRange, const uint32_t* sortedKeys, const int2* blockRange,
                    long long blockEntries, float radius, float stiffness, float damping, float4* accOut,
                    hipStream_t st);

// ---- direct sum -------------------------------------------------------------
void launch_direct_sum(const float4* pos, int n, const int32_t* targets, int nt, float eps2,
                       double* partial, int nchunks, double* out3, double gm, hipStream_t st);
int direct_chunks_for(int n);

}  // namespace sd
