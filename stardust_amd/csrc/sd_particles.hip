// sd_particles.hip -- per-particle kernels of the PM step on gfx950.
//
//   deposit   : kScatter (src/sim/Sim.cu:503-521) as an int64 fixed-point CIC
//               (exact, order-independent) or as the reference's float adds;
//   integrate : kIntegrate + sampleAcc (Sim.cu:614-629, 685-906), plain-PM
//               subset, gathering either the acceleration grid or kGridAccel's
//               differences of the potential (Sim.cu:589-611, bit-identical),
//               with an optional fused deposit of the new positions and
//               fused block partials for the momentum sum (kMomentumAccum,
//               2252-2268) and the CFL speed (kMaxSpeed, 2923-2930);
//   finalize  : deterministic reduction of those partials and the dt policy
//               (Sim.cu:4254-4296) evaluated on the device, so a step needs no
//               host round trip;
//   contact   : kContact (Sim.cu:1030-1203), contact term, with an exact
//               out-of-reach bound that skips cells without changing a bit.
//
// Arithmetic mirrors the reference as built by nvcc (fused multiply-adds
// where nvcc contracts "a*b + c"), written as explicit fmaf with contraction
// otherwise off, so results are reproducible bit for bit against the CPU
// restatement given the same mesh accelerations.
#include <hip/hip_ext.h>

#include "sd_device.h"
#include "sd_internal.h"

#include <algorithm>
#include <climits>
#include <type_traits>

#pragma clang fp contract(off)

// Block renumbering so that each XCD runs a contiguous range of tiles
// (xcd_block); 0 keeps the hardware order.
#ifndef SD_XCD_REMAP
#define SD_XCD_REMAP 1
#endif

namespace sd {

#ifndef SD_PT
#define SD_PT 256
#endif
constexpr int PT = SD_PT;  // threads per particle block

// ---- fills -------------------------------------------------------------------
// A thread stores one 16-byte chunk (the last, partial one element by
// element); a base that is not 16-byte aligned takes the one-element kernel.
template <typename T>
__global__ void k_fill_vec(T* p, long long n, T v) {
    constexpr int E = 16 / (int)sizeof(T);
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * E;
    if (i + E <= n) {
        T t[E];
#pragma unroll
        for (int e = 0; e < E; ++e) t[e] = v;
        uint4 u;
        __builtin_memcpy(&u, t, sizeof(u));
        *reinterpret_cast<uint4*>(p + i) = u;
    } else {
        for (long long j = i; j < n; ++j) p[j] = v;
    }
}
template <typename T>
__global__ void k_fill_one(T* p, long long n, T v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
template <typename T>
static void launch_fill(T* p, long long n, T v, hipStream_t st) {
    if (n <= 0) return;
    if (reinterpret_cast<uintptr_t>(p) % 16 == 0) {
        constexpr long long E = 16 / (long long)sizeof(T);
        const long long chunks = (n + E - 1) / E;
        hipLaunchKernelGGL(k_fill_vec<T>, dim3((unsigned)((chunks + 255) / 256)), dim3(256), 0, st, p, n, v);
    } else {
        hipLaunchKernelGGL(k_fill_one<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    }
}
void launch_fill_i64(long long* p, long long n, hipStream_t st) { launch_fill<long long>(p, n, 0ll, st); }
void launch_fill_f32(float* p, long long n, float v, hipStream_t st) { launch_fill<float>(p, n, v, st); }
void launch_fill_i32(int32_t* p, long long n, int v, hipStream_t st) { launch_fill<int32_t>(p, n, v, st); }

// ---- CIC helpers -------------------------------------------------------------
struct Cic {
    int ix, iy, iz;
    float fx, fy, fz;
};
__device__ __forceinline__ Cic cic_setup(float4 p, float fG) {
    Cic c;
    const float gx = fmaf(p.x, fG, -0.5f), gy = fmaf(p.y, fG, -0.5f), gz = fmaf(p.z, fG, -0.5f);
    c.ix = (int)floorf(gx);
    c.iy = (int)floorf(gy);
    c.iz = (int)floorf(gz);
    c.fx = gx - (float)c.ix;
    c.fy = gy - (float)c.iy;
    c.fz = gz - (float)c.iz;
    return c;
}

// Grid addressing in 32-bit arithmetic. Meshes have G <= 512, so a cell index
// (< 2^27) comes from 24-bit multiplies (full-rate VALU, where 32- and 64-bit
// integer multiplies run at a quarter of the rate), and every grid fits 2^32
// bytes, so an element is a 32-bit byte offset from the array's base, which
// the load or atomic adds to the (uniform) base itself.
__device__ __forceinline__ unsigned cell_idx(int x, int y, int z, int G) {
    return __umul24(__umul24((unsigned)z, (unsigned)G) + (unsigned)y, (unsigned)G) + (unsigned)x;
}
template <typename T>
__device__ __forceinline__ T* elem32(T* base, unsigned idx) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + idx * (unsigned)sizeof(T));
}

// The 2^24 mass units of a particle split exactly along x, y, z (w1 = rint(f w),
// w0 = w - w1), so the eight weights always sum to 2^24. Every partial weight
// is <= 2^24, so 32-bit integers hold them exactly and the float <-> int
// conversions are single instructions.
__device__ __forceinline__ void fixed_weights(const Cic& c, uint32_t w[8]) {
    int wx[2], wxy[4];
    wx[1] = (int)rintf(c.fx * (float)SD_FIX_ONE);
    wx[0] = (int)SD_FIX_ONE - wx[1];
#pragma unroll
    for (int ox = 0; ox < 2; ++ox) {
        const int y1 = (int)rintf(c.fy * (float)wx[ox]);
        wxy[ox + 2] = y1;
        wxy[ox] = wx[ox] - y1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int z1 = (int)rintf(c.fz * (float)wxy[j]);
        w[j + 4] = (uint32_t)z1;
        w[j] = (uint32_t)(wxy[j] - z1);
    }
}

// ---- deterministic deposit ------------------------------------------------------
// Three levels, all in exact integer arithmetic (so the grid is bit-identical
// whatever the order or the level that handled a particle):
//  1. wavefront: lanes whose CIC base node equals the previous lane's form a
//     run (the step sort orders particles by the Morton index of that node);
//     an inclusive segmented scan sums each run's eight weights into its last
//     lane;
//  2. block: run tails add into an LDS tile spanning the block's bounding box
//     of base nodes (+1 for the upper stencil corner);
//  3. grid: each non-zero tile cell is added to the global int64 grid once.
// Blocks whose box exceeds DEP_CAP cells skip level 2 and add run sums to the
// global grid directly. Same-address global atomics from consecutive waves
// were the bottleneck of a per-particle deposit; a block of KP particles
// issues one per touched cell instead.
// Tuning (compile-time): LDS tile size, particles per thread, and whether the
// streamed particle arrays bypass the caches (they are touched once per step).
#ifndef SD_DEP_CAP
#define SD_DEP_CAP 1024
#endif
#ifndef SD_KP
#define SD_KP 2
#endif
#ifndef SD_STREAM_PV
#define SD_STREAM_PV 0
#endif
constexpr int DEP_CAP = SD_DEP_CAP;  // LDS tile cells (8 B each)
constexpr int KP = SD_KP;            // particles per thread in integrate / deposit
constexpr int BP = PT * KP;          // particles per block

typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_pv(const float4* p) {
#if SD_STREAM_PV
    const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
#else
    return *p;
#endif
}
__device__ __forceinline__ void store_pv(float4* p, float4 v) {
#if SD_STREAM_PV
    const f32x4_t w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f32x4_t*>(p));
#else
    *p = v;
#endif
}

// DPP lane shift of a 32-bit value (lanes without a source read 0).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_src(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, true);
}

// Wavefront reductions on DPP (VALU, no LDS permute): quad_perm xor 1 and
// xor 2, row_half_mirror and row_mirror leave every lane of a 16-lane row with
// the row's result; row_bcast 15 (into rows 1, 3) and row_bcast 31 (into rows
// 2, 3) then leave the wavefront's result in lane 63 only. Lanes of rows a
// broadcast does not write take the identity.
#ifndef SD_DPP_REDUCE
#define SD_DPP_REDUCE 1
#endif
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_or(int identity, int v) {
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWMASK, 0xF, false);
}
#define SD_WAVE63(STEP) STEP(0xB1, 0xF) STEP(0x4E, 0xF) STEP(0x141, 0xF) STEP(0x140, 0xF) STEP(0x142, 0xA) STEP(0x143, 0xC)
// Only lane 63's result is used, and it never depends on a lane that a
// broadcast left unwritten (row 3 is written by both broadcasts, and each
// reads the last lane of a row that holds a complete partial). So the
// unwritten lanes take whatever the register holds: no identity fill.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_self(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_any(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, ROWMASK, 0xF, false);
}
__device__ __forceinline__ int wave_min63(int v) {
#define SD_STEP(C, M) v = min(v, dpp_any<C, M>(v));
    SD_WAVE63(SD_STEP)
#undef SD_STEP
    return v;
}
__device__ __forceinline__ int wave_max63(int v) {
#define SD_STEP(C, M) v = max(v, dpp_any<C, M>(v));
    SD_WAVE63(SD_STEP)
#undef SD_STEP
    return v;
}
__device__ __forceinline__ double wave_sum63(double v) {
#define SD_STEP(C, M) v += __hiloint2double(dpp_any<C, M>(__double2hiint(v)), dpp_any<C, M>(__double2loint(v)));
    SD_WAVE63(SD_STEP)
#undef SD_STEP
    return v;
}
// v >= 0 and finite: the IEEE order of such values is the integer order of
// their bits, so the maximum is an integer maximum (no NaN canonicalisation)
__device__ __forceinline__ float wave_max63(float v) {
    return __int_as_float(wave_max63(__float_as_int(v)));
}
// The previous lane's value (DPP wave_shr 1; lane 0 gets `first`).
__device__ __forceinline__ int lane_prev(int first, int v) { return dpp_or<0x138, 0xF>(first, v); }
// The same where lane 0's result is not used (it gets its own value).
__device__ __forceinline__ int lane_prev_any(int v) { return dpp_self<0x138, 0xF>(v); }

#ifndef SD_RUN_SCAN_PREFIX
#define SD_RUN_SCAN_PREFIX 1
#endif
// Segmented inclusive scan over the wavefront; returns true on the lane that
// holds a run's total in v[8]. A run lies inside one wavefront, so its sums
// are <= 64 * 2^24 = 2^30 and fit 32 bits. The scan is Hillis-Steele on DPP:
// row_shr 1, 2, 4, 8 inside each 16-lane row, then row_bcast 15 (rows 1, 3)
// and row_bcast 31 (rows 2, 3); a lane adds a source only if the source lies
// in its own run (source lane >= the run's first lane).
__device__ __forceinline__ bool wave_runs(bool alive, const Cic& c, uint32_t v[8]) {
    const int lane = threadIdx.x & 63;
    if (alive) {
        fixed_weights(c, v);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = 0u;
    }
    const int ix = alive ? c.ix : INT_MIN, iy = alive ? c.iy : INT_MIN, iz = alive ? c.iz : INT_MIN;
#if SD_DPP_REDUCE
    const int pix = lane_prev_any(ix), piy = lane_prev_any(iy), piz = lane_prev_any(iz);  // lane 0 is a head
#else
    const int pix = __shfl_up(ix, 1, 64), piy = __shfl_up(iy, 1, 64), piz = __shfl_up(iz, 1, 64);
#endif
    const bool head = lane == 0 || !alive || pix != ix || piy != iy || piz != iz;
    const uint64_t headMask = __ballot(head);
    const uint64_t le = lane == 63 ? headMask : (headMask & ((2ull << lane) - 1ull));
    const int start = 63 - __clzll(le);
    const bool tail = lane == 63 || ((headMask >> (lane + 1)) & 1ull);
#if SD_RUN_SCAN_PREFIX
    // a run's total is P(tail) - P(start - 1) for the wavefront's inclusive
    // prefix sums P (unsegmented: one DPP add per value and step); P(start - 1)
    // = P(start) - v(start) comes from the run's first lane. Integer sums, so
    // the totals are exact and equal the segmented scan's.
    uint32_t P[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = v[k];
#define SD_PSTEP(CTRL, ROWMASK, BC)                                                                 \
    _Pragma("unroll") for (int k = 0; k < 8; ++k) P[k] +=                                          \
        (uint32_t)__builtin_amdgcn_update_dpp(0, (int)P[k], CTRL, ROWMASK, 0xF, BC);
    SD_PSTEP(0x111, 0xF, true)
    SD_PSTEP(0x112, 0xF, true)
    SD_PSTEP(0x114, 0xF, true)
    SD_PSTEP(0x118, 0xF, true)
    SD_PSTEP(0x142, 0xA, false)
    SD_PSTEP(0x143, 0xC, false)
#undef SD_PSTEP
    const int src = start << 2;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        v[k] = P[k] - (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(P[k] - v[k]));
    return alive && tail;
#endif
#define SD_DPP_STEP(CTRL, ROWMASK, TAKE)                          \
    {                                                             \
        const bool take_ = (TAKE);                                \
        _Pragma("unroll") for (int k = 0; k < 8; ++k) {           \
            const uint32_t y_ = dpp_src<CTRL, ROWMASK>(v[k]);     \
            if (take_) v[k] += y_;                                \
        }                                                         \
    }
    SD_DPP_STEP(0x111, 0xF, (lane & 15) >= 1 && lane - 1 >= start)
    SD_DPP_STEP(0x112, 0xF, (lane & 15) >= 2 && lane - 2 >= start)
    SD_DPP_STEP(0x114, 0xF, (lane & 15) >= 4 && lane - 4 >= start)
    SD_DPP_STEP(0x118, 0xF, (lane & 15) >= 8 && lane - 8 >= start)
    SD_DPP_STEP(0x142, 0xA, ((lane >> 4) & 1) && (lane & ~15) - 1 >= start)
    SD_DPP_STEP(0x143, 0xC, lane >= 32 && 31 >= start)
#undef SD_DPP_STEP
    return alive && tail;
}

__device__ __forceinline__ void add_global8(const Cic& c, const uint32_t v[8], int G, int periodic,
                                            long long* rho) {
    int xs[2], ys[2], zs[2];
    xs[0] = wrap_or_clamp(c.ix, G, periodic);
    xs[1] = wrap_or_clamp(c.ix + 1, G, periodic);
    ys[0] = wrap_or_clamp(c.iy, G, periodic);
    ys[1] = wrap_or_clamp(c.iy + 1, G, periodic);
    zs[0] = wrap_or_clamp(c.iz, G, periodic);
    zs[1] = wrap_or_clamp(c.iz + 1, G, periodic);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        if (v[k] != 0u)
            atomicAdd((unsigned long long*)elem32(rho, cell_idx(xs[ox], ys[oy], zs[oz], G)), (unsigned long long)v[k]);
    }
}

// Block-wide deposit of KP particles per thread into an LDS tile of DEP_CAP
// cells (s_tile). Must be called by every thread of the block (it synchronises).
__device__ __forceinline__ void block_deposit_in(const bool alive[KP], const Cic c[KP], int G, int periodic,
                                                 long long* rho, unsigned long long* __restrict__ s_tile) {
    __shared__ int s_box[PT / 64][6];
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        if (alive[k]) {
            lo[0] = min(lo[0], c[k].ix); hi[0] = max(hi[0], c[k].ix);
            lo[1] = min(lo[1], c[k].iy); hi[1] = max(hi[1], c[k].iy);
            lo[2] = min(lo[2], c[k].iz); hi[2] = max(hi[2], c[k].iz);
        }
    }
#if SD_DPP_REDUCE
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_min63(lo[d]);
        hi[d] = wave_max63(hi[d]);
    }
    constexpr int LEAD = 63;
#else
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = max(hi[d], __shfl_xor(hi[d], o, 64));
        }
    }
    constexpr int LEAD = 0;
#endif
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == LEAD) {
        for (int d = 0; d < 3; ++d) {
            s_box[w][d] = lo[d];
            s_box[w][3 + d] = hi[d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PT / 64; ++q) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], s_box[q][d]);
            hi[d] = max(hi[d], s_box[q][3 + d]);
        }
    }
    if (lo[0] == INT_MAX) return;  // no live particle in the block (uniform)
    const long long dx = (long long)hi[0] - lo[0] + 2, dy = (long long)hi[1] - lo[1] + 2,
                    dz = (long long)hi[2] - lo[2] + 2;
    const long long vol = dx * dy * dz;
    if (vol > DEP_CAP) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            uint32_t v[8];
            if (wave_runs(alive[k], c[k], v)) add_global8(c[k], v, G, periodic, rho);
        }
        return;
    }
    const int nv = (int)vol, ndx = (int)dx, ndy = (int)dy;
    for (int t = threadIdx.x; t < nv; t += PT) s_tile[t] = 0ull;
    __syncthreads();
    // corner q of a run's base cell sits at base + off(q): the tile spans
    // fewer than 2^10 cells, so 24-bit multiplies give the base, and the
    // offsets are block-uniform; a zero weight adds nothing and is not skipped
    const unsigned txy = __umul24((unsigned)ndx, (unsigned)ndy);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        uint32_t v[8];
        if (wave_runs(alive[k], c[k], v)) {
            const unsigned bx = (unsigned)(c[k].ix - lo[0]), by = (unsigned)(c[k].iy - lo[1]),
                           bz = (unsigned)(c[k].iz - lo[2]);
            unsigned long long* t0 = s_tile + (__umul24(__umul24(bz, (unsigned)ndy) + by, (unsigned)ndx) + bx);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const unsigned off = ((q & 1) ? 1u : 0u) + (((q >> 1) & 1) ? (unsigned)ndx : 0u) + (((q >> 2) & 1) ? txy : 0u);
                atomicAdd(t0 + off, (unsigned long long)v[q]);
            }
        }
    }
    __syncthreads();
    // tile cell t -> (x, y, z): quotients from float reciprocals, exact since
    // t < 2^10 keeps (t + 1/2) / d at least 1/(2d) from an integer, far beyond
    // the float rounding
    const int nxy = ndx * ndy;
    const float invX = 1.0f / (float)ndx, invXY = 1.0f / (float)nxy;
    for (int t = threadIdx.x; t < nv; t += PT) {
        const unsigned long long v = s_tile[t];
        if (v == 0ull) continue;
        const int z = (int)(((float)t + 0.5f) * invXY);
        const int r = t - __mul24(z, nxy);
        const int y = (int)(((float)r + 0.5f) * invX);
        const int x = r - __mul24(y, ndx);
        const int gx = wrap_or_clamp(lo[0] + x, G, periodic), gy = wrap_or_clamp(lo[1] + y, G, periodic),
                  gz = wrap_or_clamp(lo[2] + z, G, periodic);
        atomicAdd((unsigned long long*)elem32(rho, cell_idx(gx, gy, gz, G)), v);
    }
}

__device__ __forceinline__ void block_deposit(const bool alive[KP], const Cic c[KP], int G, int periodic,
                                              long long* rho) {
    __shared__ unsigned long long s_tile[DEP_CAP];
    block_deposit_in(alive, c, G, periodic, rho, s_tile);
}

__global__ void __launch_bounds__(PT) k_deposit_fixed(const float4* __restrict__ pos, int n, int G, int periodic,
                                                     long long* __restrict__ rho) {
    bool alive[KP];
    Cic c[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int i = blockIdx.x * BP + k * PT + threadIdx.x;
        float4 p = make_float4(-1.f, -1.f, -1.f, 0.f);
        if (i < n) p = pos[i];
        alive[k] = p.x >= 0.f;
        c[k] = cic_setup(alive[k] ? p : make_float4(0.f, 0.f, 0.f, 0.f), (float)G);
    }
    block_deposit(alive, c, G, periodic, rho);
}

// Reference float-atomic deposit (kScatter, order-dependent) for parity runs.
__global__ void __launch_bounds__(PT) k_deposit_float(const float4* __restrict__ pos, int n, int G, int periodic,
                                                     float* __restrict__ rho) {
    const int i = blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.x < 0.f) return;
    const Cic c = cic_setup(p, (float)G);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? c.fx : 1.f - c.fx) * (oy ? c.fy : 1.f - c.fy) * (oz ? c.fz : 1.f - c.fz);
        const int x = wrap_or_clamp(c.ix + ox, G, periodic), y = wrap_or_clamp(c.iy + oy, G, periodic),
                  z = wrap_or_clamp(c.iz + oz, G, periodic);
        atomicAdd(rho + ((size_t)z * G + y) * G + x, w);
    }
}

void launch_deposit_fixed(const float4* pos, int n, int G, int periodic, long long* rho, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_deposit_fixed, dim3((n + BP - 1) / BP), dim3(PT), 0, st, pos, n, G, periodic, rho);
}
void launch_deposit_float(const float4* pos, int n, int G, int periodic, float* rho, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_deposit_float, dim3((n + PT - 1) / PT), dim3(PT), 0, st, pos, n, G, periodic, rho);
}

// ---- gather-form deposit (sparse grids) ------------------------------------------
// At under a particle per cell the block tile above overflows (a block's
// particles span more nodes than the tile holds) and every weight becomes a
// global int64 atomic into a grid larger than the caches. When the particles
// are sorted by 4^3-cell Morton blocks (contact steps: blockRange holds each
// block's sorted index range, and the integrate wrote the new positions in
// that order), the deposit can instead be gathered: a workgroup owns a
// 32 x 16 x 8 tile of nodes, reads the particles of the blocks around it, adds
// the weights that fall inside the tile into LDS (exact int64) and stores each
// node once. A particle whose new CIC base node lies within [-2, 4] cells of
// its pre-drift block's origin on every axis (any step that moves it less than
// a cell) touches only nodes of tiles that read its block; the rest ("far
// movers") are listed by the tile that owns their block and added afterwards
// with global atomics. Integer sums make the grid independent of the order, so
// it equals the atomic deposit's bit for bit.
#ifndef SD_GD_X
#define SD_GD_X 32
#endif
#ifndef SD_GD_Y
#define SD_GD_Y 16
#endif
#ifndef SD_GD_Z
#define SD_GD_Z 8
#endif
#ifndef SD_GD_THREADS
#define SD_GD_THREADS 512
#endif
constexpr int GDX = SD_GD_X, GDY = SD_GD_Y, GDZ = SD_GD_Z;  // tile nodes per axis
constexpr int GBX = GDX / 4 + 2, GBY = GDY / 4 + 2, GBZ = GDZ / 4 + 2;  // blocks read per axis
constexpr int GRUNS = GBX * GBY * GBZ;             // 240 for a 32 x 16 x 8 tile
constexpr int GTHREADS = SD_GD_THREADS;
constexpr int GLUT = 4096;  // particle visits per tile found through a run table (more: binary search)
static_assert(GRUNS <= 64 * 4 && GRUNS < 256, "runs per tile: four per lane of one wavefront, byte run table");

// Tiles must divide the mesh; on a periodic mesh the blocks a tile reads must
// be distinct after wrapping (at least GBX, GBY, GBZ blocks per axis).
bool gather_deposit_supported(int G, int periodic) {
    if (G < 32 || G % GDX || G % GDY || G % GDZ) return false;
    const int nb = G / 4;
    return !periodic || (nb >= GBX && nb >= GBY && nb >= GBZ);
}

// Offset of a new base node from its block's origin (4 * bw), folded into
// [-G/2, G/2) on a periodic mesh.
__device__ __forceinline__ int block_offset(int base, int bw, int G, int periodic) {
    int d = base - 4 * bw;
    if (periodic) {
        if (d >= G / 2) d -= G;
        if (d < -G / 2) d += G;
    }
    return d;
}

// Far-mover bookkeeping of a gather deposit. int64 output: the far movers,
// listed by the tile that owns their block (farList, farCount), are added by
// k_deposit_far. Float output (tileFlag set): the owning tile adds a far
// mover's weights to farGrid (an all-zero int64 grid) with global atomics and
// marks the tiles its eight nodes fall in; a tile marked for the first time is
// appended to tileList (count fixCount), and k_deposit_fix sums those tiles
// again with farGrid's nodes added (and cleared).
struct FarCtl {
    int32_t* farList;
    int* farCount;
    int* tileFlag;  // nullptr: int64 output
    int* tileList;
    int* fixCount;
    long long* farGrid;
};

__device__ __forceinline__ int gather_tile_id(int x, int y, int z, int G) {
    return ((z / GDZ) * (G / GDY) + y / GDY) * (G / GDX) + x / GDX;
}

// The exact int64 weights of the tile at (X0, Y0, Z0) from the particles of
// the Morton block runs around it, in s_tile (cleared here). A far mover is
// skipped; with LIST the tile that owns its block lists it (and marks the
// tiles it lands in, when f.tileFlag is set). Called by every thread of the
// block; s_tile is complete when it returns.
template <bool LIST>
__device__ __forceinline__ void gather_tile(const float4* __restrict__ pos, const int2* __restrict__ blockRange,
                                            int G, int periodic, int X0, int Y0, int Z0,
                                            unsigned long long* __restrict__ s_tile, const FarCtl& f) {
    __shared__ int s_first[GRUNS], s_pre[GRUNS + 1];
    __shared__ int s_blk[GRUNS];  // wrapped block coordinates, 10 bits each
    __shared__ unsigned char s_run[GLUT];  // visit index -> run, when the visits fit
    const int tid = threadIdx.x;
    const int NB = G / 4;
    for (int t = tid; t < GDX * GDY * GDZ; t += GTHREADS) s_tile[t] = 0ull;
    if (tid < 64) {
        // up to four runs per lane, then a wavefront scan of the lengths
        constexpr int RPL = (GRUNS + 63) / 64;
        int len[RPL];
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            len[q] = 0;
            const int r = RPL * tid + q;
            if (r >= GRUNS) continue;
            int bx = X0 / 4 - 1 + r % GBX, by = Y0 / 4 - 1 + (r / GBX) % GBY, bz = Z0 / 4 - 1 + r / (GBX * GBY);
            bool ok = true;
            if (periodic) {
                bx = (bx + NB) % NB; by = (by + NB) % NB; bz = (bz + NB) % NB;
            } else {
                ok = bx >= 0 && bx < NB && by >= 0 && by < NB && bz >= 0 && bz < NB;
            }
            int2 run = make_int2(0, 0);
            if (ok) run = blockRange[morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz)];
            s_first[r] = run.x;
            s_blk[r] = ok ? (bx | (by << 10) | (bz << 20)) : -1;
            len[q] = run.y > run.x ? run.y - run.x : 0;
        }
        int sum = 0;
#pragma unroll
        for (int q = 0; q < RPL; ++q) sum += len[q];
        int inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (tid >= o) inc += y;
        }
        int ex = inc - sum;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int r = RPL * tid + q;
            if (r < GRUNS) s_pre[r] = ex;
            ex += len[q];
        }
        if (tid == 63) s_pre[GRUNS] = inc;
    }
    __syncthreads();
    const int total = s_pre[GRUNS];
    const bool lut = total <= GLUT;
    if (lut) {
        for (int r = tid; r < GRUNS; r += GTHREADS)
            for (int j = s_pre[r]; j < s_pre[r + 1]; ++j) s_run[j] = (unsigned char)r;
        __syncthreads();
    }
    const float fG = (float)G;
    for (int j = tid; j < total; j += GTHREADS) {
        // the run holding flattened index j: the last r with s_pre[r] <= j
        int r;
        if (lut) {
            r = s_run[j];
        } else {
            int lo = 0, hi = GRUNS - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_pre[mid] <= j) lo = mid; else hi = mid - 1;
            }
            r = lo;
        }
        const int slot = s_first[r] + (j - s_pre[r]);
        const float4 p = pos[slot];
        if (p.x < 0.f) continue;
        const Cic c = cic_setup(p, fG);
        const int blk = s_blk[r];
        const int bx = blk & 1023, by = (blk >> 10) & 1023, bz = blk >> 20;
        const int dx = block_offset(c.ix, bx, G, periodic), dy = block_offset(c.iy, by, G, periodic),
                  dz = block_offset(c.iz, bz, G, periodic);
        if (dx < -2 || dx > 4 || dy < -2 || dy > 4 || dz < -2 || dz > 4) {
            // a far mover: listed once, by the tile that owns its block
            const int ri = r % GBX, rj = (r / GBX) % GBY, rk = r / (GBX * GBY);
            if (LIST && ri >= 1 && ri < GBX - 1 && rj >= 1 && rj < GBY - 1 && rk >= 1 && rk < GBZ - 1) {
                if (f.tileFlag) {
                    uint32_t w[8];
                    fixed_weights(c, w);
                    add_global8(c, w, G, periodic, f.farGrid);
                    int xs[2], ys[2], zs[2];
#pragma unroll
                    for (int o = 0; o < 2; ++o) {
                        xs[o] = wrap_or_clamp(c.ix + o, G, periodic);
                        ys[o] = wrap_or_clamp(c.iy + o, G, periodic);
                        zs[o] = wrap_or_clamp(c.iz + o, G, periodic);
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int t = gather_tile_id(xs[q & 1], ys[(q >> 1) & 1], zs[q >> 2], G);
                        if (atomicOr(f.tileFlag + t, 1) == 0) f.tileList[atomicAdd(f.fixCount, 1)] = t;
                    }
                } else {
                    f.farList[atomicAdd(f.farCount, 1)] = slot;
                }
            }
            continue;
        }
        int rx[2], ry[2], rz[2];
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            // the integrate's positions: a periodic one lies in [0, 1], so ix + o is in [-1, G]
            rx[o] = wrap_or_clamp1(c.ix + o, G, periodic) - X0;
            ry[o] = wrap_or_clamp1(c.iy + o, G, periodic) - Y0;
            rz[o] = wrap_or_clamp1(c.iz + o, G, periodic) - Z0;
        }
        const bool inx0 = (unsigned)rx[0] < GDX, inx1 = (unsigned)rx[1] < GDX;
        const bool iny0 = (unsigned)ry[0] < GDY, iny1 = (unsigned)ry[1] < GDY;
        const bool inz0 = (unsigned)rz[0] < GDZ, inz1 = (unsigned)rz[1] < GDZ;
        if (!((inx0 || inx1) && (iny0 || iny1) && (inz0 || inz1))) continue;
        uint32_t w[8];
        fixed_weights(c, w);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int ox = q & 1, oy = (q >> 1) & 1, oz = (q >> 2) & 1;
            const bool in = (ox ? inx1 : inx0) && (oy ? iny1 : iny0) && (oz ? inz1 : inz0);
            if (in && w[q] != 0u)
                atomicAdd(&s_tile[((oz ? rz[1] : rz[0]) * GDY + (oy ? ry[1] : ry[0])) * GDX + (ox ? rx[1] : rx[0])],
                          (unsigned long long)w[q]);
        }
    }
    __syncthreads();
}

// The float node value the x pass makes of an int64 node (so a float grid
// written here reads as the int64 grid would).
__device__ __forceinline__ float fix_to_float(unsigned long long v) {
    return (float)(long long)v * (1.0f / (float)SD_FIX_ONE);
}

// FOUT = false: int64 nodes into rho; FOUT = true: their float values into rhoF.
template <bool FOUT>
__global__ void __launch_bounds__(GTHREADS) k_deposit_gather(const float4* __restrict__ pos,
                                                            const int2* __restrict__ blockRange, int G, int periodic,
                                                            long long* __restrict__ rho, float* __restrict__ rhoF,
                                                            FarCtl f) {
    __shared__ unsigned long long s_tile[GDX * GDY * GDZ];
    const int X0 = blockIdx.x * GDX, Y0 = blockIdx.y * GDY, Z0 = blockIdx.z * GDZ;
    gather_tile<true>(pos, blockRange, G, periodic, X0, Y0, Z0, s_tile, f);
    for (int t = threadIdx.x; t < GDX * GDY * GDZ; t += GTHREADS) {
        const int x = t % GDX, y = (t / GDX) % GDY, z = t / (GDX * GDY);
        const unsigned cell = cell_idx(X0 + x, Y0 + y, Z0 + z, G);
        if (FOUT)
            *elem32(rhoF, cell) = fix_to_float(s_tile[t]);
        else
            *elem32(rho, cell) = (long long)s_tile[t];
    }
}

// The far movers of the gather deposit, added with global atomics after every
// tile has stored its nodes; one block, which then resets the count.
__global__ void __launch_bounds__(GTHREADS) k_deposit_far(const float4* __restrict__ pos, int G, int periodic,
                                                         long long* __restrict__ rho,
                                                         const int32_t* __restrict__ farList, int* farCount) {
    const int n = *(volatile int*)farCount;
    for (int i = threadIdx.x; i < n; i += GTHREADS) {
        const float4 p = pos[farList[i]];
        const Cic c = cic_setup(p, (float)G);
        uint32_t w[8];
        fixed_weights(c, w);
        add_global8(c, w, G, periodic, rho);
    }
    __syncthreads();
    if (threadIdx.x == 0) *farCount = 0;
}

// Float output: the tiles the far movers land in, summed again in int64 with
// farGrid's nodes added (and cleared, so farGrid is all zero again), then
// stored as floats; each tile's flag is cleared. fixCountNext, the count of
// the other parity, was consumed by the previous float deposit and is reset
// here for the next one.
__global__ void __launch_bounds__(GTHREADS) k_deposit_fix(const float4* __restrict__ pos,
                                                         const int2* __restrict__ blockRange, int G, int periodic,
                                                         float* __restrict__ rhoF, long long* __restrict__ farGrid,
                                                         const int* __restrict__ fixCount, int* __restrict__ tileFlag,
                                                         const int* __restrict__ tileList, int* __restrict__ fixCountNext) {
    __shared__ unsigned long long s_tile[GDX * GDY * GDZ];
    if (blockIdx.x == 0 && threadIdx.x == 0) *fixCountNext = 0;
    const int nfix = *fixCount;
    const int ntx = G / GDX, nty = G / GDY;
    const FarCtl none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int k = blockIdx.x; k < nfix; k += gridDim.x) {
        const int t = tileList[k];
        const int X0 = (t % ntx) * GDX, Y0 = ((t / ntx) % nty) * GDY, Z0 = (t / (ntx * nty)) * GDZ;
        gather_tile<false>(pos, blockRange, G, periodic, X0, Y0, Z0, s_tile, none);
        for (int e = threadIdx.x; e < GDX * GDY * GDZ; e += GTHREADS) {
            const int x = e % GDX, y = (e / GDX) % GDY, z = e / (GDX * GDY);
            const unsigned cell = cell_idx(X0 + x, Y0 + y, Z0 + z, G);
            long long* g = elem32(farGrid, cell);
            const unsigned long long v = s_tile[e] + (unsigned long long)*g;
            *g = 0;
            *elem32(rhoF, cell) = fix_to_float(v);
        }
        if (threadIdx.x == 0) tileFlag[t] = 0;
        __syncthreads();  // s_tile is cleared for the next tile
    }
}

void launch_deposit_gather(const float4* pos, const int2* blockRange, int G, int periodic, long long* rho,
                           int32_t* farList, StepScalars* sc, hipStream_t st) {
    const FarCtl f{farList, &sc->farCount, nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_deposit_gather<false>, dim3(G / GDX, G / GDY, G / GDZ), dim3(GTHREADS), 0, st, pos,
                       blockRange, G, periodic, rho, nullptr, f);
    hipLaunchKernelGGL(k_deposit_far, dim3(1), dim3(GTHREADS), 0, st, pos, G, periodic, rho, farList,
                       &sc->farCount);
}

int gather_deposit_tiles(int G) { return (G / GDX) * (G / GDY) * (G / GDZ); }

void launch_deposit_gather_f32(const float4* pos, const int2* blockRange, int G, int periodic, float* rhoF,
                               long long* farGrid, GatherFixBuffers& b, hipStream_t st) {
    int* cnt = b.fixCount + b.parity;
    int* cntNext = b.fixCount + (b.parity ^ 1);
    b.parity ^= 1;
    const FarCtl f{nullptr, nullptr, b.tileFlag, b.tileList, cnt, farGrid};
    hipLaunchKernelGGL(k_deposit_gather<true>, dim3(G / GDX, G / GDY, G / GDZ), dim3(GTHREADS), 0, st, pos,
                       blockRange, G, periodic, nullptr, rhoF, f);
    // a grid that walks the listed tiles (none on most steps: every block leaves at once)
    const int grid = std::min(gather_deposit_tiles(G), 256);
    hipLaunchKernelGGL(k_deposit_fix, dim3(grid), dim3(GTHREADS), 0, st, pos, blockRange, G, periodic, rhoF, farGrid,
                       cnt, b.tileFlag, b.tileList, cntNext);
}

// ---- gas physics: cooling and pressure (Sim.cu:1212-1415) ---------------------
// The reference sums each cell's velocities with float atomics (an order that
// changes run to run). Here particles are sorted by cell and the first particle
// of each cell's run sums the run serially in sorted order, so the sums are
// deterministic and equal a sequential CPU pass over the same order.
//
// k_cool_cells: kAccumCellVel + kCoolCell for one cell per run head. Mean
// velocity over the cell's particles that are not dark matter (vel.w >= -50,
// finite velocities); cells with fewer than two hold no dispersion and cool
// nothing. Each such particle adds (mean - v)^2 per axis to the cell's
// dispersion sums; gas particles (pos.w <= 0) are pulled toward the mean by
// k = min(rate * dt * 60, 0.5), dt the step's dt times the pass interval.
__global__ void __launch_bounds__(256) k_cool_cells(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                   const uint32_t* __restrict__ keys, int n, int G,
                                                   float rate, float dtMul, const StepScalars* __restrict__ sc,
                                                   float* __restrict__ dispX, float* __restrict__ dispY,
                                                   float* __restrict__ dispZ, float* __restrict__ dispCnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = keys[i];
    if (key == 0u || (i > 0 && keys[i - 1] == key)) return;  // inactive, or not the run's first
    int end = i + 1;
    while (end < n && keys[end] == key) ++end;
    float sx = 0.f, sy = 0.f, sz = 0.f, nc = 0.f;
    for (int j = i; j < end; ++j) {
        const float4 p = pos[j];
        const float4 v = vel[j];
        if (p.x < 0.f) continue;
        if (!isfinite(v.x) || !isfinite(v.y) || !isfinite(v.z)) continue;
        if (v.w < -50.f) continue;
        sx += v.x;
        sy += v.y;
        sz += v.z;
        nc += 1.0f;
    }
    if (nc < 2.f) return;
    const float4 p0 = pos[i];
    const float fG = (float)G;
    const int cx = min(max((int)(p0.x * fG), 0), G - 1);
    const int cy = min(max((int)(p0.y * fG), 0), G - 1);
    const int cz = min(max((int)(p0.z * fG), 0), G - 1);
    const int c = (cz * G + cy) * G + cx;
    const float inv = 1.f / nc;
    const float mvx = sx * inv, mvy = sy * inv, mvz = sz * inv;
    float k = rate * (sc->dt * dtMul) * 60.0f;
    if (k > 0.5f) k = 0.5f;
    float qx = 0.f, qy = 0.f, qz = 0.f, qc = 0.f;
    for (int j = i; j < end; ++j) {
        const float4 p = pos[j];
        const float4 v = vel[j];
        if (p.x < 0.f) continue;
        if (v.w < -50.f) continue;
        const float dvx = mvx - v.x, dvy = mvy - v.y, dvz = mvz - v.z;
        qx += dvx * dvx;
        qy += dvy * dvy;
        qz += dvz * dvz;
        qc += 1.0f;
        if (rate > 0.f && p.w <= 0.f) vel[j] = make_float4(fmaf(k, dvx, v.x), fmaf(k, dvy, v.y), fmaf(k, dvz, v.z), v.w);
    }
    dispX[c] = qx;
    dispY[c] = qy;
    dispZ[c] = qz;
    dispCnt[c] = qc;
}

void launch_cool_cells(const float4* pos, float4* vel, const uint32_t* keys, int n, int G, float rate, float dtMul,
                       const StepScalars* sc, float* disp, hipStream_t st) {
    if (n <= 0) return;
    const size_t cells = (size_t)G * G * G;
    hipLaunchKernelGGL(k_cool_cells, dim3((n + 255) / 256), dim3(256), 0, st, pos, vel, keys, n, G, rate, dtMul, sc,
                       disp, disp + cells, disp + 2 * cells, disp + 3 * cells);
}

// kPressure (Sim.cu:1368-1399): a = -(disp[+1] - disp[-1]) * (0.5 G k / n) per
// axis, the dispersion sum of that axis differenced along it, n the cell's
// count; clamped to +-maxAcc; zero where the cell holds no dispersion. Edges as
// kGridAccel (clamped stencil, or wrapped on a periodic mesh).
__global__ void k_pressure(const float* __restrict__ dispX, const float* __restrict__ dispY,
                           const float* __restrict__ dispZ, const float* __restrict__ dispCnt,
                           float4* __restrict__ accP, int G, int periodic, float k, float maxAcc) {
    const long long ci = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= (long long)G * G * G) return;
    const int x = (int)(ci % G), y = (int)((ci / G) % G), z = (int)(ci / ((long long)G * G));
    const float rho = dispCnt[ci];
    if (rho < 1e-6f) {
        accP[ci] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    int xm, xp, ym, yp, zm, zp;
    if (periodic) {
        xm = (x - 1 + G) % G; xp = (x + 1) % G;
        ym = (y - 1 + G) % G; yp = (y + 1) % G;
        zm = (z - 1 + G) % G; zp = (z + 1) % G;
    } else {
        xm = max(x - 1, 0); xp = min(x + 1, G - 1);
        ym = max(y - 1, 0); yp = min(y + 1, G - 1);
        zm = max(z - 1, 0); zp = min(z + 1, G - 1);
    }
    const float h = 0.5f * (float)G * k / rho;
    float ax = -(dispX[((size_t)z * G + y) * G + xp] - dispX[((size_t)z * G + y) * G + xm]) * h;
    float ay = -(dispY[((size_t)z * G + yp) * G + x] - dispY[((size_t)z * G + ym) * G + x]) * h;
    float az = -(dispZ[((size_t)zp * G + y) * G + x] - dispZ[((size_t)zm * G + y) * G + x]) * h;
    ax = fminf(fmaxf(ax, -maxAcc), maxAcc);
    ay = fminf(fmaxf(ay, -maxAcc), maxAcc);
    az = fminf(fmaxf(az, -maxAcc), maxAcc);
    accP[ci] = make_float4(ax, ay, az, 0.f);
}

void launch_pressure(const float* disp, float4* accP, int G, int periodic, float k, float maxAcc, hipStream_t st) {
    const long long cells = (long long)G * G * G;
    hipLaunchKernelGGL(k_pressure, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, disp, disp + cells,
                       disp + 2 * cells, disp + 3 * cells, accP, G, periodic, k, maxAcc);
}

// sampleAcc, Sim.cu:614-629.
__device__ __forceinline__ float3 sample_acc(const float4* __restrict__ accG, const Cic& c, int G, int periodic) {
    const int x0 = wrap_or_clamp(c.ix, G, periodic), x1 = wrap_or_clamp(c.ix + 1, G, periodic);
    const int y0 = wrap_or_clamp(c.iy, G, periodic), y1 = wrap_or_clamp(c.iy + 1, G, periodic);
    const int z0 = wrap_or_clamp(c.iz, G, periodic), z1 = wrap_or_clamp(c.iz + 1, G, periodic);
    float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? c.fx : 1.f - c.fx) * (oy ? c.fy : 1.f - c.fy) * (oz ? c.fz : 1.f - c.fz);
        const int x = ox ? x1 : x0, y = oy ? y1 : y0, z = oz ? z1 : z0;
        const float4 g = *elem32(accG, cell_idx(x, y, z, G));
        ax = fmaf(g.x, w, ax);
        ay = fmaf(g.y, w, ay);
        az = fmaf(g.z, w, az);
    }
    return make_float3(ax, ay, az);
}

// sampleAcc of the node accelerations that kGridAccel (Sim.cu:589-611) would
// write, taken from the potential itself: node (x, y, z) gets
// -(pot[.., x+1] - pot[.., x-1]) * h per axis with the gradient kernel's edge
// rule, in the same operations, so the result is bit-identical to sample_acc
// over that kernel's grid. Interior stencils (ix - 1 .. ix + 2 inside the mesh
// on every axis) load four x-rows of 4 values and eight of 2; the rest take
// the 48 single loads of the general form.
__device__ __forceinline__ float4 load4u(const float* p) {
    float4 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
}
__device__ __forceinline__ float2 load2u(const float* p) {
    float2 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
}
__device__ __forceinline__ float3 sample_acc_pot(const float* __restrict__ pot, const Cic& c, int G, int periodic,
                                                 float h) {
    float ax = 0.f, ay = 0.f, az = 0.f;
    const bool inner = c.ix >= 1 && c.ix + 2 <= G - 1 && c.iy >= 1 && c.iy + 2 <= G - 1 && c.iz >= 1 &&
                       c.iz + 2 <= G - 1;
    if (inner) {
        const unsigned g1 = (unsigned)G, g2 = (unsigned)G * (unsigned)G;
        const unsigned b = cell_idx(c.ix, c.iy, c.iz, G);  // every offset below stays >= 0 here
        float4 r[2][2];      // [oz][oy]: x = ix - 1 .. ix + 2
        float2 fy[2][2];     // [oz][0: y = iy - 1, 1: y = iy + 2]: x = ix, ix + 1
        float2 fz[2][2];     // [0: z = iz - 1, 1: z = iz + 2][oy]
#pragma unroll
        for (int oz = 0; oz < 2; ++oz)
#pragma unroll
            for (int oy = 0; oy < 2; ++oy) r[oz][oy] = load4u(elem32(pot, b + oz * g2 + oy * g1 - 1u));
#pragma unroll
        for (int oz = 0; oz < 2; ++oz) {
            fy[oz][0] = load2u(elem32(pot, b + oz * g2 - g1));
            fy[oz][1] = load2u(elem32(pot, b + oz * g2 + 2u * g1));
        }
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            fz[0][oy] = load2u(elem32(pot, b - g2 + oy * g1));
            fz[1][oy] = load2u(elem32(pot, b + 2u * g2 + oy * g1));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            const float w = (ox ? c.fx : 1.f - c.fx) * (oy ? c.fy : 1.f - c.fy) * (oz ? c.fz : 1.f - c.fz);
            const float4 R = r[oz][oy];
            const float xm = ox ? R.y : R.x, xp = ox ? R.w : R.z;
            const float4 Ro = r[oz][oy ^ 1];
            const float2 Fy = fy[oz][oy];
            const float yo = ox ? Ro.z : Ro.y;  // the other core row at this x
            const float yf = ox ? Fy.y : Fy.x;  // the face row at this x
            const float ym = oy ? yo : yf, yp = oy ? yf : yo;
            const float4 Zo = r[oz ^ 1][oy];
            const float2 Fz = fz[oz][oy];
            const float zo = ox ? Zo.z : Zo.y;
            const float zf = ox ? Fz.y : Fz.x;
            const float zm = oz ? zo : zf, zp = oz ? zf : zo;
            const float gx = -(xp - xm) * h, gy = -(yp - ym) * h, gz = -(zp - zm) * h;
            ax = fmaf(gx, w, ax);
            ay = fmaf(gy, w, ay);
            az = fmaf(gz, w, az);
        }
        return make_float3(ax, ay, az);
    }
    const int xs[2] = {wrap_or_clamp(c.ix, G, periodic), wrap_or_clamp(c.ix + 1, G, periodic)};
    const int ys[2] = {wrap_or_clamp(c.iy, G, periodic), wrap_or_clamp(c.iy + 1, G, periodic)};
    const int zs[2] = {wrap_or_clamp(c.iz, G, periodic), wrap_or_clamp(c.iz + 1, G, periodic)};
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? c.fx : 1.f - c.fx) * (oy ? c.fy : 1.f - c.fy) * (oz ? c.fz : 1.f - c.fz);
        const int x = xs[ox], y = ys[oy], z = zs[oz];
        int xm, xp, ym, yp, zm, zp;
        if (periodic) {
            xm = x == 0 ? G - 1 : x - 1; xp = x == G - 1 ? 0 : x + 1;
            ym = y == 0 ? G - 1 : y - 1; yp = y == G - 1 ? 0 : y + 1;
            zm = z == 0 ? G - 1 : z - 1; zp = z == G - 1 ? 0 : z + 1;
        } else {
            xm = max(x - 1, 0); xp = min(x + 1, G - 1);
            ym = max(y - 1, 0); yp = min(y + 1, G - 1);
            zm = max(z - 1, 0); zp = min(z + 1, G - 1);
        }
        const float gx = -(*elem32(pot, cell_idx(xp, y, z, G)) - *elem32(pot, cell_idx(xm, y, z, G))) * h;
        const float gy = -(*elem32(pot, cell_idx(x, yp, z, G)) - *elem32(pot, cell_idx(x, ym, z, G))) * h;
        const float gz = -(*elem32(pot, cell_idx(x, y, zp, G)) - *elem32(pot, cell_idx(x, y, zm, G))) * h;
        ax = fmaf(gx, w, ax);
        ay = fmaf(gy, w, ay);
        az = fmaf(gz, w, az);
    }
    return make_float3(ax, ay, az);
}

// ---- the potential box of a block in LDS (sparse meshes) --------------------------
// Under half a particle per cell, neighbouring particles' stencil rows rarely
// share cache lines, so the global gather fetches the potential several times
// over. Instead a block copies the box its particles' stencils cover (nodes
// ix - 1 .. ix + 2 of every live particle, x widened to whole float4 lines)
// into LDS once, and gathers from there with the same operations. The x origin
// and extent are multiples of 4, so each float4 of the copy is one aligned
// global load and one aligned LDS store, and (G a multiple of 4) never crosses
// the periodic boundary. Values outside an isolated mesh are clamped copies:
// only particles whose stencil lies inside the mesh read the box there.
#ifndef SD_BOX_CAP
#define SD_BOX_CAP 9216
#endif
#ifndef SD_BOX_GATHER
#define SD_BOX_GATHER 1
#endif
constexpr int BOX_CAP = SD_BOX_CAP;  // floats (36 KB: four blocks per CU); the deposit's tile reuses the memory

struct PotBox {
    int x0, y0, z0;  // origin (node coordinates, unwrapped)
    int dx, dxy;     // row and plane pitch (floats)
    bool ok;         // the block has live particles and its box fits BOX_CAP (uniform)
};

// Exact quotient q = t / d for 0 <= t < 2^15, d >= 1: the float estimate is within
// one of it, and one correction each way fixes it.
__device__ __forceinline__ int quot_small(int t, int d, float invD) {
    int q = (int)(((float)t + 0.5f) * invD);
    if (__mul24(q, d) > t) --q;
    if (__mul24(q + 1, d) <= t) ++q;
    return q;
}

// Every thread of the block calls it (it synchronises).
__device__ __forceinline__ PotBox load_pot_box(const bool alive[KP], const Cic c[KP], const float* __restrict__ pot,
                                               int G, int periodic, float* __restrict__ box) {
    __shared__ int s_bb[PT / 64][6];
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        if (alive[k]) {
            lo[0] = min(lo[0], c[k].ix); hi[0] = max(hi[0], c[k].ix);
            lo[1] = min(lo[1], c[k].iy); hi[1] = max(hi[1], c[k].iy);
            lo[2] = min(lo[2], c[k].iz); hi[2] = max(hi[2], c[k].iz);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_min63(lo[d]);
        hi[d] = wave_max63(hi[d]);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) {
        for (int d = 0; d < 3; ++d) {
            s_bb[w][d] = lo[d];
            s_bb[w][3 + d] = hi[d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PT / 64; ++q) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], s_bb[q][d]);
            hi[d] = max(hi[d], s_bb[q][3 + d]);
        }
    }
    PotBox b;
    b.ok = false;
    b.x0 = b.y0 = b.z0 = b.dx = b.dxy = 0;
    if (lo[0] == INT_MAX) return b;  // no live particle (uniform)
    // x from a multiple of 4 at or below ix - 1 to one above ix + 2 (floor division)
    b.x0 = (lo[0] - 1) & ~3;
    const int x1 = ((hi[0] + 3) + 3) & ~3;
    b.y0 = lo[1] - 1;
    b.z0 = lo[2] - 1;
    const long long dx = x1 - b.x0, dy = (long long)hi[1] - lo[1] + 4, dz = (long long)hi[2] - lo[2] + 4;
    if (dx * dy * dz > BOX_CAP) return b;
    // box coordinates in [-G, 2G): the wrap below needs no division
    if (b.x0 < -G || x1 > 2 * G || b.y0 < -G || hi[1] + 3 >= 2 * G || b.z0 < -G || hi[2] + 3 >= 2 * G) return b;
    b.ok = true;
    b.dx = (int)dx;
    b.dxy = (int)(dx * dy);
    const int nx4 = (int)dx >> 2, rows = (int)(dy * dz), units = nx4 * rows;
    const float inv4 = 1.0f / (float)nx4, invY = 1.0f / (float)dy;
    const unsigned g1 = (unsigned)G;
    for (int u = threadIdx.x; u < units; u += PT) {
        const int row = quot_small(u, nx4, inv4), x4 = u - __mul24(row, nx4);
        const int z = quot_small(row, (int)dy, invY), y = row - __mul24(z, (int)dy);
        const int gx = wrap_or_clamp1(b.x0 + 4 * x4, G, periodic), gy = wrap_or_clamp1(b.y0 + y, G, periodic),
                  gz = wrap_or_clamp1(b.z0 + z, G, periodic);
        float4 v;
        if (periodic || (b.x0 + 4 * x4 >= 0 && b.x0 + 4 * x4 + 3 < G)) {
            v = *reinterpret_cast<const float4*>(elem32(pot, __umul24(__umul24((unsigned)gz, g1) + (unsigned)gy, g1) +
                                                             (unsigned)gx));
        } else {  // an isolated mesh's edge: clamp each x
            const unsigned rb = __umul24(__umul24((unsigned)gz, g1) + (unsigned)gy, g1);
            float e[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) e[q] = *elem32(pot, rb + (unsigned)wrap_or_clamp1(b.x0 + 4 * x4 + q, G, 0));
            v = make_float4(e[0], e[1], e[2], e[3]);
        }
        *reinterpret_cast<float4*>(box + __mul24(row, (int)dx) + 4 * x4) = v;
    }
    __syncthreads();
    return b;
}

// sample_acc_pot's interior form on the LDS box (same loads, same operations).
__device__ __forceinline__ float3 sample_acc_box(const float* __restrict__ box, const PotBox& b, const Cic& c,
                                                 float h) {
    float ax = 0.f, ay = 0.f, az = 0.f;
    const int g1 = b.dx, g2 = b.dxy;
    const int base = __mul24(c.iz - b.z0, g2) + __mul24(c.iy - b.y0, g1) + (c.ix - b.x0);
    float4 r[2][2];   // [oz][oy]: x = ix - 1 .. ix + 2
    float2 fy[2][2];  // [oz][0: y = iy - 1, 1: y = iy + 2]: x = ix, ix + 1
    float2 fz[2][2];  // [0: z = iz - 1, 1: z = iz + 2][oy]
#pragma unroll
    for (int oz = 0; oz < 2; ++oz)
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
            const float* q = box + base + oz * g2 + oy * g1 - 1;
            r[oz][oy] = make_float4(q[0], q[1], q[2], q[3]);
        }
#pragma unroll
    for (int oz = 0; oz < 2; ++oz) {
        const float* q0 = box + base + oz * g2 - g1;
        const float* q1 = box + base + oz * g2 + 2 * g1;
        fy[oz][0] = make_float2(q0[0], q0[1]);
        fy[oz][1] = make_float2(q1[0], q1[1]);
    }
#pragma unroll
    for (int oy = 0; oy < 2; ++oy) {
        const float* q0 = box + base - g2 + oy * g1;
        const float* q1 = box + base + 2 * g2 + oy * g1;
        fz[0][oy] = make_float2(q0[0], q0[1]);
        fz[1][oy] = make_float2(q1[0], q1[1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? c.fx : 1.f - c.fx) * (oy ? c.fy : 1.f - c.fy) * (oz ? c.fz : 1.f - c.fz);
        const float4 R = r[oz][oy];
        const float xm = ox ? R.y : R.x, xp = ox ? R.w : R.z;
        const float4 Ro = r[oz][oy ^ 1];
        const float2 Fy = fy[oz][oy];
        const float yo = ox ? Ro.z : Ro.y;
        const float yf = ox ? Fy.y : Fy.x;
        const float ym = oy ? yo : yf, yp = oy ? yf : yo;
        const float4 Zo = r[oz ^ 1][oy];
        const float2 Fz = fz[oz][oy];
        const float zo = ox ? Zo.z : Zo.y;
        const float zf = ox ? Fz.y : Fz.x;
        const float zm = oz ? zo : zf, zp = oz ? zf : zo;
        const float gx = -(xp - xm) * h, gy = -(yp - ym) * h, gz = -(zp - zm) * h;
        ax = fmaf(gx, w, ax);
        ay = fmaf(gy, w, ay);
        az = fmaf(gz, w, az);
    }
    return make_float3(ax, ay, az);
}

// Block reduction of (sum v, count, max |v|) into one partial per block.
__device__ __forceinline__ void block_partial(double sx, double sy, double sz, double cnt, float vmax,
                                              ReducePartial* out, int blk) {
    __shared__ double s_d[PT / 64][4];
    __shared__ float s_f[PT / 64];
    vmax = vmax >= 0.f ? vmax : 0.f;  // a NaN speed counts as 0, as fmaxf would drop it
#if SD_DPP_REDUCE
    sx = wave_sum63(sx);
    sy = wave_sum63(sy);
    sz = wave_sum63(sz);
    cnt = wave_sum63(cnt);
    vmax = wave_max63(vmax);
    constexpr int LEAD = 63;
#else
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    cnt = wave_sum(cnt);
    vmax = wave_max(vmax);
    constexpr int LEAD = 0;
#endif
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == LEAD) {
        s_d[w][0] = sx;
        s_d[w][1] = sy;
        s_d[w][2] = sz;
        s_d[w][3] = cnt;
        s_f[w] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ReducePartial r;
        r.sx = r.sy = r.sz = r.cnt = 0.0;
        r.vmax = 0.f;
        for (int k = 0; k < PT / 64; ++k) {
            r.sx += s_d[k][0];
            r.sy += s_d[k][1];
            r.sz += s_d[k][2];
            r.cnt += s_d[k][3];
            r.vmax = fmaxf(r.vmax, s_f[k]);
        }
        r.pad[0] = r.pad[1] = r.pad[2] = 0.f;
        out[blk] = r;
    }
}

__device__ __forceinline__ float speed3(float x, float y, float z) {
    return sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
}

// kIntegrate, Sim.cu:685-906 (no black holes). Each thread advances KP
// particles (i = block * BP + k * PT + thread, coalesced for every k); the new
// positions feed the block deposit when depositTo is set. GAS adds the gas
// pressure gather (a.accPress); it is a separate instantiation because its
// registers would cost the plain step a wave per SIMD (79 against 71 VGPRs).
template <bool GAS, bool BOX = false>
__device__ __forceinline__ void integrate_body(const IntegrateArgs& a, int blk) {
    double sx = 0.0, sy = 0.0, sz = 0.0, cnt = 0.0;
    float vm2 = 0.f;  // the largest squared speed: sqrt is monotone and correctly rounded, so
                      // sqrt(max v^2) is the largest speed3() bit for bit
    bool depAlive[KP];
    Cic depC[KP];
    // phase 1: every particle load of this thread in flight together
    float4 P[KP], V[KP];
    bool act[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int i = blk * BP + k * PT + threadIdx.x;
        act[k] = i < a.n;
        if (a.order) {
            const int j = act[k] ? a.order[i] : 0;
            P[k] = act[k] ? load_pv(a.srcPos + j) : make_float4(-1.f, -1.f, -1.f, 0.f);
            V[k] = act[k] ? load_pv(a.srcVel + j) : make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            P[k] = act[k] ? load_pv(a.pos + i) : make_float4(-1.f, -1.f, -1.f, 0.f);
            V[k] = act[k] ? load_pv(a.vel + i) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    // phase 2: the CIC gathers (8 x KP loads, no stores in between)
    const float fG = (float)a.G;
    float3 ACC[KP];
    // BOX: the potential box of the block in LDS; the deposit's tile reuses it
    __shared__ __attribute__((aligned(16))) unsigned long long s_boxmem[BOX ? BOX_CAP / 2 : 2];
    if constexpr (BOX) {
        float* box = reinterpret_cast<float*>(s_boxmem);
        bool live[KP];
        Cic cc[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            live[k] = P[k].x >= 0.f;
            cc[k] = cic_setup(P[k], fG);
        }
        const PotBox pb = load_pot_box(live, cc, a.pot, a.G, a.periodic, box);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            ACC[k] = make_float3(0.f, 0.f, 0.f);
            if (!live[k]) continue;
            const Cic& c = cc[k];
            const bool inner = c.ix >= 1 && c.ix + 2 <= a.G - 1 && c.iy >= 1 && c.iy + 2 <= a.G - 1 && c.iz >= 1 &&
                               c.iz + 2 <= a.G - 1;
            ACC[k] = pb.ok && (a.periodic || inner) ? sample_acc_box(box, pb, c, a.gradH)
                                                    : sample_acc_pot(a.pot, c, a.G, a.periodic, a.gradH);
        }
    } else {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            ACC[k] = make_float3(0.f, 0.f, 0.f);
            if (P[k].x >= 0.f)
                ACC[k] = a.pot ? sample_acc_pot(a.pot, cic_setup(P[k], fG), a.G, a.periodic, a.gradH)
                               : sample_acc(a.accG, cic_setup(P[k], fG), a.G, a.periodic);
        }
    }
    const float kick = a.sc->kick, drift = a.sc->dt;
    // phase 3: kick, drift, boundaries, stores
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        depAlive[k] = false;
        depC[k].ix = depC[k].iy = depC[k].iz = 0;
        depC[k].fx = depC[k].fy = depC[k].fz = 0.f;
        const int i = blk * BP + k * PT + threadIdx.x;
        if (!act[k]) continue;
        float4 p = P[k];
        if (p.x >= 0.f) {
            float4 v = V[k];
            if (a.applyMeanSub) {
                v.x -= a.sc->meanSub[0];
                v.y -= a.sc->meanSub[1];
                v.z -= a.sc->meanSub[2];
            }
            float3 acc = ACC[k];
            if (a.accContact) {
                const float4 c = a.accContact[i];
                acc.x += c.x;
                acc.y += c.y;
                acc.z += c.z;
            }
            if (a.accCorr) {
                const float4 c = a.accCorr[i];
                acc.x += c.x;
                acc.y += c.y;
                acc.z += c.z;
            }
            if (a.darkEnergy != 0.f) {
                float k = 1.0f;
                if (a.softBoundR > 0.f) {
                    const float dx = p.x - 0.5f, dy = p.y - 0.5f, dz = p.z - 0.5f;
                    const float r = speed3(dx, dy, dz);
                    if (r > a.softBoundR) {
                        const float R1 = 0.497f;
                        k = 1.0f - fminf((r - a.softBoundR) / fmaxf(R1 - a.softBoundR, 1e-6f), 1.0f);
                    }
                }
                if (k > 0.f) {
                    acc.x = fmaf(a.darkEnergy * k, p.x - 0.5f, acc.x);
                    acc.y = fmaf(a.darkEnergy * k, p.y - 0.5f, acc.y);
                    acc.z = fmaf(a.darkEnergy * k, p.z - 0.5f, acc.z);
                }
            }
            // gas pressure (Sim.cu:740-743): gas particles only (w <= 0), not dark matter
            if (GAS && p.w <= 0.f && v.w >= -50.f) {
                const float3 pr = sample_acc(a.accPress, cic_setup(p, fG), a.G, a.periodic);
                acc.x += pr.x;
                acc.y += pr.y;
                acc.z += pr.z;
            }
            v.x = fmaf(acc.x, kick, v.x);
            v.y = fmaf(acc.y, kick, v.y);
            v.z = fmaf(acc.z, kick, v.z);
            {
                const float sp2 = fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x));
                if (sp2 > a.c2) {
                    const float k = sqrtf(a.c2 / sp2);
                    v.x *= k;
                    v.y *= k;
                    v.z *= k;
                }
            }
            p.x = fmaf(v.x, drift, p.x);
            p.y = fmaf(v.y, drift, p.y);
            p.z = fmaf(v.z, drift, p.z);
            bool alive = true;
            if (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z) || !isfinite(v.x) || !isfinite(v.y) ||
                !isfinite(v.z)) {
                p = make_float4(-1.f, -1.f, -1.f, 0.f);
                v = make_float4(0.f, 0.f, 0.f, 0.f);
                alive = false;
            } else if (a.periodic) {
                p.x -= floorf(p.x);
                p.y -= floorf(p.y);
                p.z -= floorf(p.z);
            } else if (a.softBoundR > 0.f) {
                const float dx = p.x - 0.5f, dy = p.y - 0.5f, dz = p.z - 0.5f;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float R1 = 0.497f;
                if (r2 > a.softBoundR * a.softBoundR) {
                    const float r = sqrtf(r2);
                    const float inv = 1.0f / fmaxf(r, 1e-6f);
                    const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
                    const float t = fminf((r - a.softBoundR) / fmaxf(R1 - a.softBoundR, 1e-6f), 1.0f);
                    const float vr = fmaf(v.z, nz, fmaf(v.y, ny, v.x * nx));
                    if (vr > 0.f) {
                        const float cut = vr * t * t;
                        v.x = fmaf(-nx, cut, v.x);
                        v.y = fmaf(-ny, cut, v.y);
                        v.z = fmaf(-nz, cut, v.z);
                    }
                    const float pull = 2.0f * t * t;
                    v.x = fmaf(-(nx * pull), drift, v.x);
                    v.y = fmaf(-(ny * pull), drift, v.y);
                    v.z = fmaf(-(nz * pull), drift, v.z);
                    if (r > R1) {
                        const float k = R1 * inv;
                        p.x = fmaf(dx, k, 0.5f);
                        p.y = fmaf(dy, k, 0.5f);
                        p.z = fmaf(dz, k, 0.5f);
                    }
                }
            } else {
                if (p.x < 0.002f) { p.x = 0.002f; v.x = fabsf(v.x) * 0.25f; }
                if (p.x > 0.998f) { p.x = 0.998f; v.x = -fabsf(v.x) * 0.25f; }
                if (p.y < 0.002f) { p.y = 0.002f; v.y = fabsf(v.y) * 0.25f; }
                if (p.y > 0.998f) { p.y = 0.998f; v.y = -fabsf(v.y) * 0.25f; }
                if (p.z < 0.002f) { p.z = 0.002f; v.z = fabsf(v.z) * 0.25f; }
                if (p.z > 0.998f) { p.z = 0.998f; v.z = -fabsf(v.z) * 0.25f; }
            }
            store_pv(a.pos + i, p);
            store_pv(a.vel + i, v);
            if (alive) {
                depAlive[k] = true;
                depC[k] = cic_setup(p, fG);
                sx += (double)v.x;
                sy += (double)v.y;
                sz += (double)v.z;
                cnt += 1.0;
                vm2 = fmaxf(vm2, fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x)));
            }
        } else if (a.order) {
            // an inactive slot is left untouched in place, but a fused reorder
            // writes another buffer: carry it over
            store_pv(a.pos + i, p);
            store_pv(a.vel + i, V[k]);
        }
    }
    if constexpr (BOX) {
        static_assert(BOX_CAP / 2 >= DEP_CAP, "the deposit tile fits the box memory");
        if (a.depositTo) block_deposit_in(depAlive, depC, a.G, a.periodic, a.depositTo, s_boxmem);
    } else {
        if (a.depositTo) block_deposit(depAlive, depC, a.G, a.periodic, a.depositTo);
    }
    if (a.partials) block_partial(sx, sy, sz, cnt, sqrtf(vm2), a.partials, blk);
}

// Consecutive particle tiles (Morton-sorted: neighbouring stencils) run on one
// XCD, so the potential lines their gathers share stay in that XCD's L2.
template <bool GAS, bool BOX = false>
__global__ void __launch_bounds__(PT) k_integrate(IntegrateArgs a) {
    integrate_body<GAS, BOX>(a, SD_XCD_REMAP ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x);
}

// Integrate of many simulations in one launch: block b belongs to job j with
// prefix[j] <= b < prefix[j + 1] and runs that job's block b - prefix[j], so
// every job sees exactly its own single-launch decomposition.
__global__ void __launch_bounds__(PT) k_integrate_batch(const IntegrateArgs* __restrict__ jobs,
                                                        const int* __restrict__ prefix, int njobs) {
    const int j = find_job(prefix, njobs, (int)blockIdx.x);
    const IntegrateArgs a = jobs[j];
    integrate_body<false>(a, (int)blockIdx.x - prefix[j]);
}

int integrate_blocks(int n) { return (n + BP - 1) / BP; }
int particle_blocks(int n) { return (n + PT - 1) / PT; }

// evStart / evStop (optional): events the dispatch itself records at the
// kernel's start and end (no marker packets of their own in the queue).
void launch_integrate(const IntegrateArgs& a, hipStream_t st, hipEvent_t evStart, hipEvent_t evStop) {
    if (a.n <= 0) return;
    const dim3 grid(integrate_blocks(a.n)), block(PT);
    // the LDS box gather serves the potential gather of sparse meshes (no gas term)
    const bool box = SD_BOX_GATHER && a.boxGather && a.pot && !a.accPress;
    if (evStart && evStop) {
        if (a.accPress)
            hipExtLaunchKernelGGL(k_integrate<true>, grid, block, 0, st, evStart, evStop, 0, a);
        else if (box)
            hipExtLaunchKernelGGL((k_integrate<false, true>), grid, block, 0, st, evStart, evStop, 0, a);
        else
            hipExtLaunchKernelGGL(k_integrate<false>, grid, block, 0, st, evStart, evStop, 0, a);
        return;
    }
    if (a.accPress)
        hipLaunchKernelGGL(k_integrate<true>, grid, block, 0, st, a);
    else if (box)
        hipLaunchKernelGGL((k_integrate<false, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_integrate<false>, grid, block, 0, st, a);
}

// kMaxSpeed (Sim.cu:2923-2930) as block partials.
__global__ void __launch_bounds__(PT) k_vmax(const float4* __restrict__ pos, const float4* __restrict__ vel, int n,
                                            ReducePartial* partials) {
    const int i = blockIdx.x * PT + threadIdx.x;
    float vm = 0.f;
    if (i < n && pos[i].x >= 0.f) {
        const float4 v = vel[i];
        vm = speed3(v.x, v.y, v.z);
    }
    block_partial(0.0, 0.0, 0.0, 0.0, vm, partials, blockIdx.x);
}

// kMomentumAccum (Sim.cu:2252-2268) as block partials.
__global__ void __launch_bounds__(PT) k_momentum(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                int n, ReducePartial* partials) {
    const int i = blockIdx.x * PT + threadIdx.x;
    double sx = 0.0, sy = 0.0, sz = 0.0, cnt = 0.0;
    if (i < n && pos[i].x >= 0.f) {
        const float4 v = vel[i];
        if (isfinite(v.x) && isfinite(v.y) && isfinite(v.z)) {
            sx = (double)v.x;
            sy = (double)v.y;
            sz = (double)v.z;
            cnt = 1.0;
        }
    }
    block_partial(sx, sy, sz, cnt, 0.f, partials, blockIdx.x);
}

void launch_vmax(const float4* pos, const float4* vel, int n, ReducePartial* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_vmax, dim3(particle_blocks(n > 0 ? n : 1)), dim3(PT), 0, st, pos, vel, n, partials);
}
void launch_momentum(const float4* pos, const float4* vel, int n, ReducePartial* partials, hipStream_t st) {
    hipLaunchKernelGGL(k_momentum, dim3(particle_blocks(n > 0 ? n : 1)), dim3(PT), 0, st, pos, vel, n, partials);
}

__global__ void __launch_bounds__(FIN) k_finalize(const ReducePartial* __restrict__ parts, int np, int mode,
                                                  DtPolicy pol, StepScalars* sc) {
    finalize_body<FIN>(parts, np, mode, pol, sc);
}

// One block per simulation.
__global__ void __launch_bounds__(FIN) k_finalize_batch(const FinalizeJob* __restrict__ jobs) {
    const FinalizeJob j = jobs[blockIdx.x];
    finalize_body<FIN>(j.partials, j.nparts, j.mode, j.pol, j.sc);
}

void launch_finalize(const ReducePartial* partials, int nparts, int mode, DtPolicy pol, StepScalars* sc,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN), 0, st, partials, nparts, mode, pol, sc);
}

void launch_integrate_batch(const IntegrateArgs* jobs, const int* prefix, int njobs, int totalBlocks,
                            hipStream_t st) {
    if (totalBlocks <= 0) return;
    hipLaunchKernelGGL(k_integrate_batch, dim3(totalBlocks), dim3(PT), 0, st, jobs, prefix, njobs);
}

void launch_finalize_batch(const FinalizeJob* jobs, int njobs, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_finalize_batch, dim3(njobs), dim3(FIN), 0, st, jobs);
}

__global__ void __launch_bounds__(256) k_copy_f4(const float4* __restrict__ src, float4* __restrict__ dst,
                                                 long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_copy_f4(const float4* src, float4* dst, long long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_f4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}

// kSubtractMeanVelDev (Sim.cu:2282-2291); the pending mean is cleared after use
// by the last block to finish (a device-wide counter).
__global__ void __launch_bounds__(PT) k_subtract_mean(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                     int n, StepScalars* sc) {
    const int i = blockIdx.x * PT + threadIdx.x;
    const float mx = sc->meanSub[0], my = sc->meanSub[1], mz = sc->meanSub[2];
    if (i < n && pos[i].x >= 0.f) {
        float4 v = vel[i];
        v.x -= mx;
        v.y -= my;
        v.z -= mz;
        vel[i] = v;
    }
}
__global__ void k_clear_mean(StepScalars* sc) {
    sc->meanSub[0] = sc->meanSub[1] = sc->meanSub[2] = 0.f;
}
void launch_subtract_mean(const float4* pos, float4* vel, int n, StepScalars* sc, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_subtract_mean, dim3(particle_blocks(n)), dim3(PT), 0, st, pos, vel, n, sc);
    hipLaunchKernelGGL(k_clear_mean, dim3(1), dim3(1), 0, st, sc);
}

// Velocity read-out: leapfrog state holds v(n-1/2); synchronise with half a
// kick of the current mesh force. KD state is returned as stored. A pending
// momentum correction is applied in both cases.
__global__ void __launch_bounds__(PT) k_sync_velocity(const float4* __restrict__ accG, const float4* __restrict__ pos,
                                                     const float4* __restrict__ vel, int n, int G, int periodic,
                                                     const StepScalars* sc, int halfKick, float4* __restrict__ out) {
    const int i = blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    float4 v = vel[i];
    if (p.x >= 0.f) {
        v.x -= sc->meanSub[0];
        v.y -= sc->meanSub[1];
        v.z -= sc->meanSub[2];
        if (halfKick && !sc->kdkOpen) {
            const float h = 0.5f * sc->dtPrev;
            const float3 a = sample_acc(accG, cic_setup(p, (float)G), G, periodic);
            v.x = fmaf(a.x, h, v.x);
            v.y = fmaf(a.y, h, v.y);
            v.z = fmaf(a.z, h, v.z);
        }
    }
    out[i] = v;
}
void launch_sync_velocity(const float4* accG, const float4* pos, const float4* vel, int n, int G, int periodic,
                          const StepScalars* sc, float4* velOut, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sync_velocity, dim3(particle_blocks(n)), dim3(PT), 0, st, accG, pos, vel, n, G, periodic,
                       sc, accG ? 1 : 0, velOut);
}

// ---- contact: kContact, Sim.cu:1030-1203 (contact term) ---------------------------
// x, y, z in [-1, G] (a neighbour of a cell inside the mesh).
__device__ __forceinline__ int cell_slot(int x, int y, int z, int G, int periodic, int keyMode) {
    x = wrap_or_clamp1(x, G, periodic);
    y = wrap_or_clamp1(y, G, periodic);
    z = wrap_or_clamp1(z, G, periodic);
    return keyMode ? (int)morton3((uint32_t)x, (uint32_t)y, (uint32_t)z) : (z * G + y) * G + x;
}

// Squared distance bound (in cells) along one axis from a particle at offset o
// (cell units, [0, 1)) inside cell c to the cells c - 1, c, c + 1: every
// particle keyed to cell c - 1 lies at least o away, to c + 1 at least 1 - o.
// A neighbour that folds onto another cell (the clamped edge of an isolated
// box) gets a large negative value, so that no sum including it passes the
// pruning test.
__device__ __forceinline__ void axis_gaps(float o, int c, int G, int periodic, float g[3]) {
    const float lo = fmaxf(o, 0.f), hi = fmaxf(1.f - o, 0.f);
    g[0] = (periodic || c > 0) ? lo * lo : -1e30f;
    g[1] = 0.f;
    g[2] = (periodic || c < G - 1) ? hi * hi : -1e30f;
}

// Distinct blocks (at most 3) holding the cells c - 1, c, c + 1 along one axis.
__device__ __forceinline__ int axis_blocks(int c, int G, int periodic, int b[3]) {
    const int b0 = wrap_or_clamp1(c - 1, G, periodic) >> 2, b1 = c >> 2, b2 = wrap_or_clamp1(c + 1, G, periodic) >> 2;
    int m = 0;
    b[m++] = b0;
    if (b1 != b0) b[m++] = b1;
    if (b2 != b0 && b2 != b1) b[m++] = b2;
    return m;
}

// Sorted indices [first, end) of the particles with key k (cell k - 1) inside
// the run br of that cell's 4^3-cell block (Morton keys ascend, so the cell's
// particles are one sub-run of the block's run): a binary search for the
// first key >= k, then, if that key is k, one for the first key > k. Every
// probe lies in [br.x, br.y).
__device__ __forceinline__ int2 run_of_key(const uint32_t* __restrict__ keys, int2 br, uint32_t k) {
    int a = br.x, len = br.y - br.x;
    while (len > 0) {
        const int h = len >> 1;
        if (keys[a + h] < k) {
            a += h + 1;
            len -= h + 1;
        } else {
            len = h;
        }
    }
    if (a >= br.y || keys[a] != k) return make_int2(a, a);
    int e = a + 1;
    len = br.y - e;
    while (len > 0) {
        const int h = len >> 1;
        if (keys[e + h] == k) {
            e += h + 1;
            len -= h + 1;
        } else {
            len = h;
        }
    }
    return make_int2(a, e);
}

// One cell's candidates [s0, e0) in the reference's order (kContact, Sim.cu:1063-1100):
// each counts toward the cap of 96 (self excluded), in range or not.
__device__ __forceinline__ void contact_cell(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                             const int32_t* __restrict__ order, int i, float4 p, float4 v, int s0,
                                             int e0, int periodic, float d0, float d02, float stiffness,
                                             float damping, int& seen, float& ax, float& ay, float& az) {
    constexpr int kMaxPeers = 96;
    for (int j = s0; j < e0 && seen < kMaxPeers; ++j) {
        if (j == i) continue;
        ++seen;
        const int sj = order ? order[j] : j;  // storage slot of sorted particle j
        const float4 q = pos[sj];
        if (q.x < 0.f) continue;
        float ex = p.x - q.x, ey = p.y - q.y, ez = p.z - q.z;
        if (periodic) {
            if (ex > 0.5f) ex -= 1.f; else if (ex < -0.5f) ex += 1.f;
            if (ey > 0.5f) ey -= 1.f; else if (ey < -0.5f) ey += 1.f;
            if (ez > 0.5f) ez -= 1.f; else if (ez < -0.5f) ez += 1.f;
        }
        const float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
        if (dd >= d02 || dd < 1e-16f) continue;
        const float d = sqrtf(dd);
        const float nx = ex / d, ny = ey / d, nz = ez / d;
        const float4 vj = vel[sj];
        const float vn = fmaf(v.z - vj.z, nz, fmaf(v.y - vj.y, ny, (v.x - vj.x) * nx));
        float f = 0.f;
        if (d < d0) {
            float fc = fmaf(stiffness, d0 - d, -(damping * vn));
            if (fc < 0.f) fc = 0.f;
            f += fc;
        }
        ax = fmaf(f, nx, ax);
        ay = fmaf(f, ny, ay);
        az = fmaf(f, nz, az);
    }
}

// The reference visits the 27 cells around the particle's cell in z, y, x
// order and stops after 96 candidates (self excluded; every candidate counts,
// in range or not). A cell whose nearest point lies beyond the contact
// diameter plus 1e-3 cell holds only particles that fail the distance test;
// the margin is far above the float rounding of keys, offsets and distances
// (about 3e-5 cell at G = 512).
// A cell's particles come from the cell table (row-major keys, range[c]) or,
// with Morton keys (BLOCKS), from a search inside its 4^3-cell block's run of
// the sorted keys (blockRange[c >> 6]); there is then no per-cell table.
// Two visits give exactly the reference's forces:
// * when the blocks covering the 27 cells hold at most 97 particles (the
//   particle itself and at most 96 candidates), the cap cannot be reached and
//   out-of-reach cells are skipped outright (Morton keys only);
// * otherwise the nine entries of a z-plane load together and an out-of-reach
//   cell adds only its count to the tally (capped at 96 as the reference's
//   loop caps it).
// Isolated boxes take the second visit at their edge cells, where the clamped
// neighbour folds onto the particle's own cell (periodic meshes below 3 cells
// likewise).
// order (optional): the sort's permutation, sorted particle j living in slot
// order[j] of pos / vel (the arrays not yet reordered); accOut is in sorted order.
template <bool BLOCKS>
__global__ void __launch_bounds__(PT) k_contact(const float4* __restrict__ pos, const float4* __restrict__ vel,
                                               const int32_t* __restrict__ order, int n, int G, int periodic,
                                               const int2* __restrict__ range, const uint32_t* __restrict__ keys,
                                               const int2* __restrict__ blockRange, long long blockEntries,
                                               float radius, float stiffness, float damping,
                                               float4* __restrict__ accOut) {
    // the blocks' cell-occupancy words follow the block runs (launch_cell_range)
    const unsigned long long* occ = reinterpret_cast<const unsigned long long*>(blockRange + blockEntries);
    // consecutive (sorted: neighbouring) particles on one XCD share candidates in its L2
    const int i = (SD_XCD_REMAP ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x) * PT + threadIdx.x;
    if (i >= n) return;
    const int si = order ? order[i] : i;
    const float4 p = pos[si];
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (p.x < 0.f) {
        accOut[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4 v = vel[si];
    const float d0 = radius * 2.0f;
    const float d02 = d0 * d0;
    const float fG = (float)G;
    const int cx = min(max((int)(p.x * fG), 0), G - 1);
    const int cy = min(max((int)(p.y * fG), 0), G - 1);
    const int cz = min(max((int)(p.z * fG), 0), G - 1);
    float gx[3], gy[3], gz[3];
    axis_gaps(p.x * fG - (float)cx, cx, G, periodic, gx);
    axis_gaps(p.y * fG - (float)cy, cy, G, periodic, gy);
    axis_gaps(p.z * fG - (float)cz, cz, G, periodic, gz);
    const float reach = fmaf(d0, fG, 1e-3f);  // contact diameter in cells, plus the margin
    const float reach2 = reach * reach;
    const int kMaxPeers = 96;
    const int keyMode = BLOCKS ? 1 : 0;
    int seen = 0;

    bool sparse = BLOCKS && (periodic ? G >= 3 : (cx > 0 && cx < G - 1 && cy > 0 && cy < G - 1 && cz > 0 && cz < G - 1));
    if (sparse) {
        int bx[3], by[3], bz[3];
        const int mx = axis_blocks(cx, G, periodic, bx), my = axis_blocks(cy, G, periodic, by),
                  mz = axis_blocks(cz, G, periodic, bz);
        // Morton codes of the blocks from their per-axis parts
        uint32_t sx[3], sy[3], sz[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            sx[q] = morton_spread10((uint32_t)bx[q]);
            sy[q] = morton_spread10((uint32_t)by[q]) << 1;
            sz[q] = morton_spread10((uint32_t)bz[q]) << 2;
        }
        int total = 0;
        for (int c = 0; c < mz; ++c)
            for (int b = 0; b < my; ++b)
                for (int a = 0; a < mx; ++a) {
                    const int2 r = blockRange[sx[a] | sy[b] | sz[c]];
                    total += r.y - r.x;
                }
        sparse = total <= kMaxPeers + 1;
    }
    if (sparse) {
        // Morton codes of the 27 cells from the per-axis parts of the three
        // neighbouring coordinates (wrapped on a periodic mesh; an isolated
        // mesh takes this path only inside, where no neighbour is clamped)
        // the particle's own cell holds only itself unless a sorted neighbour shares its key
        const uint32_t ownKey = keys[i];
        const bool alone = (i == 0 || keys[i - 1] != ownKey) && (i + 1 >= n || keys[i + 1] != ownKey);
        uint32_t px[3], py[3], pz[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            px[q] = morton_spread10((uint32_t)wrap_or_clamp1(cx + q - 1, G, periodic));
            py[q] = morton_spread10((uint32_t)wrap_or_clamp1(cy + q - 1, G, periodic)) << 1;
            pz[q] = morton_spread10((uint32_t)wrap_or_clamp1(cz + q - 1, G, periodic)) << 2;
        }
        // bit q = 9 (dz + 1) + 3 (dy + 1) + (dx + 1): the cell is in reach and occupied
        // (the 27 occupancy loads issue together; q ascending is the visiting order)
        uint32_t visit = 0u;
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
            if (gx[dx + 1] + gy[dy + 1] + gz[dz + 1] > reach2) continue;
            if (alone && q == 13) continue;  // the own cell holds no candidate but itself
            const uint32_t c = px[dx + 1] | py[dy + 1] | pz[dz + 1];
            if ((occ[c >> 6] >> (c & 63u)) & 1ull) visit |= 1u << q;  // an empty cell: no search
        }
        while (visit) {
            const int q = __builtin_ctz(visit);
            visit &= visit - 1u;
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
            const int c = (int)(px[dx + 1] | py[dy + 1] | pz[dz + 1]);
            const int2 br = blockRange[c >> 6];
            if (br.y <= br.x) continue;
            const int2 r = run_of_key(keys, br, (uint32_t)c + 1u);
            if (r.y <= r.x) continue;
            contact_cell(pos, vel, order, i, p, v, r.x, r.y, periodic, d0, d02, stiffness, damping, seen, ax, ay, az);
        }
    } else {
        for (int dz = -1; dz <= 1 && seen < kMaxPeers; ++dz) {
            int cs[9];
            int2 r[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                cs[q] = cell_slot(cx + q % 3 - 1, cy + q / 3 - 1, cz + dz, G, periodic, keyMode);
                r[q] = BLOCKS ? blockRange[cs[q] >> 6] : range[cs[q]];
            }
            if (!BLOCKS) {
                // a table entry counts only when its first index holds the cell's key
                uint32_t kf[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) kf[q] = (r[q].x >= 0 && r[q].x < n) ? keys[r[q].x] : 0u;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    if (kf[q] != (uint32_t)cs[q] + 1u) r[q] = make_int2(0, 0);  // empty cell (stale or unwritten)
            }
            for (int q = 0; q < 9 && seen < kMaxPeers; ++q) {
                if (r[q].y <= r[q].x) continue;
                const int2 cr = BLOCKS ? run_of_key(keys, r[q], (uint32_t)cs[q] + 1u) : r[q];
                const int s0 = cr.x, e0 = cr.y;
                if (e0 <= s0) continue;
                if (gx[q % 3] + gy[q / 3] + gz[dz + 1] > reach2) {
                    seen = min(seen + (e0 - s0), kMaxPeers);  // out of reach: the particle is not in this cell
                    continue;
                }
                contact_cell(pos, vel, order, i, p, v, s0, e0, periodic, d0, d02, stiffness, damping, seen, ax, ay, az);
            }
        }
    }
    accOut[i] = make_float4(ax, ay, az, 0.f);
}

void launch_contact(const float4* pos, const float4* vel, const int32_t* order, int n, int G, int periodic,
                    int keyMode, const int2* cellRange, const uint32_t* sortedKeys, const int2* blockRange,
                    long long blockEntries, float radius, float stiffness, float damping, float4* accOut,
                    hipStream_t st) {
    if (n <= 0) return;
    if (keyMode)
        hipLaunchKernelGGL(k_contact<true>, dim3(particle_blocks(n)), dim3(PT), 0, st, pos, vel, order, n, G,
                           periodic, nullptr, sortedKeys, blockRange, blockEntries, radius, stiffness, damping,
                           accOut);
    else
        hipLaunchKernelGGL(k_contact<false>, dim3(particle_blocks(n)), dim3(PT), 0, st, pos, vel, order, n, G,
                           periodic, cellRange, sortedKeys, nullptr, 0ll, radius, stiffness, damping, accOut);
}

// ---- direct sum (method of proto/cuda/x3.cu:84-103, 3D, Plummer) -------------------
constexpr int DS_CHUNK = 16384;  // sources per block

__global__ void __launch_bounds__(256) k_direct_partial(const float4* __restrict__ pos, int n,
                                                       const int32_t* __restrict__ targets, int nt, float eps2,
                                                       double* __restrict__ partial) {
    __shared__ float4 s_src[256];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int chunk = blockIdx.y;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < nt) {
        // a target outside [0, n) gets NaN accelerations (the host entry point
        // rejects such targets; device-side target lists are checked here)
        const int ti = targets[t];
        p = (ti >= 0 && ti < n) ? pos[ti] : make_float4(__builtin_nanf(""), 0.f, 0.f, 0.f);
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    const int j0 = chunk * DS_CHUNK, j1 = min(n, j0 + DS_CHUNK);
    for (int base = j0; base < j1; base += 256) {
        const int j = base + threadIdx.x;
        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);  // w = 0 -> no contribution
        if (j < j1) {
            q = pos[j];
            q.w = (q.x < 0.f) ? 0.f : 1.f;
        }
        s_src[threadIdx.x] = q;
        __syncthreads();
        const int lim = min(256, j1 - base);
        for (int k = 0; k < lim; ++k) {
            const float4 s = s_src[k];
            const float dx = s.x - p.x, dy = s.y - p.y, dz = s.z - p.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            const float inv = rsqrtf(r2);
            const float f = s.w * inv * inv * inv;
            ax = fmaf(dx, f, ax);
            ay = fmaf(dy, f, ay);
            az = fmaf(dz, f, az);
        }
        __syncthreads();
    }
    if (t < nt) {
        double* o = partial + ((size_t)chunk * nt + t) * 3;
        o[0] = ax;
        o[1] = ay;
        o[2] = az;
    }
}

__global__ void k_direct_reduce(const double* __restrict__ partial, int nt, int nchunks, double gm,
                                double* __restrict__ out3) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < nchunks; ++c)
        for (int q = 0; q < 3; ++q) s[q] += partial[((size_t)c * nt + t) * 3 + q];
    for (int q = 0; q < 3; ++q) out3[(size_t)t * 3 + q] = gm * s[q];
}

int direct_chunks_for(int n) { return (n + DS_CHUNK - 1) / DS_CHUNK; }

void launch_direct_sum(const float4* pos, int n, const int32_t* targets, int nt, float eps2, double* partial,
                       int nchunks, double* out3, double gm, hipStream_t st) {
    if (nt <= 0) return;
    hipLaunchKernelGGL(k_direct_partial, dim3((nt + 255) / 256, nchunks), dim3(256), 0, st, pos, n, targets, nt,
                       eps2, partial);
    hipLaunchKernelGGL(k_direct_reduce, dim3((nt + 255) / 256), dim3(256), 0, st, partial, nt, nchunks, gm, out3);
}

}  // namespace sd
