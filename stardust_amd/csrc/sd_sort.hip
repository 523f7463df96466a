// sd_sort.hip -- cell keys, stable LSD radix sort, reorder and cell ranges.
//
// Replaces Impl::sortParticles (src/sim/Sim.cu:3656-3671): kCellKey
// (2476-2486), cub::DeviceRadixSort::SortPairs (3660-3667) and kReorder
// (2488-2495); and the contact cell table kFillInt + kBuildCellRange
// (1004-1011, 3689-3691).
//
// Key modes: 0 row-major cell (kCellKey), 1 Morton index of the cell
// (contact hash grid), 2 Morton index of the CIC base node floor(p G - 0.5)
// clamped to [0, G-1] (step order: particles sharing a deposit / gather
// stencil are contiguous).
// Keys are unsigned "reference key + 1" (inactive -> 0), so an unsigned sort
// reproduces CUB's signed order (the -1 keys first) and only
// ceil(log2(maxKey + 1)) bits are sorted. Each 8-bit pass is three launches:
//   hist    : per-tile digit counts (wavefront ballot match -> one LDS add per
//             run of equal digits) written digit-major; for the first pass
//             the key kernel counts them (k_cell_keys_hist),
//   scan    : exclusive scan of each digit's row of tile counts, a block per
//             digit, and the digit totals,
//   scatter : digit bases from the totals, stable in-tile ranks (same ballot
//             match, running per-digit counters across rounds and waves) ->
//             global positions.
#include <cmath>

#include "sd_device.h"
#include "sd_internal.h"

namespace sd {

constexpr int SORT_THREADS = 256;
// Keys per thread per tile (R): 8 up to SORT_SMALL_N keys, 16 above (4.2 M
// keys: 2048-key tiles 1% faster per step than 4096; 16.7 M keys: 0.4%
// slower, 1024-key tiles 3% slower).
constexpr int SORT_ROUNDS_SMALL = 8, SORT_ROUNDS_LARGE = 16;
constexpr int SORT_SMALL_N = 1 << 23;
static int sort_rounds_for(int n) { return n <= SORT_SMALL_N ? SORT_ROUNDS_SMALL : SORT_ROUNDS_LARGE; }
constexpr int SCAN_CHUNK = 1024;  // hist entries per scan block
constexpr int WAVES = SORT_THREADS / 64;

// Key of one particle: cell (mode 0 row-major, 1 Morton) or CIC base node
// (mode 2, Morton), plus one; 0 for an inactive slot.
__device__ __forceinline__ uint32_t cell_key(float4 p, int G, int mode) {
    if (p.x < 0.f) return 0u;
    const float fG = (float)G;
    int cx, cy, cz;
    if (mode == 2) {
        // CIC base node floor(p G - 0.5) (the deposit / gather stencil origin)
        cx = min(max((int)floorf(fmaf(p.x, fG, -0.5f)), 0), G - 1);
        cy = min(max((int)floorf(fmaf(p.y, fG, -0.5f)), 0), G - 1);
        cz = min(max((int)floorf(fmaf(p.z, fG, -0.5f)), 0), G - 1);
    } else {
        cx = min(max((int)(p.x * fG), 0), G - 1);
        cy = min(max((int)(p.y * fG), 0), G - 1);
        cz = min(max((int)(p.z * fG), 0), G - 1);
    }
    uint32_t k = mode ? morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz)
                      : (uint32_t)((cz * G + cy) * G + cx);
    return k + 1u;
}

// Keys of many particle sets in one launch (see KeyJob): the key of set j is
// its own cell key with the set's tag above it; idx is the global index.
__global__ void __launch_bounds__(256) k_cell_keys_batch(const float4* __restrict__ pos,
                                                         const KeyJob* __restrict__ jobs,
                                                         const int* __restrict__ prefix, int njobs,
                                                         uint32_t* __restrict__ keys,
                                                         int32_t* __restrict__ idx) {
    const int j = find_job(prefix, njobs, (int)blockIdx.x);
    const KeyJob J = jobs[j];
    const int li = ((int)blockIdx.x - prefix[j]) * 256 + (int)threadIdx.x;
    if (li >= J.n) return;
    const int i = J.base + li;
    idx[i] = i;
    keys[i] = J.tag | cell_key(pos[i], J.G, J.mode);
}

// Lanes of this wavefront whose digit equals mine (8-bit digits).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    return m;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <int R>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                            uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_THREADS * R;
    // every key load of the thread in flight at once, then the counting
    uint32_t kk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + threadIdx.x;
        kk[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + threadIdx.x;
        const bool valid = i < n;
        const uint32_t d = (kk[r] >> shift) & 255u;
        const uint64_t active = __ballot(valid);
        if (valid) {
            const uint64_t m = match_digit(d, active);
            const int lane = threadIdx.x & 63;
            // highest lane of the group adds the group size
            if ((m >> lane) == 1ull) atomicAdd(&cnt[d], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Keys and identity values of one sort tile, plus the tile's histogram of the
// first radix digit (what k_sort_hist would count for shift 0).
template <int R>
__global__ void __launch_bounds__(SORT_THREADS) k_cell_keys_hist(const float4* __restrict__ pos, int n, int G,
                                                                 int mode, uint32_t* __restrict__ keys,
                                                                 int32_t* __restrict__ idx,
                                                                 uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_THREADS * R;
    // every position load of the thread in flight at once, then keys and counts
    float4 pp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + threadIdx.x;
        pp[r] = i < n ? pos[i] : make_float4(-1.f, -1.f, -1.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + threadIdx.x;
        const bool valid = i < n;
        uint32_t d = 0u;
        if (valid) {
            const uint32_t k = cell_key(pp[r], G, mode);
            keys[i] = k;
            idx[i] = i;
            d = k & 255u;
        }
        const uint64_t active = __ballot(valid);
        if (valid) {
            const uint64_t m = match_digit(d, active);
            if ((m >> (threadIdx.x & 63)) == 1ull) atomicAdd(&cnt[d], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Block-wide exclusive scan of SCAN_CHUNK entries (4 per thread), returns total.
__device__ uint32_t block_scan_excl(uint32_t* v4, uint32_t* s_wave) {
    uint32_t loc = v4[0] + v4[1] + v4[2] + v4[3];
    // inclusive wave scan
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = loc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for (int k = 0; k < WAVES; ++k) {
        if (k < w) wbase += s_wave[k];
        total += s_wave[k];
    }
    __syncthreads();
    uint32_t run = wbase + x - loc;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t t = v4[k];
        v4[k] = run;
        run += t;
    }
    return total;
}

// Exclusive scan of each digit's row of tile counts (hist is digit-major:
// hist[d * ntiles + tile]), in place, one block per digit; totals[d] = the
// digit's key count. The scatter adds the digit's base (the exclusive scan of
// totals) itself.
__global__ void __launch_bounds__(256) k_scan_digit_rows(uint32_t* __restrict__ h, int ntiles,
                                                         uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_wave[WAVES];
    uint32_t* row = h + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int off = 0; off < ntiles; off += SCAN_CHUNK) {  // uniform
        const int base = off + threadIdx.x * 4;
        uint32_t v4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v4[k] = (base + k < ntiles) ? row[base + k] : 0u;
        const uint32_t t = block_scan_excl(v4, s_wave);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (base + k < ntiles) row[base + k] = v4[k] + carry;
        carry += t;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

#ifndef SD_SORT_LOCAL
#define SD_SORT_LOCAL 1
#endif

// Scatter of one tile. A digit's first output slot is its base (exclusive scan
// of the digit totals, done by every block) plus the tile's entry in the
// digit's scanned row. Every key's rank within its digit in the tile comes
// from the ballot match and running per-digit counters. LOCAL first orders the
// tile by digit in LDS (the tile's digit counts are the differences of
// consecutive entries of the digit's scanned row), so that each digit's
// keys leave as one contiguous run instead of one store per key per round:
// that pays when a round's 256 keys spread over many digits (sparse keys), and
// costs LDS traffic and occupancy when equal digits already come in long runs.
template <bool LOCAL, int R>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(const uint32_t* __restrict__ keys,
                                                               const int32_t* __restrict__ vals, int n, int shift,
                                                               const uint32_t* __restrict__ offs,
                                                               const uint32_t* __restrict__ totals, int ntiles,
                                                               uint32_t* __restrict__ keysOut,
                                                               int32_t* __restrict__ valsOut) {
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_wcnt[WAVES][256];
    __shared__ uint32_t s_wbase[WAVES][256];
    __shared__ uint32_t s_wave[WAVES];
    const int t = threadIdx.x;
    const size_t flat = (size_t)t * ntiles + blockIdx.x;
    const uint32_t loc = offs[flat], tot = totals[t];
    {
        uint32_t v4[4] = {tot, 0u, 0u, 0u};
        block_scan_excl(v4, s_wave);
        s_off[t] = v4[0] + loc;
    }
    s_run[t] = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s_wcnt[w][t] = 0;
    __shared__ uint32_t s_start[LOCAL ? 256 : 1];  // first tile slot of each digit
    __shared__ uint32_t s_lk[LOCAL ? SORT_THREADS * R : 1];
    __shared__ int32_t s_lv[LOCAL ? SORT_THREADS * R : 1];
    if constexpr (LOCAL) {
        const uint32_t next = (int)blockIdx.x + 1 < ntiles ? offs[flat + 1] : tot;
        uint32_t v4[4] = {next - loc, 0u, 0u, 0u};
        block_scan_excl(v4, s_wave);
        s_start[t] = v4[0];
    }
    __syncthreads();
    const int base = blockIdx.x * SORT_THREADS * R;
    const int lane = t & 63, wv = t >> 6;
    // every key and value load of the thread in flight at once
    uint32_t kk[R];
    int32_t vv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        kk[r] = 0u;
        vv[r] = 0;
        if (i < n) {
            kk[r] = keys[i];
            vv[r] = vals[i];
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        const bool valid = i < n;
        const uint32_t key = kk[r];
        const int32_t val = vv[r];
        const uint32_t d = (key >> shift) & 255u;
        const uint64_t active = __ballot(valid);
        uint32_t rank = 0;
        if (valid) {
            const uint64_t m = match_digit(d, active);
            rank = (uint32_t)__popcll(m & lanemask_lt());
            if ((m >> lane) == 1ull) s_wcnt[wv][d] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        {
            uint32_t run = s_run[t];
#pragma unroll
            for (int w = 0; w < WAVES; ++w) {
                s_wbase[w][t] = run;
                run += s_wcnt[w][t];
                s_wcnt[w][t] = 0;
            }
            s_run[t] = run;
        }
        __syncthreads();
        if (valid) {
            if constexpr (LOCAL) {
                const uint32_t slot = s_start[d] + s_wbase[wv][d] + rank;
                s_lk[slot] = key;
                s_lv[slot] = val;
            } else {
                const uint32_t dst = s_off[d] + s_wbase[wv][d] + rank;
                keysOut[dst] = key;
                valsOut[dst] = val;
            }
        }
        // s_wbase is rewritten next round only after the barrier that follows
        // the next round's counting, so no extra barrier is needed here.
    }
    if constexpr (LOCAL) {
        __syncthreads();
        const int inTile = min(SORT_THREADS * R, n - base);
        for (int e = t; e < inTile; e += SORT_THREADS) {
            const uint32_t key = s_lk[e];
            const uint32_t d = (key >> shift) & 255u;
            const uint32_t dst = s_off[d] + (uint32_t)e - s_start[d];
            keysOut[dst] = key;
            valsOut[dst] = s_lv[e];
        }
    }
}

__global__ void __launch_bounds__(256) k_reorder(const float4* __restrict__ srcP, const float4* __restrict__ srcV,
                                                 const int32_t* __restrict__ order, float4* __restrict__ dstP,
                                                 float4* __restrict__ dstV, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = order[i];
    dstP[i] = srcP[j];
    dstV[i] = srcV[j];
}

__global__ void __launch_bounds__(256) k_cell_range(const uint32_t* __restrict__ keys, int n,
                                                    int2* __restrict__ range, int2* __restrict__ blockRange) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (k == 0u) return;
    const int c = (int)(k - 1u);
    if (range) {
        const bool first = i == 0 || keys[i - 1] != k, last = i == n - 1 || keys[i + 1] != k;
        if (first && last)
            range[c] = make_int2(i, i + 1);  // a cell of one particle: one 8-byte store
        else if (first)
            range[c].x = i;
        else if (last)
            range[c].y = i + 1;
    }
    if (blockRange) {
        // keys ascend, so keys[i + 1] >= k > 0; an inactive predecessor (key 0) starts a run
        const uint32_t b = (uint32_t)c >> 6;
        const bool bfirst = i == 0 || keys[i - 1] == 0u || ((keys[i - 1] - 1u) >> 6) != b;
        const bool blast = i == n - 1 || ((keys[i + 1] - 1u) >> 6) != b;
        if (bfirst && blast)
            blockRange[b] = make_int2(i, i + 1);
        else if (bfirst)
            blockRange[b].x = i;
        else if (blast)
            blockRange[b].y = i + 1;
    }
}

// ---------------------------------------------------------------------------

void launch_cell_keys_batch(const float4* pos, const KeyJob* jobs, const int* prefix, int njobs,
                            int totalBlocks, uint32_t* keys, int32_t* idx, hipStream_t st) {
    if (totalBlocks <= 0) return;
    hipLaunchKernelGGL(k_cell_keys_batch, dim3(totalBlocks), dim3(256), 0, st, pos, jobs, prefix, njobs, keys, idx);
}

int sort_tiles_for(int n) {
    const int tile = SORT_THREADS * sort_rounds_for(n);
    return (n + tile - 1) / tile;
}
int sort_chunks_for(int) { return 256; }  // the digit totals

void launch_cell_keys_hist(const float4* pos, int n, int G, int mode, SortBuffers& buf, hipStream_t st) {
    if (n <= 0) return;
    const int ntiles = sort_tiles_for(n);
    if (sort_rounds_for(n) == SORT_ROUNDS_SMALL)
        hipLaunchKernelGGL(k_cell_keys_hist<SORT_ROUNDS_SMALL>, dim3(ntiles), dim3(SORT_THREADS), 0, st, pos, n, G,
                           mode, buf.keys, buf.idx, buf.hist, ntiles);
    else
        hipLaunchKernelGGL(k_cell_keys_hist<SORT_ROUNDS_LARGE>, dim3(ntiles), dim3(SORT_THREADS), 0, st, pos, n, G,
                           mode, buf.keys, buf.idx, buf.hist, ntiles);
}

template <int R>
static void radix_passes(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady);

void radix_sort_pairs(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady) {
    if (n <= 1) return;
    if (sort_rounds_for(n) == SORT_ROUNDS_SMALL)
        radix_passes<SORT_ROUNDS_SMALL>(buf, n, bits, st, firstHistReady);
    else
        radix_passes<SORT_ROUNDS_LARGE>(buf, n, bits, st, firstHistReady);
}

template <int R>
static void radix_passes(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady) {
    if (n <= 1) return;
    const int ntiles = sort_tiles_for(n);
    // Keys sparse in their code space (under one key per two codes) spread each
    // round over many digits, and so do the lowest digits of any keys: order
    // the tile in LDS first there, except in the last pass, whose few digit
    // values come in long runs. (G=512 cosmic web, 16.7 M
    // keys of 27 bits: 214/342/365/86 us per pass direct, 136/133/128/104
    // ordered; two-disk G=128, 4.2 M keys of 21 bits: 43/24/23 direct,
    // 38/34/33 ordered.)
    const bool sparse = SD_SORT_LOCAL && 2.0 * (double)n < std::ldexp(1.0, bits);
    for (int shift = 0; shift < bits; shift += 8) {
        if (shift > 0 || !firstHistReady)
            hipLaunchKernelGGL(k_sort_hist<R>, dim3(ntiles), dim3(SORT_THREADS), 0, st, buf.keys, n, shift,
                               buf.hist, ntiles);
        hipLaunchKernelGGL(k_scan_digit_rows, dim3(256), dim3(256), 0, st, buf.hist, ntiles, buf.chunk);
        if (SD_SORT_LOCAL && (sparse || shift == 0) && shift + 8 < bits)
            hipLaunchKernelGGL((k_sort_scatter<true, R>), dim3(ntiles), dim3(SORT_THREADS), 0, st, buf.keys, buf.idx, n,
                               shift, buf.hist, buf.chunk, ntiles, buf.keysAlt, buf.idxAlt);
        else
            hipLaunchKernelGGL((k_sort_scatter<false, R>), dim3(ntiles), dim3(SORT_THREADS), 0, st, buf.keys, buf.idx, n,
                               shift, buf.hist, buf.chunk, ntiles, buf.keysAlt, buf.idxAlt);
        uint32_t* tk = buf.keys;
        buf.keys = buf.keysAlt;
        buf.keysAlt = tk;
        int32_t* ti = buf.idx;
        buf.idx = buf.idxAlt;
        buf.idxAlt = ti;
    }
}

void launch_reorder(const float4* srcP, const float4* srcV, const int32_t* order, float4* dstP,
                    float4* dstV, int n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_reorder, dim3((n + 255) / 256), dim3(256), 0, st, srcP, srcV, order, dstP, dstV, n);
}

void launch_cell_range(const uint32_t* keys, int n, int2* range, int2* blockRange, long long blockEntries,
                       hipStream_t st) {
    if (blockRange) launch_fill_i32(reinterpret_cast<int32_t*>(blockRange), 2 * blockEntries, 0, st);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_range, dim3((n + 255) / 256), dim3(256), 0, st, keys, n, range, blockRange);
}

}  // namespace sd
