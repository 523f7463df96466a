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
#include <algorithm>
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
#ifndef SD_SORT_DEV_GRID
#define SD_SORT_DEV_GRID 512
#endif
constexpr int SORT_DEV_GRID = SD_SORT_DEV_GRID;  // blocks of a sort on a device count
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

// Adds one round of digits (one per lane) to the block's LDS counts. Equal
// digits in consecutive lanes form a run, and each run's first lane adds the
// run's length: the keys are nearly sorted, so the higher digits come in long
// runs (and the lowest ones mostly alone, with few same-address atomics).
// Invalid lanes form runs of their own and add nothing.
__device__ __forceinline__ void count_digit_runs(uint32_t d, bool valid, uint32_t* cnt) {
    const int lane = threadIdx.x & 63;
    const uint32_t dv = valid ? d : 0xFFFFFFFFu;
    // the previous lane's digit (wave_shr 1); lane 0 gets ~dv, so it heads a run
    const uint32_t dp = (uint32_t)__builtin_amdgcn_update_dpp((int)~dv, (int)dv, 0x138, 0xF, 0xF, false);
    const bool head = dv != dp;
    const uint64_t H = __ballot(head);
    if (head && valid) {
        const uint64_t above = lane == 63 ? 0ull : H >> (lane + 1);
        const int len = above ? __builtin_ctzll(above) + 1 : 64 - lane;
        atomicAdd(&cnt[d], (uint32_t)len);
    }
}

// Element count of a sort: the launch's n, or a count on the device (nDev) when
// the list was built on the device; launches are then sized for the largest
// count, and tiles beyond the device count leave at once.
__device__ __forceinline__ int sort_count(int n, const uint32_t* nDev) { return nDev ? (int)*nDev : n; }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <int R>
__device__ __forceinline__ void sort_hist_tile(const uint32_t* __restrict__ keys, int n, int shift,
                                               uint32_t* __restrict__ hist, int ntiles, int tile) {
    __shared__ uint32_t cnt[256];
    const int base = tile * SORT_THREADS * R;
    cnt[threadIdx.x] = 0;
    __syncthreads();
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
        count_digit_runs((kk[r] >> shift) & 255u, valid, cnt);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
}

// A block per tile; with a count on the device (nDev), a grid of fewer blocks
// walks the tiles in use.
template <int R>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                            uint32_t* __restrict__ hist, int ntiles,
                                                            const uint32_t* __restrict__ nDev) {
    if (!nDev) {
        sort_hist_tile<R>(keys, n, shift, hist, ntiles, blockIdx.x);
        return;
    }
    n = (int)*nDev;
    const int used = (n + SORT_THREADS * R - 1) / (SORT_THREADS * R);
    for (int tile = blockIdx.x; tile < used; tile += gridDim.x) {  // uniform
        sort_hist_tile<R>(keys, n, shift, hist, ntiles, tile);
        __syncthreads();  // the tile's counts are read out before the next clears them
    }
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
        count_digit_runs(d, valid, cnt);
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
__global__ void __launch_bounds__(256) k_scan_digit_rows(uint32_t* __restrict__ h, int stride,
                                                         uint32_t* __restrict__ totals,
                                                         const uint32_t* __restrict__ nDev, int tile) {
    __shared__ uint32_t s_wave[WAVES];
    // rows hold `stride` tile entries; the tiles in use cover the sort's count
    const int ntiles = nDev ? (int)((*nDev + (uint32_t)tile - 1u) / (uint32_t)tile) : stride;
    uint32_t* row = h + (size_t)blockIdx.x * stride;
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
__device__ __forceinline__ void sort_scatter_tile(const uint32_t* __restrict__ keys,
                                                  const int32_t* __restrict__ vals, int n, int shift,
                                                  const uint32_t* __restrict__ offs,
                                                  const uint32_t* __restrict__ totals, int ntiles,
                                                  uint32_t* __restrict__ keysOut, int32_t* __restrict__ valsOut,
                                                  int tile, int usedTiles) {
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_wcnt[WAVES][256];
    __shared__ uint32_t s_wbase[WAVES][256];
    __shared__ uint32_t s_wave[WAVES];
    const int t = threadIdx.x;
    const size_t flat = (size_t)t * ntiles + tile;  // ntiles: the row stride of offs
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
        const uint32_t next = tile + 1 < usedTiles ? offs[flat + 1] : tot;
        uint32_t v4[4] = {next - loc, 0u, 0u, 0u};
        block_scan_excl(v4, s_wave);
        s_start[t] = v4[0];
    }
    __syncthreads();
    const int base = tile * SORT_THREADS * R;
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

// A block per tile; with a count on the device (nDev), a grid of fewer blocks
// walks the tiles in use.
template <bool LOCAL, int R>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(const uint32_t* __restrict__ keys,
                                                               const int32_t* __restrict__ vals, int n, int shift,
                                                               const uint32_t* __restrict__ offs,
                                                               const uint32_t* __restrict__ totals, int ntiles,
                                                               uint32_t* __restrict__ keysOut,
                                                               int32_t* __restrict__ valsOut,
                                                               const uint32_t* __restrict__ nDev) {
    if (!nDev) {
        sort_scatter_tile<LOCAL, R>(keys, vals, n, shift, offs, totals, ntiles, keysOut, valsOut, blockIdx.x, ntiles);
        return;
    }
    n = (int)*nDev;
    const int used = (n + SORT_THREADS * R - 1) / (SORT_THREADS * R);
    for (int tile = blockIdx.x; tile < used; tile += gridDim.x) {  // uniform
        sort_scatter_tile<LOCAL, R>(keys, vals, n, shift, offs, totals, ntiles, keysOut, valsOut, tile, used);
        __syncthreads();  // the tile's LDS is read out before the next tile writes it
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

// occ (with blockRange): one bit per cell of each 4^3-cell block (Morton order
// inside the block), set when the cell holds a particle. A wavefront's keys are
// sorted, so the lanes of one block are consecutive: an OR scan over them leaves
// the block's bits in its last lane, which adds them with one atomic.
// Keys per thread (CR_R chunks of 256 consecutive keys per block, each
// chunk's loads issued before any is used).
#ifndef SD_CR_R
#define SD_CR_R 4
#endif
constexpr int CR_R = SD_CR_R;

__device__ __forceinline__ void cell_range_one(const uint32_t* __restrict__ keys, int n, int i, uint32_t k,
                                               int2* __restrict__ range, int2* __restrict__ blockRange,
                                               unsigned long long* __restrict__ occ) {
    const int c = (int)(k - 1u);
    if (occ) {  // every lane of the wavefront takes part in the scan
        const bool valid = k != 0u;
        unsigned long long v = valid ? 1ull << (c & 63) : 0ull;
        const uint32_t bk = valid ? (uint32_t)c >> 6 : 0xFFFFFFFFu;
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long u = __shfl_up(v, o, 64);
            const uint32_t bo = __shfl_up(bk, o, 64);
            if (lane >= o && bo == bk) v |= u;
        }
        const uint32_t bn = __shfl_down(bk, 1, 64);
        if (valid && (lane == 63 || bn != bk)) atomicOr(&occ[bk], v);
    }
    if (i >= n || k == 0u) return;
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

__global__ void __launch_bounds__(256) k_cell_range(const uint32_t* __restrict__ keys, int n,
                                                    int2* __restrict__ range, int2* __restrict__ blockRange,
                                                    unsigned long long* __restrict__ occ) {
    const int base = blockIdx.x * (256 * CR_R) + (int)threadIdx.x;
    uint32_t k[CR_R];
#pragma unroll
    for (int r = 0; r < CR_R; ++r) {
        const int i = base + r * 256;
        k[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < CR_R; ++r) cell_range_one(keys, n, base + r * 256, k[r], range, blockRange, occ);
}

// ---- sort of nearly sorted keys ------------------------------------------------
// Between two step sorts a particle moves a fraction of a cell, so most keys
// do not change. Let prev be the previous sort's key array (any ascending
// array of n keys). The slots whose new key equals prev[i] ("stayers") are in
// ascending (key, index) order already: their keys ascend with the index. The
// others ("movers") are gathered in index order and sorted by a stable radix
// sort, so they too are in (key, index) order; a merge path of the two lists by
// (key, index) is then the stable sort of all n keys, whatever prev held. Only
// the speed depends on prev matching the particles.
constexpr int NS_R = 8;                       // keys per thread per tile
constexpr int NS_TILE = SORT_THREADS * NS_R;  // 2048
constexpr int MERGE_T = 2048;                 // outputs per merge block (8 per thread)

// New keys into nk and each tile's mover count.
__global__ void __launch_bounds__(SORT_THREADS) k_keys_movers(const float4* __restrict__ pos, int n, int G, int mode,
                                                              const uint32_t* __restrict__ prev,
                                                              uint32_t* __restrict__ nk,
                                                              uint32_t* __restrict__ tileMovers) {
    __shared__ uint32_t s_w[WAVES];
    const int base = blockIdx.x * NS_TILE, t = threadIdx.x;
    float4 pp[NS_R];
    uint32_t pk[NS_R];
#pragma unroll
    for (int r = 0; r < NS_R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        pp[r] = i < n ? pos[i] : make_float4(-1.f, -1.f, -1.f, 0.f);
        pk[r] = i < n ? prev[i] : 0u;
    }
    int cnt = 0;
#pragma unroll
    for (int r = 0; r < NS_R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        if (i < n) {
            const uint32_t k = cell_key(pp[r], G, mode);
            nk[i] = k;
            cnt += k != pk[r] ? 1 : 0;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((t & 63) == 0) s_w[t >> 6] = (uint32_t)cnt;
    __syncthreads();
    if (t == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < WAVES; ++w) sum += s_w[w];
        tileMovers[blockIdx.x] = sum;
    }
}

// Exclusive scan of the tile mover counts in place (one block); counts[0] =
// movers, counts[1] = stayers.
__global__ void __launch_bounds__(1024) k_scan_tile_movers(uint32_t* __restrict__ tileMovers, int ntiles, int n,
                                                           uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_w[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t carry = 0;
    for (int off = 0; off < ntiles; off += 1024) {  // uniform
        const int i = off + t;
        const uint32_t v = i < ntiles ? tileMovers[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            if (k < w) wb += s_w[k];
            tot += s_w[k];
        }
        if (i < ntiles) tileMovers[i] = carry + wb + x - v;
        carry += tot;
        __syncthreads();
    }
    if (t == 0) {
        counts[0] = carry;
        counts[1] = (uint32_t)n - carry;
    }
}

// Stayers and movers of each tile to their lists, both in index order.
__global__ void __launch_bounds__(SORT_THREADS) k_split_movers(const uint32_t* __restrict__ nk,
                                                               const uint32_t* __restrict__ prev, int n,
                                                               const uint32_t* __restrict__ tileBase,
                                                               uint32_t* __restrict__ sKeys, int32_t* __restrict__ sIdx,
                                                               uint32_t* __restrict__ mKeys,
                                                               int32_t* __restrict__ mIdx) {
    __shared__ uint32_t s_w[NS_R][WAVES];
    const int base = blockIdx.x * NS_TILE, t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t kk[NS_R], pk[NS_R];
#pragma unroll
    for (int r = 0; r < NS_R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        kk[r] = i < n ? nk[i] : 0u;
        pk[r] = i < n ? prev[i] : 0u;
    }
    bool mv[NS_R];
    uint32_t pre[NS_R];
#pragma unroll
    for (int r = 0; r < NS_R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        mv[r] = i < n && kk[r] != pk[r];
        const uint64_t b = __ballot(mv[r]);
        pre[r] = (uint32_t)__popcll(b & lanemask_lt());
        if (lane == 0) s_w[r][w] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    const uint32_t mBase = tileBase[blockIdx.x];
    const uint32_t sBase = (uint32_t)base - mBase;  // stayers of the tiles before this one
    uint32_t before = 0;                            // movers of this tile in earlier rounds
#pragma unroll
    for (int r = 0; r < NS_R; ++r) {
        const int i = base + r * SORT_THREADS + t;
        uint32_t wb = 0, rt = 0;
#pragma unroll
        for (int k = 0; k < WAVES; ++k) {
            const uint32_t c = s_w[r][k];
            wb += k < w ? c : 0u;
            rt += c;
        }
        const uint32_t rank = before + wb + pre[r];  // movers of the tile before slot i
        if (mv[r]) {
            mKeys[mBase + rank] = kk[r];
            mIdx[mBase + rank] = i;
        } else if (i < n) {
            const uint32_t sr = (uint32_t)(r * SORT_THREADS + t) - rank;
            sKeys[sBase + sr] = kk[r];
            sIdx[sBase + sr] = i;
        }
        before += rt;
    }
}

__device__ __forceinline__ uint64_t key_idx(uint32_t k, int32_t i) { return ((uint64_t)k << 32) | (uint32_t)i; }

// Merge path of the stayers (nS) and the sorted movers (nM) into the stable
// order: block b writes outputs [b T, b T + T). Its first (last) stayer is the
// first a in [max(0, d - nM), min(d, nS)] with a = nS, a = d or
// S[a] > M[d - 1 - a] (a monotone predicate), found by a 128-ary search of
// each half of the block; the tile then merges in LDS, 8 outputs per thread.
__global__ void __launch_bounds__(256) k_merge_movers(const uint32_t* __restrict__ sKeys,
                                                      const int32_t* __restrict__ sIdx,
                                                      const uint32_t* __restrict__ mKeys,
                                                      const int32_t* __restrict__ mIdx,
                                                      const uint32_t* __restrict__ counts, int n,
                                                      uint32_t* __restrict__ outKeys,
                                                      int32_t* __restrict__ outIdx) {
    // element e at PADM(e) = e + e/8: the 8-element runs of consecutive threads
    // (64 B apart) then fall in distinct banks
    __shared__ uint64_t s_v[MERGE_T + MERGE_T / 8];
    __shared__ int s_lo[2], s_hi[2];
    __shared__ uint64_t s_first[2][2];  // per half: (found lanes of the two waves)
    const int t = threadIdx.x, half = t >> 7, ht = t & 127;
    const int nM = (int)counts[0], nS = n - nM;
    const int d0 = blockIdx.x * MERGE_T, d1 = min(d0 + MERGE_T, n);
    const int d = half ? d1 : d0;
    if (ht == 0) {
        s_lo[half] = max(0, d - nM);
        s_hi[half] = min(d, nS);
    }
    __syncthreads();
    // each half narrows its range [lo, hi] (P(hi) true) by 128 probes a round
    for (;;) {
        const int lo = s_lo[half], hi = s_hi[half];
        const bool done0 = s_lo[0] >= s_hi[0], done1 = s_lo[1] >= s_hi[1];
        if (done0 && done1) break;  // uniform
        bool p = false;
        int c = hi;
        if (lo < hi) {
            // probe c = lo + (ht + 1) * (hi - lo) / 129 ... the last probe is hi - 1 at most
            c = lo + (int)(((long long)(ht + 1) * (hi - lo)) / 129);
            if (c >= hi) c = hi - 1;
            const int b = d - 1 - c;
            p = c >= nS || b < 0 || key_idx(sKeys[c], sIdx[c]) > key_idx(mKeys[b], mIdx[b]);
        }
        const uint64_t bal = __ballot(p);
        __syncthreads();  // every lane has read s_lo / s_hi
        if ((t & 63) == 0) s_first[half][(t >> 6) & 1] = bal;
        __syncthreads();
        if (ht == 0 && lo < hi) {
            // first true probe of this half (probes ascend with ht); none: the answer is hi
            const uint64_t b0 = s_first[half][0], b1 = s_first[half][1];
            const int f = b0 ? __ffsll((long long)b0) - 1 : (b1 ? 64 + __ffsll((long long)b1) - 1 : -1);
            auto probe = [&](int q) {
                int cc = lo + (int)(((long long)(q + 1) * (hi - lo)) / 129);
                return cc >= hi ? hi - 1 : cc;
            };
            if (f < 0) {
                s_lo[half] = probe(127) + 1;
            } else {
                s_hi[half] = probe(f);
                if (f > 0) s_lo[half] = probe(f - 1) + 1;
            }
        }
        __syncthreads();
    }
    const int a0 = s_lo[0], a1 = s_lo[1];
    const int b0 = d0 - a0, b1 = d1 - a1;
    const int ns = a1 - a0, nm = b1 - b0, nt = d1 - d0;
    if (nm == 0 || ns == 0) {  // (uniform) one list only: a copy
        const uint32_t* k = nm == 0 ? sKeys + a0 : mKeys + b0;
        const int32_t* x = nm == 0 ? sIdx + a0 : mIdx + b0;
        for (int e = t; e < nt; e += 256) {
            outKeys[d0 + e] = k[e];
            outIdx[d0 + e] = x[e];
        }
        return;
    }
    auto PADM = [](int e) { return e + (e >> 3); };
    for (int e = t; e < nt; e += 256) {
        if (e < ns)
            s_v[PADM(e)] = key_idx(sKeys[a0 + e], sIdx[a0 + e]);
        else
            s_v[PADM(e)] = key_idx(mKeys[b0 + e - ns], mIdx[b0 + e - ns]);
    }
    __syncthreads();
    // this thread's outputs [8t, 8t + 8): its split x (stayers among the first 8t)
    const int q = 8 * t;
    uint64_t out[8];
    int cnt = 0;
    if (q < nt) {
        int lo = max(0, q - nm), hi = min(q, ns);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const int b = q - 1 - mid;
            if (b < 0 || s_v[PADM(mid)] > s_v[PADM(ns + b)])
                hi = mid;
            else
                lo = mid + 1;
        }
        int i = lo, j = q - lo;
        cnt = min(8, nt - q);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (e < cnt) {
                const uint64_t vs = i < ns ? s_v[PADM(i)] : ~0ull, vm = j < nm ? s_v[PADM(ns + j)] : ~0ull;
                const bool takeS = j >= nm || (i < ns && vs < vm);
                out[e] = takeS ? vs : vm;
                i += takeS ? 1 : 0;
                j += takeS ? 0 : 1;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e)
        if (e < cnt) s_v[PADM(q + e)] = out[e];
    __syncthreads();
    for (int e = t; e < nt; e += 256) {
        const uint64_t v = s_v[PADM(e)];
        outKeys[d0 + e] = (uint32_t)(v >> 32);
        outIdx[d0 + e] = (int32_t)(uint32_t)v;
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
static void radix_passes(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady,
                         const uint32_t* nDev);

void radix_sort_pairs(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady) {
    if (n <= 1) return;
    if (sort_rounds_for(n) == SORT_ROUNDS_SMALL)
        radix_passes<SORT_ROUNDS_SMALL>(buf, n, bits, st, firstHistReady, nullptr);
    else
        radix_passes<SORT_ROUNDS_LARGE>(buf, n, bits, st, firstHistReady, nullptr);
}

// n: the element count, or with nDev the largest count the device count can
// take (launch sizes; the kernels read the count itself).
template <int R>
static void radix_passes(SortBuffers& buf, int n, int bits, hipStream_t st, bool firstHistReady,
                         const uint32_t* nDev) {
    if (n <= 0 || (n <= 1 && !nDev)) return;
    const int ntiles = sort_tiles_for(n);
    constexpr int TILE = SORT_THREADS * R;
    // Keys sparse in their code space (under one key per two codes) spread each
    // round over many digits, and so do the lowest digits of any keys: order
    // the tile in LDS first there, except in the last pass, whose few digit
    // values come in long runs. (G=512 cosmic web, 16.7 M
    // keys of 27 bits: 214/342/365/86 us per pass direct, 136/133/128/104
    // ordered; two-disk G=128, 4.2 M keys of 21 bits: 43/24/23 direct,
    // 38/34/33 ordered.)
    const bool sparse = SD_SORT_LOCAL && 2.0 * (double)n < std::ldexp(1.0, bits);
    // with a device count: a grid of at most SORT_DEV_GRID blocks walks the tiles
    const int grid = nDev ? std::min(ntiles, SORT_DEV_GRID) : ntiles;
    for (int shift = 0; shift < bits; shift += 8) {
        if (shift > 0 || !firstHistReady)
            hipLaunchKernelGGL(k_sort_hist<R>, dim3(grid), dim3(SORT_THREADS), 0, st, buf.keys, n, shift,
                               buf.hist, ntiles, nDev);
        hipLaunchKernelGGL(k_scan_digit_rows, dim3(256), dim3(256), 0, st, buf.hist, ntiles, buf.chunk, nDev, TILE);
        if (SD_SORT_LOCAL && (sparse || shift == 0) && shift + 8 < bits)
            hipLaunchKernelGGL((k_sort_scatter<true, R>), dim3(grid), dim3(SORT_THREADS), 0, st, buf.keys, buf.idx, n,
                               shift, buf.hist, buf.chunk, ntiles, buf.keysAlt, buf.idxAlt, nDev);
        else
            hipLaunchKernelGGL((k_sort_scatter<false, R>), dim3(grid), dim3(SORT_THREADS), 0, st, buf.keys, buf.idx, n,
                               shift, buf.hist, buf.chunk, ntiles, buf.keysAlt, buf.idxAlt, nDev);
        uint32_t* tk = buf.keys;
        buf.keys = buf.keysAlt;
        buf.keysAlt = tk;
        int32_t* ti = buf.idx;
        buf.idx = buf.idxAlt;
        buf.idxAlt = ti;
    }
}

int mover_tiles_for(int n) { return (n + NS_TILE - 1) / NS_TILE; }

void sort_nearly_sorted(SortBuffers& buf, const MoverBuffers& mb, const float4* pos, int n, int G, int mode,
                        int bits, hipStream_t st) {
    if (n <= 0) return;
    const int ntiles = mover_tiles_for(n);
    // buf.keys: the previous (ascending) keys; buf.keysAlt: the new keys
    hipLaunchKernelGGL(k_keys_movers, dim3(ntiles), dim3(SORT_THREADS), 0, st, pos, n, G, mode, buf.keys, buf.keysAlt,
                       mb.tileMovers);
    hipLaunchKernelGGL(k_scan_tile_movers, dim3(1), dim3(1024), 0, st, mb.tileMovers, ntiles, n, mb.counts);
    hipLaunchKernelGGL(k_split_movers, dim3(ntiles), dim3(SORT_THREADS), 0, st, buf.keysAlt, buf.keys, n,
                       mb.tileMovers, mb.sKeys, mb.sIdx, mb.mKeys, mb.mIdx);
    // the movers' stable sort, sized for n, on the device count
    SortBuffers ms{mb.mKeys, mb.mIdx, mb.mKeysAlt, mb.mIdxAlt, buf.hist, buf.chunk, buf.maxTiles};
    if (sort_rounds_for(n) == SORT_ROUNDS_SMALL)
        radix_passes<SORT_ROUNDS_SMALL>(ms, n, bits, st, false, mb.counts);
    else
        radix_passes<SORT_ROUNDS_LARGE>(ms, n, bits, st, false, mb.counts);
    hipLaunchKernelGGL(k_merge_movers, dim3((n + MERGE_T - 1) / MERGE_T), dim3(256), 0, st, mb.sKeys, mb.sIdx, ms.keys,
                       ms.idx, mb.counts, n, buf.keys, buf.idx);
}

void launch_reorder(const float4* srcP, const float4* srcV, const int32_t* order, float4* dstP,
                    float4* dstV, int n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_reorder, dim3((n + 255) / 256), dim3(256), 0, st, srcP, srcV, order, dstP, dstV, n);
}

void launch_cell_range(const uint32_t* keys, int n, int2* range, int2* blockRange, long long blockEntries,
                       hipStream_t st) {
    // blockRange holds 2 * blockEntries int2: the runs, then the cells' occupancy bits
    unsigned long long* occ = blockRange ? reinterpret_cast<unsigned long long*>(blockRange + blockEntries) : nullptr;
    if (blockRange) launch_fill_i32(reinterpret_cast<int32_t*>(blockRange), 4 * blockEntries, 0, st);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_range, dim3((n + 256 * CR_R - 1) / (256 * CR_R)), dim3(256), 0, st, keys, n, range,
                       blockRange, occ);
}

}  // namespace sd
