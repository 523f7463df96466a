// sd_device.h -- device-side helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SD_FIX_SHIFT 24
#define SD_FIX_ONE (1LL << SD_FIX_SHIFT)

namespace sd {

// Cell-index clamp/wrap of the reference's gidx3 (Sim.cu:96-100) on a G-sided grid.
// XCD-aware block order. Workgroups of a launch are dealt to the 8 XCDs in
// turn (block b to XCD b % 8), and each XCD has its own L2. Renumbering so
// that XCD x runs one contiguous range of logical blocks keeps neighbouring
// tiles (and the data they share, e.g. a potential stencil halo) in one L2.
// A bijection on [0, nb): the first nb % 8 XCDs (all, when 8 divides nb) get
// ceil(nb / 8) blocks, the others one fewer.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int per = (nb + NUM_XCD - 1) / NUM_XCD;
    const int tall = nb % NUM_XCD == 0 ? NUM_XCD : nb % NUM_XCD;
    const int x = b % NUM_XCD, local = b / NUM_XCD;
    return x < tall ? x * per + local : tall * per + (x - tall) * (per - 1) + local;
}

__device__ __forceinline__ int wrap_or_clamp(int c, int G, int periodic) {
    if (periodic) {
        // G is a power of two whenever the boundary is periodic; the modulo form
        // also serves 3*2^k meshes.
        c %= G;
        return c < 0 ? c + G : c;
    }
    return min(max(c, 0), G - 1);
}

// The same for c in [-G, 2G) (for instance a cell index inside the mesh plus
// or minus a few cells): no division.
__device__ __forceinline__ int wrap_or_clamp1(int c, int G, int periodic) {
    if (periodic) return c < 0 ? c + G : (c >= G ? c - G : c);
    return min(max(c, 0), G - 1);
}

__device__ __forceinline__ uint32_t morton_spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return morton_spread10(x) | (morton_spread10(y) << 1) | (morton_spread10(z) << 2);
}

// Gradient of the potential at cell i: kGridAccel, Sim.cu:589-611.
// a = -(pot[+1] - pot[-1]) * (0.5 * G * potScale); isolated clamps the stencil,
// periodic wraps. pot and acc are G^3 [z][y][x].
__device__ __forceinline__ void grid_accel_body(const float* __restrict__ pot, float4* __restrict__ acc, int G,
                                                float h, int periodic, long long i) {
    const long long cells = (long long)G * G * G;
    if (i >= cells) return;
    const int x = (int)(i % G), y = (int)((i / G) % G), z = (int)(i / ((long long)G * G));
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
    const size_t row = ((size_t)z * G + y) * G;
    const float ax = -(pot[row + xp] - pot[row + xm]) * h;
    const float ay = -(pot[((size_t)z * G + yp) * G + x] - pot[((size_t)z * G + ym) * G + x]) * h;
    const float az = -(pot[((size_t)zp * G + y) * G + x] - pot[((size_t)zm * G + y) * G + x]) * h;
    acc[i] = make_float4(ax, ay, az, 0.f);
}

// Job of a block in a batched launch: the j with prefix[j] <= b < prefix[j + 1]
// (prefix[0] = 0, nondecreasing; empty jobs are skipped). The search runs on
// uniform values, so the loads are scalar.
__device__ __forceinline__ int find_job(const int* __restrict__ prefix, int njobs, int b) {
    int lo = 0, hi = njobs;  // prefix[lo] <= b < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Wavefront-wide reductions (64 lanes on CDNA).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Per-step scalars that live on the device so a step never waits on the host.
struct StepScalars {
    float dt;          // dt of the step being executed
    float kick;        // kick weight applied by the integrate kernel
    float dtPrev;      // previous step's dt (leapfrog kick = (dtPrev + dt)/2)
    float vmax;        // max |v| measured for the CFL rule
    float meanSub[4];  // pending momentum correction (x, y, z, unused)
    double simTime;
    double mom[4];     // last momentum sums (x, y, z, count)
    long long stepCount;
    int kdkOpen;       // 1 until the first leapfrog step has been taken
    int farCount;      // far movers listed by the gather deposit (0 between steps)
};

struct ReducePartial {  // one per integrate block
    double sx, sy, sz, cnt;
    float vmax;
    float pad[3];
};

struct DtPolicy {
    float fixedDt;      // > 0: fixed step
    float baseDt;       // 0.0016 * dtScaleFor(timeScale)
    float cell;         // 1/G
    float dtStiff;      // contact cap (0 = none)
    int leapfrog;       // kick = (dtPrev + dt)/2
};

// One reduction of integrate partials into the step scalars.
struct FinalizeJob {
    const ReducePartial* partials;  // nullptr: none
    int nparts;
    int mode;  // bit0 momentum, bit1 vmax, bit2 next dt/kick, bit3 advance time
    DtPolicy pol;
    StepScalars* sc;
};

// Partials reduction and dt policy (Sim.cu:4254-4296) in one fixed summation
// order: FIN virtual threads, virtual thread v adding partials v + i FIN in
// ascending i, a butterfly inside each virtual wavefront of 64, then the
// FIN / 64 wavefronts in index order. A block of TH threads (TH divides FIN)
// plays the virtual threads j TH + t for j = 0 .. FIN / TH - 1 in turn, so
// every block size gives the same bits.
constexpr int FIN = 1024;

// Virtual thread vb + threadIdx.x's sums, then its wavefront's (every lane).
__device__ __forceinline__ void finalize_waves(const ReducePartial* __restrict__ parts, int np, int vb, double d[4],
                                               float& vm) {
    constexpr int UNR = 4;  // loads in flight
    d[0] = d[1] = d[2] = d[3] = 0.0;
    vm = 0.f;
#pragma unroll 1
    for (int base = vb + (int)threadIdx.x; base < np; base += UNR * FIN) {
        ReducePartial r[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int k = base + u * FIN;
            if (k < np) r[u] = parts[k];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (base + u * FIN < np) {
                d[0] += r[u].sx;
                d[1] += r[u].sy;
                d[2] += r[u].sz;
                d[3] += r[u].cnt;
                vm = fmaxf(vm, r[u].vmax);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = wave_sum(d[q]);
    vm = wave_max(vm);
}

// The step scalars from the totals (sums t, max speed m).
__device__ __forceinline__ void finalize_apply(const double t[4], float m, int mode, const DtPolicy& pol,
                                               StepScalars* sc) {
    if (mode & 1) {
        for (int q = 0; q < 4; ++q) sc->mom[q] = t[q];
        const double c = t[3];
        if (c >= 1.0) {
            sc->meanSub[0] = (float)(t[0] / c);
            sc->meanSub[1] = (float)(t[1] / c);
            sc->meanSub[2] = (float)(t[2] / c);
        } else {
            sc->meanSub[0] = sc->meanSub[1] = sc->meanSub[2] = 0.f;
        }
    }
    if (mode & 2) sc->vmax = m;
    if (mode & 8) {  // advance time by the step just taken
        sc->simTime += (double)sc->dt;
        sc->dtPrev = sc->dt;
        sc->kdkOpen = 0;
        sc->stepCount += 1;
    }
    if (mode & 4) {  // dt (and kick) of the next integrate
        float dt;
        if (pol.fixedDt > 0.f) {
            dt = pol.fixedDt;
        } else {
            const float vmax = sc->vmax;
            dt = pol.baseDt;
            const float dtMax = (vmax > 1e-6f) ? (0.8f * pol.cell / vmax) : dt;
            if (dt > dtMax) dt = dtMax;
            if (pol.dtStiff > 0.f && dt > pol.dtStiff) dt = pol.dtStiff;
        }
        sc->dt = dt;
        sc->kick = pol.leapfrog ? 0.5f * ((sc->kdkOpen ? 0.f : sc->dtPrev) + dt) : dt;
    }
}

// The whole reduction in one block of TH threads.
template <int TH>
__device__ __forceinline__ void finalize_body(const ReducePartial* __restrict__ parts, int np, int mode,
                                              const DtPolicy& pol, StepScalars* sc) {
    static_assert(FIN % TH == 0 && TH % 64 == 0, "block size must divide FIN");
    __shared__ double s_w[FIN / 64][4];
    __shared__ float s_v[FIN / 64];
#pragma unroll 1
    for (int j = 0; j < FIN / TH; ++j) {
        double d[4];
        float vm;
        finalize_waves(parts, np, j * TH, d, vm);
        if ((threadIdx.x & 63) == 0) {
            const int w = (j * TH + (int)threadIdx.x) >> 6;
#pragma unroll
            for (int q = 0; q < 4; ++q) s_w[w][q] = d[q];
            s_v[w] = vm;
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    float m = 0.f;
    for (int w = 0; w < FIN / 64; ++w) {
        for (int q = 0; q < 4; ++q) t[q] += s_w[w][q];
        m = fmaxf(m, s_v[w]);
    }
    finalize_apply(t, m, mode, pol, sc);
}

}  // namespace sd
