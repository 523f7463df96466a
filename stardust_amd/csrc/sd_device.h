// sd_device.h -- device-side helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SD_FIX_SHIFT 24
#define SD_FIX_ONE (1LL << SD_FIX_SHIFT)

namespace sd {

// Cell-index clamp/wrap of the reference's gidx3 (Sim.cu:96-100) on a G-sided grid.
__device__ __forceinline__ int wrap_or_clamp(int c, int G, int periodic) {
    if (periodic) {
        // G is a power of two whenever the boundary is periodic; the modulo form
        // also serves 3*2^k meshes.
        c %= G;
        return c < 0 ? c + G : c;
    }
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
    int pad;
};

struct ReducePartial {  // one per integrate block
    double sx, sy, sz, cnt;
    float vmax;
    float pad[3];
};

}  // namespace sd
