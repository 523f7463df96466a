/*
 * sd_oracle.c -- CPU oracle for the Stardust particle-mesh gravity step.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library as the checker; the product never does.
 *
 * Each function restates one piece of the reference pipeline
 * (src/sim/Sim.cu, chdnl0420-svg/stardust) and names the lines it follows.
 * Arithmetic conventions:
 *   - Per-particle FP32 code follows the reference's operation order. The
 *     reference is built by nvcc with its default contraction (-fmad=true), so
 *     every "a*b + c" there is a fused multiply-add; this file writes those as
 *     explicit fmaf() so the restatement does not depend on the host compiler
 *     (the file is compiled with -ffp-contract=off).
 *   - The mesh (FFT Poisson solve) is done in FP64 with this file's own FFT;
 *     the reference uses FP32 cuFFT, whose rounding is not reproducible here,
 *     so mesh results are compared by tolerance.
 *
 * Parity pins (tests/test_oracle_*.py): every function listed here is checked
 * against closed forms, invariants, numpy, or brute force on small inputs.
 */
#define _GNU_SOURCE
#include "sd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* gidx3, Sim.cu:96-100: periodic wraps with &(G-1), isolated clamps to the edge;
 * the index is into a grid of side S. */
static inline int64_t gidx3(int x, int y, int z, int G, int S, int periodic) {
    if (periodic) {
        x &= (G - 1);
        y &= (G - 1);
        z &= (G - 1);
    } else {
        x = clampi(x, 0, G - 1);
        y = clampi(y, 0, G - 1);
        z = clampi(z, 0, G - 1);
    }
    return ((int64_t)z * S + y) * S + x;
}

/* ------------------------------------------------------------------------ */
/* Keys, sort, reorder                                                       */
/* ------------------------------------------------------------------------ */

/* Morton interleave (x in bit 0, y in bit 1, z in bit 2), 10 bits per axis. */
uint32_t sdo_morton3(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) {
        r |= ((x >> b) & 1u) << (3 * b);
        r |= ((y >> b) & 1u) << (3 * b + 1);
        r |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return r;
}

/* kCellKey, Sim.cu:2476-2486: c = clamp((int)(p*G), 0, G-1), key = (cz*G+cy)*G+cx,
 * inactive (p.x < 0) -> -1, idx[i] = i. Stored here as key+1 (unsigned). Morton
 * mode interleaves the same cell coordinates; Morton-node mode interleaves the
 * CIC base node clamp(floor(p*G - 0.5), 0, G-1) instead (no reference
 * counterpart: the product's step order). */
void sdo_cell_keys(const float* pos4, int n, int G, int mode, uint32_t* keys, int32_t* idx) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        idx[i] = i;
        if (p[0] < 0.f) {
            keys[i] = 0u;
            continue;
        }
        const float fG = (float)G;
        int cx, cy, cz;
        if (mode == SDO_KEY_MORTON_NODE) {
            cx = clampi((int)floorf(fmaf(p[0], fG, -0.5f)), 0, G - 1);
            cy = clampi((int)floorf(fmaf(p[1], fG, -0.5f)), 0, G - 1);
            cz = clampi((int)floorf(fmaf(p[2], fG, -0.5f)), 0, G - 1);
        } else {
            cx = clampi((int)(p[0] * fG), 0, G - 1);
            cy = clampi((int)(p[1] * fG), 0, G - 1);
            cz = clampi((int)(p[2] * fG), 0, G - 1);
        }
        uint32_t k;
        if (mode == SDO_KEY_MORTON || mode == SDO_KEY_MORTON_NODE)
            k = sdo_morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
        else
            k = (uint32_t)((cz * G + cy) * G + cx);
        keys[i] = k + 1u;
    }
}

/* Stable ascending sort of (key, value) pairs: the documented semantics of
 * cub::DeviceRadixSort::SortPairs used at Sim.cu:3660-3667. LSD radix, 8 bits. */
void sdo_stable_sort_pairs(uint32_t* keys, int32_t* vals, int n) {
    if (n <= 1) return;
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    int32_t* v2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    uint32_t kmax = 0;
    for (int i = 0; i < n; ++i)
        if (keys[i] > kmax) kmax = keys[i];
    uint32_t* ks = keys;
    int32_t* vs = vals;
    uint32_t* kd = k2;
    int32_t* vd = v2;
    for (int shift = 0; shift < 32 && (shift == 0 || (kmax >> shift) != 0); shift += 8) {
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int i = 0; i < n; ++i) {
            const size_t o = cnt[(ks[i] >> shift) & 255u]++;
            kd[o] = ks[i];
            vd[o] = vs[i];
        }
        uint32_t* tk = ks; ks = kd; kd = tk;
        int32_t* tv = vs; vs = vd; vd = tv;
    }
    if (ks != keys) {
        memcpy(keys, ks, sizeof(uint32_t) * (size_t)n);
        memcpy(vals, vs, sizeof(int32_t) * (size_t)n);
    }
    free(k2);
    free(v2);
}

/* kReorder, Sim.cu:2488-2495: dst[i] = src[order[i]] for float4 arrays. */
void sdo_reorder4(const float* src4, const int32_t* order, int n, float* dst4) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) memcpy(dst4 + 4 * (size_t)i, src4 + 4 * (size_t)order[i], 16);
}

/* ------------------------------------------------------------------------ */
/* Deposit                                                                   */
/* ------------------------------------------------------------------------ */

/* kScatter, Sim.cu:503-521 (cleared by kClearF, 487-490): cell-centred CIC with
 * g = p*G - 0.5 (a fused multiply-add under nvcc), mass 1 per particle, the
 * eight weights are products of (1-f) / f, isolated clamps out-of-range nodes
 * onto the edge (mass kept), periodic wraps. Adds in particle order -- one of
 * the orders the reference's float atomics may take. rho has side S. */
void sdo_deposit_float(const float* pos4, int n, int G, int S, int periodic, float* rho) {
    const size_t cells = (size_t)S * S * S;
    memset(rho, 0, sizeof(float) * cells);
    const float fG = (float)G;
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        if (p[0] < 0.f) continue;
        const float gx = fmaf(p[0], fG, -0.5f), gy = fmaf(p[1], fG, -0.5f),
                    gz = fmaf(p[2], fG, -0.5f);
        const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
        const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
        for (int k = 0; k < 8; ++k) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            const float w = (ox ? fx : 1.f - fx) * (oy ? fy : 1.f - fy) * (oz ? fz : 1.f - fz);
            rho[gidx3(ix + ox, iy + oy, iz + oz, G, S, periodic)] += w;
        }
    }
}

/* Fixed-point CIC weights (no reference counterpart: the deterministic deposit).
 * The particle's 2^24 mass units are split exactly along x, then y, then z:
 *   w1 = rint(f * w), w0 = w - w1   (f in FP32, rint half-to-even),
 * so the eight integer weights always sum to 2^24. The fractional offsets
 * f and base nodes are the reference CIC's (Sim.cu:507-509). base3 receives
 * (ix, iy, iz); w8 is ordered k = ox + 2*oy + 4*oz. */
void sdo_cic_weights_fixed(float px, float py, float pz, int G, int* base3, int64_t* w8) {
    const float fG = (float)G;
    const float gx = fmaf(px, fG, -0.5f), gy = fmaf(py, fG, -0.5f), gz = fmaf(pz, fG, -0.5f);
    const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    base3[0] = ix;
    base3[1] = iy;
    base3[2] = iz;
    int64_t wx[2], wxy[4];
    wx[1] = (int64_t)rintf(fx * (float)SD_FIX_ONE);
    wx[0] = SD_FIX_ONE - wx[1];
    for (int ox = 0; ox < 2; ++ox) {
        const int64_t y1 = (int64_t)rintf(fy * (float)wx[ox]);
        wxy[ox + 2] = y1;
        wxy[ox] = wx[ox] - y1;
    }
    for (int j = 0; j < 4; ++j) {
        const int64_t z1 = (int64_t)rintf(fz * (float)wxy[j]);
        w8[j + 4] = z1;
        w8[j] = wxy[j] - z1;
    }
}

/* Deterministic deposit: the reference CIC stencil and edge rules (Sim.cu:503-521)
 * with the fixed-point weights above, summed in int64 (order-independent, exact).
 * rho has side G (the G-corner of the reference's padded grid). */
void sdo_deposit_fixed(const float* pos4, int n, int G, int periodic, int64_t* rho) {
    const size_t cells = (size_t)G * G * G;
    memset(rho, 0, sizeof(int64_t) * cells);
    /* integer adds are exact, so the atomic adds of several threads give the
     * same grid in any order */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        if (p[0] < 0.f) continue;
        int b[3];
        int64_t w[8];
        sdo_cic_weights_fixed(p[0], p[1], p[2], G, b, w);
        for (int k = 0; k < 8; ++k) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            __atomic_fetch_add(&rho[gidx3(b[0] + ox, b[1] + oy, b[2] + oz, G, G, periodic)], w[k],
                               __ATOMIC_RELAXED);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FFT (own mixed-radix, FP64)                                               */
/* ------------------------------------------------------------------------ */

/* Mixed-radix FFT in FP64 (sdo_fft1d / sdo_fft3d) and FP32 (sdo_fft3d_f32,
 * the CPU baseline's single-precision mesh): one source, compiled twice. */
#define REAL double
#define SUF(name) name
#include "sd_oracle_fft.inc"
#undef REAL
#undef SUF
#define REAL float
#define SUF(name) name##_f32
#include "sd_oracle_fft.inc"
#undef REAL
#undef SUF

/* kGreen, Sim.cu:563-573: g = -1/(sqrt(dx^2+dy^2+dz^2) + soft) with wrapped
 * distances d = x <= S/2 ? x : S - x. Evaluated in FP64. */
void sdo_green_real(int S, float soft, double* g) {
    for (int z = 0; z < S; ++z)
        for (int y = 0; y < S; ++y)
            for (int x = 0; x < S; ++x) {
                const int dx = x <= S / 2 ? x : S - x;
                const int dy = y <= S / 2 ? y : S - y;
                const int dz = z <= S / 2 ? z : S - z;
                const double r = sqrt((double)(dx * dx + dy * dy + dz * dz)) + (double)soft;
                g[((size_t)z * S + y) * S + x] = -1.0 / r;
            }
}

/* Isolated solve, Sim.cu:3569-3603 (buildGreen + solvePoisson + kMulSpec):
 * (the Green spectrum is cached per (S, soft) like the reference's specGreen)
 * zero-pad the G^3 density into S = 2G, circular convolution with the Green
 * table (soft = max(softeningCells, 0.5), Sim.cu:3572), normalised by 1/S^3,
 * keep the G-corner. rho and pot are G^3 [z][y][x]. */
static double* g_greenRe = NULL; /* cached spectrum of the Green table */
static double* g_greenIm = NULL;
static int g_greenS = 0;
static float g_greenSoft = -1.f;

void sdo_poisson_iso(const double* rho, int G, float softCells, double* pot) {
    const int S = 2 * G;
    const size_t N = (size_t)S * S * S;
    const float soft = softCells > 0.5f ? softCells : 0.5f;
    if (g_greenS != S || g_greenSoft != soft) {
        free(g_greenRe);
        free(g_greenIm);
        g_greenRe = (double*)malloc(N * sizeof(double));
        g_greenIm = (double*)calloc(N, sizeof(double));
        sdo_green_real(S, soft, g_greenRe);
        sdo_fft3d(g_greenRe, g_greenIm, S, S, S, 0);
        g_greenS = S;
        g_greenSoft = soft;
    }
    const double* gr = g_greenRe;
    const double* gi = g_greenIm;
    double* ar = (double*)calloc(N, sizeof(double));
    double* ai = (double*)calloc(N, sizeof(double));
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x)
                ar[((size_t)z * S + y) * S + x] = rho[((size_t)z * G + y) * G + x];
    sdo_fft3d(ar, ai, S, S, S, 0);
    const double norm = 1.0 / (double)N;
    for (size_t i = 0; i < N; ++i) {
        const double xr = ar[i], xi = ai[i];
        ar[i] = (xr * gr[i] - xi * gi[i]) * norm;
        ai[i] = (xr * gi[i] + xi * gr[i]) * norm;
    }
    sdo_fft3d(ar, ai, S, S, S, 1);
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x)
                pot[((size_t)z * G + y) * G + x] = ar[((size_t)z * S + y) * S + x];
    free(ar);
    free(ai);
}

/* Periodic solve, Sim.cu:528-557 with the scale of Sim.cu:3593-3597:
 * phi_k = rho_k * (-(1/G^3)/(4 pi^2)) / k^2 * exp(-k^2 (2 pi soft/G)^2 / 2),
 * integer wavenumbers folded to [-G/2, G/2], k = 0 -> 0, no softening floor. */
void sdo_poisson_periodic(const double* rho, int G, float softCells, double* pot) {
    const size_t N = (size_t)G * G * G;
    double* ar = (double*)malloc(N * sizeof(double));
    double* ai = (double*)calloc(N, sizeof(double));
    memcpy(ar, rho, N * sizeof(double));
    sdo_fft3d(ar, ai, G, G, G, 0);
    const double scale = (1.0 / (double)N) / (4.0 * M_PI * M_PI);
    const double s = 2.0 * M_PI * (double)softCells / (double)G;
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                const int kx = x <= G / 2 ? x : x - G;
                const int ky = y <= G / 2 ? y : y - G;
                const int kz = z <= G / 2 ? z : z - G;
                const double k2 = (double)(kx * kx + ky * ky + kz * kz);
                const size_t i = ((size_t)z * G + y) * G + x;
                double m = 0.0;
                if (k2 > 0.0) {
                    m = -scale / k2;
                    if (softCells > 0.f) m *= exp(-0.5 * k2 * s * s);
                }
                ar[i] *= m;
                ai[i] *= m;
            }
    sdo_fft3d(ar, ai, G, G, G, 1);
    memcpy(pot, ar, N * sizeof(double));
    free(ar);
    free(ai);
}

/* Single-precision mesh of the CPU baseline (no reference counterpart beyond
 * cuFFT's FP32 transforms at Sim.cu:3592, 3601): the same isolated and
 * periodic solves as above with FP32 arrays and transforms (twiddles rounded
 * from FP64), the Green spectrum cached per (S, soft). Used only for timing. */
static float* g_greenRe32 = NULL;
static float* g_greenIm32 = NULL;
static int g_greenS32 = 0;
static float g_greenSoft32 = -1.f;

void sdo_poisson_iso_f32(const float* rho, int G, float softCells, float* pot) {
    const int S = 2 * G;
    const size_t N = (size_t)S * S * S;
    const float soft = softCells > 0.5f ? softCells : 0.5f;
    if (g_greenS32 != S || g_greenSoft32 != soft) {
        free(g_greenRe32);
        free(g_greenIm32);
        double* g = (double*)malloc(N * sizeof(double));
        sdo_green_real(S, soft, g);
        g_greenRe32 = (float*)malloc(N * sizeof(float));
        g_greenIm32 = (float*)calloc(N, sizeof(float));
        for (size_t i = 0; i < N; ++i) g_greenRe32[i] = (float)g[i];
        free(g);
        sdo_fft3d_f32(g_greenRe32, g_greenIm32, S, S, S, 0);
        g_greenS32 = S;
        g_greenSoft32 = soft;
    }
    float* ar = (float*)calloc(N, sizeof(float));
    float* ai = (float*)calloc(N, sizeof(float));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            memcpy(ar + ((size_t)z * S + y) * S, rho + ((size_t)z * G + y) * G, sizeof(float) * (size_t)G);
    sdo_fft3d_f32(ar, ai, S, S, S, 0);
    const float norm = (float)(1.0 / (double)N);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)N; ++i) {
        const float xr = ar[i], xi = ai[i], gr = g_greenRe32[i], gi = g_greenIm32[i];
        ar[i] = (xr * gr - xi * gi) * norm;
        ai[i] = (xr * gi + xi * gr) * norm;
    }
    sdo_fft3d_f32(ar, ai, S, S, S, 1);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            memcpy(pot + ((size_t)z * G + y) * G, ar + ((size_t)z * S + y) * S, sizeof(float) * (size_t)G);
    free(ar);
    free(ai);
}

void sdo_poisson_periodic_f32(const float* rho, int G, float softCells, float* pot) {
    const size_t N = (size_t)G * G * G;
    float* ar = (float*)malloc(N * sizeof(float));
    float* ai = (float*)calloc(N, sizeof(float));
    memcpy(ar, rho, N * sizeof(float));
    sdo_fft3d_f32(ar, ai, G, G, G, 0);
    const float scale = (float)((1.0 / (double)N) / (4.0 * M_PI * M_PI));
    const float sg = (float)(2.0 * M_PI * (double)softCells / (double)G);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                const int kx = x <= G / 2 ? x : x - G;
                const int ky = y <= G / 2 ? y : y - G;
                const int kz = z <= G / 2 ? z : z - G;
                const float k2 = (float)(kx * kx + ky * ky + kz * kz);
                const size_t i = ((size_t)z * G + y) * G + x;
                float m = 0.f;
                if (k2 > 0.f) {
                    m = -scale / k2;
                    if (softCells > 0.f) m *= expf(-0.5f * k2 * sg * sg);
                }
                ar[i] *= m;
                ai[i] *= m;
            }
    sdo_fft3d_f32(ar, ai, G, G, G, 1);
    memcpy(pot, ar, N * sizeof(float));
    free(ar);
    free(ai);
}

/* kGridAccel, Sim.cu:589-611: a = -(pot[+1] - pot[-1]) * (0.5*G*potScale);
 * isolated clamps the stencil at the edge, periodic wraps. pot is G^3. */
void sdo_grid_accel(const double* pot, int G, float potScale, int periodic, float* acc4) {
    const double h = (double)(0.5f * (float)G * potScale);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                int xm, xp, ym, yp, zm, zp;
                if (periodic) {
                    xm = (x - 1) & (G - 1); xp = (x + 1) & (G - 1);
                    ym = (y - 1) & (G - 1); yp = (y + 1) & (G - 1);
                    zm = (z - 1) & (G - 1); zp = (z + 1) & (G - 1);
                } else {
                    xm = x > 0 ? x - 1 : 0; xp = x < G - 1 ? x + 1 : G - 1;
                    ym = y > 0 ? y - 1 : 0; yp = y < G - 1 ? y + 1 : G - 1;
                    zm = z > 0 ? z - 1 : 0; zp = z < G - 1 ? z + 1 : G - 1;
                }
#define P(X, Y, Z) pot[((size_t)(Z) * G + (Y)) * G + (X)]
                float* a = acc4 + 4 * (((size_t)z * G + y) * G + x);
                a[0] = (float)(-(P(xp, y, z) - P(xm, y, z)) * h);
                a[1] = (float)(-(P(x, yp, z) - P(x, ym, z)) * h);
                a[2] = (float)(-(P(x, y, zp) - P(x, y, zm)) * h);
                a[3] = 0.f;
#undef P
            }
}

/* The same gradient from the single-precision potential (the CPU baseline). */
void sdo_grid_accel_f32(const float* pot, int G, float potScale, int periodic, float* acc4) {
    const float h = 0.5f * (float)G * potScale;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                int xm, xp, ym, yp, zm, zp;
                if (periodic) {
                    xm = (x - 1) & (G - 1); xp = (x + 1) & (G - 1);
                    ym = (y - 1) & (G - 1); yp = (y + 1) & (G - 1);
                    zm = (z - 1) & (G - 1); zp = (z + 1) & (G - 1);
                } else {
                    xm = x > 0 ? x - 1 : 0; xp = x < G - 1 ? x + 1 : G - 1;
                    ym = y > 0 ? y - 1 : 0; yp = y < G - 1 ? y + 1 : G - 1;
                    zm = z > 0 ? z - 1 : 0; zp = z < G - 1 ? z + 1 : G - 1;
                }
#define P(X, Y, Z) pot[((size_t)(Z) * G + (Y)) * G + (X)]
                float* a = acc4 + 4 * (((size_t)z * G + y) * G + x);
                a[0] = -(P(xp, y, z) - P(xm, y, z)) * h;
                a[1] = -(P(x, yp, z) - P(x, ym, z)) * h;
                a[2] = -(P(x, y, zp) - P(x, y, zm)) * h;
                a[3] = 0.f;
#undef P
            }
}

/* ------------------------------------------------------------------------ */
/* Particles                                                                 */
/* ------------------------------------------------------------------------ */

/* sampleAcc, Sim.cu:614-629: CIC gather with the deposit's weights and edge
 * rules; "a += g*w" is fused as nvcc does. */
void sdo_sample_acc(const float* acc4, float px, float py, float pz, int G, int periodic,
                    float* out3) {
    const float fG = (float)G;
    const float gx = fmaf(px, fG, -0.5f), gy = fmaf(py, fG, -0.5f), gz = fmaf(pz, fG, -0.5f);
    const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = (ox ? fx : 1.f - fx) * (oy ? fy : 1.f - fy) * (oz ? fz : 1.f - fz);
        const float* g = acc4 + 4 * gidx3(ix + ox, iy + oy, iz + oz, G, G, periodic);
        ax = fmaf(g[0], w, ax);
        ay = fmaf(g[1], w, ay);
        az = fmaf(g[2], w, az);
    }
    out3[0] = ax;
    out3[1] = ay;
    out3[2] = az;
}

/* sampleAcc at every live particle (zero for inactive slots): out4[n*4]. */
void sdo_sample_acc_all(const float* acc4, const float* pos4, int n, int G, int periodic, float* out4) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        float* o = out4 + 4 * (size_t)i;
        o[0] = o[1] = o[2] = o[3] = 0.f;
        if (p[0] < 0.f) continue;
        sdo_sample_acc(acc4, p[0], p[1], p[2], G, periodic, o);
    }
}

/* kMaxSpeed, Sim.cu:2923-2930: max |v| over live particles. */
float sdo_max_speed(const float* pos4, const float* vel4, int n) {
    float m = 0.f;
    for (int i = 0; i < n; ++i) {
        if (pos4[4 * (size_t)i] < 0.f) continue;
        const float* v = vel4 + 4 * (size_t)i;
        const float s = sqrtf(fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0])));
        if (s > m) m = s;
    }
    return m;
}

/* kIntegrate, Sim.cu:685-906, plain-PM subset (no black holes, no pressure):
 *   a = CIC(accG) [+ accContact] [+ dark energy];
 *   v += a*kick; |v| clamped to sqrt(c2); p += v*drift;
 *   non-finite -> slot killed (p = -1, v = 0);
 *   periodic: p -= floor(p); isolated: soft sphere if softBoundR > 0, else cube
 *   walls at 0.002 / 0.998 with v = +-0.25|v|.
 * kick == drift == dt is the reference's kick-drift step. meanSub3 (optional)
 * is a pending momentum correction subtracted from v on load. */
void sdo_integrate(const float* acc4, float* pos4, float* vel4, int n, int G, float kick,
                   float drift, int periodic, float c2, const float* accContact4,
                   float softBoundR, float darkEnergy, const float* meanSub3, const float* accPress4) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        float* P = pos4 + 4 * (size_t)i;
        float* V = vel4 + 4 * (size_t)i;
        float p[4] = {P[0], P[1], P[2], P[3]};
        if (p[0] < 0.f) continue;
        float v[4] = {V[0], V[1], V[2], V[3]};
        if (meanSub3) {
            v[0] -= meanSub3[0];
            v[1] -= meanSub3[1];
            v[2] -= meanSub3[2];
        }
        float a[3];
        sdo_sample_acc(acc4, p[0], p[1], p[2], G, periodic, a);
        if (accContact4) {
            const float* c = accContact4 + 4 * (size_t)i;
            a[0] += c[0];
            a[1] += c[1];
            a[2] += c[2];
        }
        if (darkEnergy != 0.f) {
            float k = 1.0f;
            if (softBoundR > 0.f) {
                const float dx = p[0] - 0.5f, dy = p[1] - 0.5f, dz = p[2] - 0.5f;
                const float r = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
                if (r > softBoundR) {
                    const float R1 = 0.497f;
                    const float den = (R1 - softBoundR) > 1e-6f ? (R1 - softBoundR) : 1e-6f;
                    const float t = (r - softBoundR) / den;
                    k = 1.0f - (t < 1.0f ? t : 1.0f);
                }
            }
            if (k > 0.f) {
                a[0] = fmaf(darkEnergy * k, p[0] - 0.5f, a[0]);
                a[1] = fmaf(darkEnergy * k, p[1] - 0.5f, a[1]);
                a[2] = fmaf(darkEnergy * k, p[2] - 0.5f, a[2]);
            }
        }
        /* gas pressure, Sim.cu:740-743: gas (w <= 0), not dark matter */
        if (accPress4 && p[3] <= 0.f && v[3] >= -50.f) {
            float pr[3];
            sdo_sample_acc(accPress4, p[0], p[1], p[2], G, periodic, pr);
            a[0] += pr[0];
            a[1] += pr[1];
            a[2] += pr[2];
        }
        v[0] = fmaf(a[0], kick, v[0]);
        v[1] = fmaf(a[1], kick, v[1]);
        v[2] = fmaf(a[2], kick, v[2]);
        {
            const float sp2 = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
            if (sp2 > c2) {
                const float k = sqrtf(c2 / sp2);
                v[0] *= k;
                v[1] *= k;
                v[2] *= k;
            }
        }
        p[0] = fmaf(v[0], drift, p[0]);
        p[1] = fmaf(v[1], drift, p[1]);
        p[2] = fmaf(v[2], drift, p[2]);
        if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2]) || !isfinite(v[0]) ||
            !isfinite(v[1]) || !isfinite(v[2])) {
            P[0] = P[1] = P[2] = -1.f;
            P[3] = 0.f;
            V[0] = V[1] = V[2] = V[3] = 0.f;
            continue;
        }
        if (periodic) {
            p[0] -= floorf(p[0]);
            p[1] -= floorf(p[1]);
            p[2] -= floorf(p[2]);
        } else if (softBoundR > 0.f) {
            const float dx = p[0] - 0.5f, dy = p[1] - 0.5f, dz = p[2] - 0.5f;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float R1 = 0.497f;
            if (r2 > softBoundR * softBoundR) {
                const float r = sqrtf(r2);
                const float inv = 1.0f / (r > 1e-6f ? r : 1e-6f);
                const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
                const float den = (R1 - softBoundR) > 1e-6f ? (R1 - softBoundR) : 1e-6f;
                float t = (r - softBoundR) / den;
                t = t < 1.0f ? t : 1.0f;
                const float vr = fmaf(v[2], nz, fmaf(v[1], ny, v[0] * nx));
                if (vr > 0.f) {
                    const float cut = vr * t * t;
                    v[0] = fmaf(-nx, cut, v[0]);
                    v[1] = fmaf(-ny, cut, v[1]);
                    v[2] = fmaf(-nz, cut, v[2]);
                }
                const float pull = 2.0f * t * t;
                v[0] = fmaf(-(nx * pull), drift, v[0]);
                v[1] = fmaf(-(ny * pull), drift, v[1]);
                v[2] = fmaf(-(nz * pull), drift, v[2]);
                if (r > R1) {
                    const float k = R1 * inv;
                    p[0] = fmaf(dx, k, 0.5f);
                    p[1] = fmaf(dy, k, 0.5f);
                    p[2] = fmaf(dz, k, 0.5f);
                }
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                if (p[d] < 0.002f) { p[d] = 0.002f; v[d] = fabsf(v[d]) * 0.25f; }
                if (p[d] > 0.998f) { p[d] = 0.998f; v[d] = -fabsf(v[d]) * 0.25f; }
            }
        }
        memcpy(P, p, 16);
        memcpy(V, v, 16);
    }
}

/* kMomentumAccum, Sim.cu:2252-2268: sum of v (FP64) and count over live
 * particles with finite velocity. out4 = (sx, sy, sz, count). */
void sdo_momentum(const float* pos4, const float* vel4, int n, double* out4) {
    double s[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        if (pos4[4 * (size_t)i] < 0.f) continue;
        const float* v = vel4 + 4 * (size_t)i;
        if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) continue;
        s[0] += (double)v[0];
        s[1] += (double)v[1];
        s[2] += (double)v[2];
        s[3] += 1.0;
    }
    memcpy(out4, s, sizeof(s));
}

/* kSubtractMeanVelDev, Sim.cu:2282-2291: v -= (float)(sum/count) on live particles. */
void sdo_subtract_mean(const float* pos4, float* vel4, int n, const double* acc4) {
    const double cnt = acc4[3];
    if (cnt < 1.0) return;
    const float mx = (float)(acc4[0] / cnt), my = (float)(acc4[1] / cnt),
                mz = (float)(acc4[2] / cnt);
    for (int i = 0; i < n; ++i) {
        if (pos4[4 * (size_t)i] < 0.f) continue;
        float* v = vel4 + 4 * (size_t)i;
        v[0] -= mx;
        v[1] -= my;
        v[2] -= mz;
    }
}

/* ------------------------------------------------------------------------ */
/* Contact hash grid                                                         */
/* ------------------------------------------------------------------------ */

/* kFillInt(-1) x2 + kBuildCellRange, Sim.cu:1004-1011, 3689-3691, on keys
 * sorted ascending (key 0 = inactive, cell c = key - 1). */
void sdo_cell_range(const uint32_t* keys, int n, int ncells, int32_t* start, int32_t* end) {
    for (int c = 0; c < ncells; ++c) start[c] = end[c] = -1;
    for (int i = 0; i < n; ++i) {
        if (keys[i] == 0u) continue;
        const int c = (int)(keys[i] - 1u);
        if (i == 0 || keys[i - 1] != keys[i]) start[c] = i;
        if (i == n - 1 || keys[i + 1] != keys[i]) end[c] = i + 1;
    }
}

static inline int cell_slot(int x, int y, int z, int G, int periodic, int keyMode) {
    if (periodic) {
        x &= (G - 1); y &= (G - 1); z &= (G - 1);
    } else {
        x = clampi(x, 0, G - 1); y = clampi(y, 0, G - 1); z = clampi(z, 0, G - 1);
    }
    if (keyMode == SDO_KEY_MORTON) return (int)sdo_morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
    return (z * G + y) * G + x;
}

/* kContact, Sim.cu:1030-1203, contact term only (the strong / EM forces are
 * off in every plain-PM configuration): scan the 27 neighbour cells (same
 * clamp / wrap as gidx3), at most 96 peers, spring-dashpot
 * f = max(0, k (d0 - d) - c vn) along the pair normal, d0 = 2 radius. */
void sdo_contact(const float* pos4, const float* vel4, int n, int G, int periodic,
                 int keyMode, const int32_t* cellStart, const int32_t* cellEnd, float radius,
                 float stiffness, float damping, float* accOut4) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; ++i) {
        const float* p = pos4 + 4 * (size_t)i;
        float a[3] = {0.f, 0.f, 0.f};
        float* out = accOut4 + 4 * (size_t)i;
        if (p[0] < 0.f) {
            out[0] = out[1] = out[2] = out[3] = 0.f;
            continue;
        }
        const float* v = vel4 + 4 * (size_t)i;
        const float d0 = radius * 2.0f;
        const float d02 = d0 * d0;
        const float fG = (float)G;
        const int cx = clampi((int)(p[0] * fG), 0, G - 1);
        const int cy = clampi((int)(p[1] * fG), 0, G - 1);
        const int cz = clampi((int)(p[2] * fG), 0, G - 1);
        const int kMaxPeers = 96;
        int seen = 0;
        for (int dz = -1; dz <= 1 && seen < kMaxPeers; ++dz)
            for (int dy = -1; dy <= 1 && seen < kMaxPeers; ++dy)
                for (int dx = -1; dx <= 1 && seen < kMaxPeers; ++dx) {
                    const int c = cell_slot(cx + dx, cy + dy, cz + dz, G, periodic, keyMode);
                    const int s0 = cellStart[c], e0 = cellEnd[c];
                    if (s0 < 0 || e0 <= s0) continue;
                    for (int j = s0; j < e0 && seen < kMaxPeers; ++j) {
                        if (j == i) continue;
                        ++seen;
                        const float* q = pos4 + 4 * (size_t)j;
                        if (q[0] < 0.f) continue;
                        float ex = p[0] - q[0], ey = p[1] - q[1], ez = p[2] - q[2];
                        if (periodic) {
                            if (ex > 0.5f) ex -= 1.f; else if (ex < -0.5f) ex += 1.f;
                            if (ey > 0.5f) ey -= 1.f; else if (ey < -0.5f) ey += 1.f;
                            if (ez > 0.5f) ez -= 1.f; else if (ez < -0.5f) ez += 1.f;
                        }
                        const float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                        if (dd >= d02 || dd < 1e-16f) continue;
                        const float d = sqrtf(dd);
                        const float nx = ex / d, ny = ey / d, nz = ez / d;
                        const float* vj = vel4 + 4 * (size_t)j;
                        const float vn =
                            fmaf(v[2] - vj[2], nz, fmaf(v[1] - vj[1], ny, (v[0] - vj[0]) * nx));
                        float f = 0.f;
                        if (d < d0) {
                            float fc = fmaf(stiffness, d0 - d, -(damping * vn));
                            if (fc < 0.f) fc = 0.f;
                            f += fc;
                        }
                        a[0] = fmaf(f, nx, a[0]);
                        a[1] = fmaf(f, ny, a[1]);
                        a[2] = fmaf(f, nz, a[2]);
                    }
                }
        out[0] = a[0];
        out[1] = a[1];
        out[2] = a[2];
        out[3] = 0.f;
    }
}

/* ------------------------------------------------------------------------ */
/* Whole step (reference order, Sim.cu:4240-4525)                            */
/* ------------------------------------------------------------------------ */

/* Gas cooling, kAccumCellVel + kCoolCell (Sim.cu:1212-1366), over particles
 * sorted by cell (keys = cell key + 1, 0 inactive), each cell's run summed in
 * sorted order (the reference's float atomics take some order; this is one).
 * Mean velocity over the cell's particles that are not dark matter (vel.w >=
 * -50, finite velocities); cells under two particles are skipped. Every such
 * particle adds (mean - v)^2 to the cell's dispersion sums (disp4: x, y, z,
 * count grids of G^3 each, zeroed by the caller); gas (pos.w <= 0) moves
 * toward the mean by k = min(rate dt 60, 0.5), "v + k dv" fused as nvcc does. */
void sdo_cool_cells(const float* pos4, float* vel4, const uint32_t* keys, int n, int G, float rate, float dt,
                    float* disp4) {
    const size_t cells = (size_t)G * G * G;
    int i = 0;
    while (i < n) {
        const uint32_t key = keys[i];
        int end = i + 1;
        while (end < n && keys[end] == key) ++end;
        if (key != 0u) {
            float sx = 0.f, sy = 0.f, sz = 0.f, nc = 0.f;
            for (int j = i; j < end; ++j) {
                const float* p = pos4 + 4 * (size_t)j;
                const float* v = vel4 + 4 * (size_t)j;
                if (p[0] < 0.f) continue;
                if (!isfinite(v[0]) || !isfinite(v[1]) || !isfinite(v[2])) continue;
                if (v[3] < -50.f) continue;
                sx += v[0];
                sy += v[1];
                sz += v[2];
                nc += 1.0f;
            }
            if (nc >= 2.f) {
                const float* p0 = pos4 + 4 * (size_t)i;
                const float fG = (float)G;
                const int cx = clampi((int)(p0[0] * fG), 0, G - 1);
                const int cy = clampi((int)(p0[1] * fG), 0, G - 1);
                const int cz = clampi((int)(p0[2] * fG), 0, G - 1);
                const size_t c = ((size_t)cz * G + cy) * G + cx;
                const float inv = 1.f / nc;
                const float mvx = sx * inv, mvy = sy * inv, mvz = sz * inv;
                float k = rate * dt * 60.0f;
                if (k > 0.5f) k = 0.5f;
                float qx = 0.f, qy = 0.f, qz = 0.f, qc = 0.f;
                for (int j = i; j < end; ++j) {
                    const float* p = pos4 + 4 * (size_t)j;
                    float* v = vel4 + 4 * (size_t)j;
                    if (p[0] < 0.f) continue;
                    if (v[3] < -50.f) continue;
                    const float dvx = mvx - v[0], dvy = mvy - v[1], dvz = mvz - v[2];
                    qx += dvx * dvx;
                    qy += dvy * dvy;
                    qz += dvz * dvz;
                    qc += 1.0f;
                    if (rate > 0.f && p[3] <= 0.f) {
                        v[0] = fmaf(k, dvx, v[0]);
                        v[1] = fmaf(k, dvy, v[1]);
                        v[2] = fmaf(k, dvz, v[2]);
                    }
                }
                disp4[c] = qx;
                disp4[cells + c] = qy;
                disp4[2 * cells + c] = qz;
                disp4[3 * cells + c] = qc;
            }
        }
        i = end;
    }
}

/* kPressure, Sim.cu:1368-1399: per cell with dispersion, a = -(D[+1] - D[-1]) *
 * (0.5 G k / n) on each axis's own dispersion grid, clamped to +-maxAcc; zero
 * elsewhere. Isolated edges clamp the stencil, periodic wraps. */
void sdo_pressure(const float* disp4, int G, int periodic, float k, float maxAcc, float* accP4) {
    const size_t cells = (size_t)G * G * G;
    const float* dX = disp4;
    const float* dY = disp4 + cells;
    const float* dZ = disp4 + 2 * cells;
    const float* dC = disp4 + 3 * cells;
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                const size_t c = ((size_t)z * G + y) * G + x;
                float* a = accP4 + 4 * c;
                a[0] = a[1] = a[2] = a[3] = 0.f;
                const float rho = dC[c];
                if (rho < 1e-6f) continue;
                int xm, xp, ym, yp, zm, zp;
                if (periodic) {
                    xm = (x - 1 + G) % G; xp = (x + 1) % G;
                    ym = (y - 1 + G) % G; yp = (y + 1) % G;
                    zm = (z - 1 + G) % G; zp = (z + 1) % G;
                } else {
                    xm = x > 0 ? x - 1 : 0; xp = x < G - 1 ? x + 1 : G - 1;
                    ym = y > 0 ? y - 1 : 0; yp = y < G - 1 ? y + 1 : G - 1;
                    zm = z > 0 ? z - 1 : 0; zp = z < G - 1 ? z + 1 : G - 1;
                }
                const float h = 0.5f * (float)G * k / rho;
                float ax = -(dX[((size_t)z * G + y) * G + xp] - dX[((size_t)z * G + y) * G + xm]) * h;
                float ay = -(dY[((size_t)z * G + yp) * G + x] - dY[((size_t)z * G + ym) * G + x]) * h;
                float az = -(dZ[((size_t)zp * G + y) * G + x] - dZ[((size_t)zm * G + y) * G + x]) * h;
                a[0] = fminf(fmaxf(ax, -maxAcc), maxAcc);
                a[1] = fminf(fmaxf(ay, -maxAcc), maxAcc);
                a[2] = fminf(fmaxf(az, -maxAcc), maxAcc);
            }
}

/* dt policy, Sim.cu:4259-4283 with dtScaleFor (Sim.h:89-93):
 * dt = min(0.0016 * dtScaleFor(timeScale), 0.8 * (1/G) / vmax), contact caps
 * dt at 0.5/sqrt(stiffness). fixedDt > 0 replaces the first term and the CFL
 * term (the fixed-step mode used by the benchmark configurations). */
float sdo_dt_policy(const sdo_params* prm, float vmax) {
    if (prm->fixedDt > 0.f) return prm->fixedDt;
    const float ts = prm->timeScale;
    const float kMaxStepReps = 3.0f;
    const float scale = ts <= 1.0f ? ts : (ts <= kMaxStepReps ? 1.0f : ts / kMaxStepReps);
    const float cell = 1.0f / (float)prm->G;
    float dt = 0.0016f * scale;
    const float dtMax = (vmax > 1e-6f) ? (0.8f * cell / vmax) : dt;
    if (dt > dtMax) dt = dtMax;
    if (prm->contact) {
        const float k = prm->contactStiffness > 1.0f ? prm->contactStiffness : 1.0f;
        const float dtStiff = 0.5f / sqrtf(k);
        if (dt > dtStiff) dt = dtStiff;
    }
    return dt;
}

/* computeAccel, Sim.cu:3605-3633: deposit -> Poisson -> gradient. potScale =
 * gravity/allocN (Sim.cu:3107-3109). The density fed to the FP64 solver is the
 * fixed-point grid / 2^24, or the particle-order float deposit. */
void sdo_compute_accel(const sdo_params* prm, const float* pos4, int n, float* accG4,
                       int64_t* rhoFixOut, double* potOut) {
    const int G = prm->G;
    const size_t cells = (size_t)G * G * G;
    const float potScale = prm->gravity / (float)(prm->allocN > 0 ? prm->allocN : 1);
    if (prm->fp32Mesh && !potOut) {
        /* the CPU baseline's single-precision mesh */
        int64_t* rf = rhoFixOut ? rhoFixOut : (int64_t*)malloc(cells * sizeof(int64_t));
        float* r32 = (float*)malloc(cells * sizeof(float));
        float* p32 = (float*)malloc(cells * sizeof(float));
        sdo_deposit_fixed(pos4, n, G, prm->periodic, rf);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)cells; ++i) r32[i] = (float)rf[i] * (1.0f / (float)SD_FIX_ONE);
        if (prm->periodic)
            sdo_poisson_periodic_f32(r32, G, prm->softeningCells, p32);
        else
            sdo_poisson_iso_f32(r32, G, prm->softeningCells, p32);
        sdo_grid_accel_f32(p32, G, potScale, prm->periodic, accG4);
        if (!rhoFixOut) free(rf);
        free(r32);
        free(p32);
        return;
    }
    double* rho = (double*)malloc(cells * sizeof(double));
    double* pot = potOut ? potOut : (double*)malloc(cells * sizeof(double));
    if (prm->fixedDeposit) {
        int64_t* rf = rhoFixOut ? rhoFixOut : (int64_t*)malloc(cells * sizeof(int64_t));
        sdo_deposit_fixed(pos4, n, G, prm->periodic, rf);
        for (size_t i = 0; i < cells; ++i) rho[i] = (double)rf[i] / (double)SD_FIX_ONE;
        if (!rhoFixOut) free(rf);
    } else {
        float* rf = (float*)malloc(cells * sizeof(float));
        sdo_deposit_float(pos4, n, G, G, prm->periodic, rf);
        for (size_t i = 0; i < cells; ++i) rho[i] = (double)rf[i];
        free(rf);
    }
    if (prm->periodic)
        sdo_poisson_periodic(rho, G, prm->softeningCells, pot);
    else
        sdo_poisson_iso(rho, G, prm->softeningCells, pot);
    sdo_grid_accel(pot, G, potScale, prm->periodic, accG4);
    free(rho);
    if (!potOut) free(pot);
}

/* keyMode 1 sorts the step by the CIC-node Morton key and the contact table
 * by the cell Morton key; keyMode 0 uses kCellKey for both. */
static void sort_particles(const sdo_params* prm, float* pos4, float* vel4, int n,
                           float* tmpP, float* tmpV, uint32_t* keys, int32_t* idx, int forContact) {
    const int mode = prm->keyMode ? (forContact ? SDO_KEY_MORTON : SDO_KEY_MORTON_NODE) : SDO_KEY_ROWMAJOR;
    sdo_cell_keys(pos4, n, prm->G, mode, keys, idx);
    sdo_stable_sort_pairs(keys, idx, n);
    sdo_reorder4(pos4, idx, n, tmpP);
    sdo_reorder4(vel4, idx, n, tmpV);
    memcpy(pos4, tmpP, sizeof(float) * 4 * (size_t)n);
    memcpy(vel4, tmpV, sizeof(float) * 4 * (size_t)n);
}

/* Sim::step, Sim.cu:4240-4525, plain-PM flags: sort every sortInterval steps
 * (skipped when contact sorts itself), computeAccel, contact (sort + cell
 * ranges + kContact), CFL dt, integrate, momentum removal, time advance. */
int sdo_step(const sdo_params* prm, sdo_state* st, float* pos4, float* vel4, int n,
             float* scratchPos4, float* scratchVel4, float* gas4) {
    const int G = prm->G;
    const size_t cells = (size_t)G * G * G;
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (prm->sortInterval > 0 && (st->stepCount % prm->sortInterval) == 0 && !prm->contact)
        sort_particles(prm, pos4, vel4, n, scratchPos4, scratchVel4, keys, idx, 0);
    float* accG = (float*)malloc(sizeof(float) * 4 * cells);
    sdo_compute_accel(prm, pos4, n, accG, NULL, NULL);
    /* the pressure grid from the last cooling pass's dispersion (Sim.cu:3617-3631) */
    const int gas = (prm->pressureEnabled || prm->coolingEnabled) && gas4;
    float* accP = NULL;
    if (gas && prm->pressureEnabled) {
        accP = (float*)malloc(sizeof(float) * 4 * cells);
        sdo_pressure(gas4, G, prm->periodic, prm->pressureK, 5.0f, accP);
    }
    float* accC = NULL;
    if (prm->contact && !((double)prm->allocN > 0.764 * (double)G * G * G)) {
        sort_particles(prm, pos4, vel4, n, scratchPos4, scratchVel4, keys, idx, 1);
        /* Morton cell codes of a G that is not a power of two run over the
         * enclosing power-of-two cube */
        size_t tableCells = cells;
        if (prm->keyMode == SDO_KEY_MORTON) {
            size_t g = 1;
            while (g < (size_t)G) g <<= 1;
            tableCells = g * g * g;
        }
        int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * tableCells);
        int32_t* ce = (int32_t*)malloc(sizeof(int32_t) * tableCells);
        sdo_cell_range(keys, n, (int)tableCells, cs, ce);
        accC = (float*)malloc(sizeof(float) * 4 * (size_t)(n > 0 ? n : 1));
        sdo_contact(pos4, vel4, n, G, prm->periodic, prm->keyMode, cs, ce,
                    prm->contactRadiusCells / (float)G, prm->contactStiffness,
                    prm->contactDamping, accC);
        free(cs);
        free(ce);
    }
    const float vmax = sdo_max_speed(pos4, vel4, n);
    const float dt = sdo_dt_policy(prm, vmax);
    if (gas) {
        /* doCooling (Sim.cu:3732-3830): every 1 / 2 / 8 steps, dt times that */
        const int every = prm->contact ? 1 : (prm->pressureEnabled ? 2 : 8);
        if ((st->stepCount % every) == 0) {
            if (!accC) sort_particles(prm, pos4, vel4, n, scratchPos4, scratchVel4, keys, idx, 1);
            memset(gas4, 0, sizeof(float) * 4 * cells);
            const float rate = prm->coolingEnabled && prm->coolingRate > 0.f ? prm->coolingRate : 0.f;
            sdo_cool_cells(pos4, vel4, keys, n, G, rate, dt * (float)every, gas4);
        }
    }
    sdo_integrate(accG, pos4, vel4, n, G, dt, dt, prm->periodic, prm->c2, accC,
                  prm->periodic ? 0.f : (prm->softBoundR > 0.f ? prm->softBoundR : 0.f),
                  prm->darkEnergy, NULL, accP);
    free(accP);
    if (prm->momentumZero) {
        double m[4];
        sdo_momentum(pos4, vel4, n, m);
        sdo_subtract_mean(pos4, vel4, n, m);
    }
    st->simTime += dt;
    st->stepCount += 1;
    st->lastDt = dt;
    st->lastVmax = vmax;
    free(accG);
    free(accC);
    free(keys);
    free(idx);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Direct sum and the force-error metric (method of proto/cuda/x3.cu)        */
/* ------------------------------------------------------------------------ */

/* Plummer-softened direct sum (x3.cu:84-103 in 3D, FP64):
 * a_i = gm * sum_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^(3/2), j over all live
 * particles (the j == i term is zero). */
void sdo_direct_sum(const float* pos4, int n, const int32_t* targets, int nt, double gm,
                    double eps, double* acc3) {
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < nt; ++t) {
        const float* p = pos4 + 4 * (size_t)targets[t];
        double ax = 0, ay = 0, az = 0;
        for (int j = 0; j < n; ++j) {
            const float* q = pos4 + 4 * (size_t)j;
            if (q[0] < 0.f) continue;
            const double dx = (double)q[0] - p[0], dy = (double)q[1] - p[1],
                         dz = (double)q[2] - p[2];
            const double r2 = dx * dx + dy * dy + dz * dz + e2;
            if (r2 <= 0.0) continue;
            const double inv = 1.0 / sqrt(r2);
            const double inv3 = inv * inv * inv;
            ax += dx * inv3;
            ay += dy * inv3;
            az += dz * inv3;
        }
        acc3[3 * (size_t)t + 0] = gm * ax;
        acc3[3 * (size_t)t + 1] = gm * ay;
        acc3[3 * (size_t)t + 2] = gm * az;
    }
}

/* rel-RMS = sqrt(sum |a - t|^2) / sqrt(sum |t|^2) (un-fitted). lsqScale, if
 * given, receives x3.cu:168-177's least-squares factor a = <p,t>/<p,p>. */
double sdo_rel_rms(const double* approx3, const double* truth3, int n, double* lsqScale) {
    double se = 0, sr = 0, num = 0, den = 0;
    for (size_t i = 0; i < 3 * (size_t)n; ++i) {
        const double d = approx3[i] - truth3[i];
        se += d * d;
        sr += truth3[i] * truth3[i];
        num += approx3[i] * truth3[i];
        den += approx3[i] * approx3[i];
    }
    if (lsqScale) *lsqScale = den > 0 ? num / den : 0.0;
    return sr > 0 ? sqrt(se / sr) : 0.0;
}

/* ------------------------------------------------------------------------ */
/* MLP (no reference counterpart): 30 -> H -> H -> H -> 3                   */
/* ------------------------------------------------------------------------ */
/* Layout: W = [W1 (H x 30) | W2 (H x H) | W3 (H x H) | W4 (3 x H)], row-major
 * (out x in); b = [b1 (H) | b2 (H) | b3 (H) | b4 (3)]. Activation GELU in its
 * tanh form: gelu(u) = 0.5 u (1 + tanh(sqrt(2/pi) (u + 0.044715 u^3))).
 * Loss = 0.5/nb * sum |y - target|^2. */

#define MLP_IN 30
#define MLP_OUT 3

size_t sdo_mlp_wcount(int H) { return (size_t)H * MLP_IN + 2 * (size_t)H * H + 3 * (size_t)H; }
size_t sdo_mlp_bcount(int H) { return 3 * (size_t)H + 3; }

static double gelu(double u) {
    const double c = 0.7978845608028654;
    return 0.5 * u * (1.0 + tanh(c * (u + 0.044715 * u * u * u)));
}
static double gelu_d(double u) {
    const double c = 0.7978845608028654;
    const double inner = c * (u + 0.044715 * u * u * u);
    const double t = tanh(inner);
    const double dinner = c * (1.0 + 3.0 * 0.044715 * u * u);
    return 0.5 * (1.0 + t) + 0.5 * u * (1.0 - t * t) * dinner;
}

static void mlp_layers(int H, const double* W, const double* b, const double** Wl,
                       const double** bl, int* din, int* dout) {
    Wl[0] = W;
    Wl[1] = W + (size_t)H * MLP_IN;
    Wl[2] = Wl[1] + (size_t)H * H;
    Wl[3] = Wl[2] + (size_t)H * H;
    bl[0] = b;
    bl[1] = b + H;
    bl[2] = b + 2 * H;
    bl[3] = b + 3 * H;
    din[0] = MLP_IN; dout[0] = H;
    din[1] = H; dout[1] = H;
    din[2] = H; dout[2] = H;
    din[3] = H; dout[3] = MLP_OUT;
}

void sdo_mlp_forward(const double* W, const double* b, int H, const double* x, int nb,
                     double* y) {
    const double* Wl[4];
    const double* bl[4];
    int din[4], dout[4];
    mlp_layers(H, W, b, Wl, bl, din, dout);
    double* h0 = (double*)malloc(sizeof(double) * (size_t)(H > MLP_IN ? H : MLP_IN));
    double* h1 = (double*)malloc(sizeof(double) * (size_t)H);
    for (int s = 0; s < nb; ++s) {
        memcpy(h0, x + (size_t)s * MLP_IN, sizeof(double) * MLP_IN);
        for (int l = 0; l < 4; ++l) {
            double* out = (l == 3) ? (y + (size_t)s * MLP_OUT) : h1;
            for (int o = 0; o < dout[l]; ++o) {
                double acc = bl[l][o];
                for (int i = 0; i < din[l]; ++i) acc += Wl[l][(size_t)o * din[l] + i] * h0[i];
                out[o] = (l == 3) ? acc : gelu(acc);
            }
            if (l < 3) memcpy(h0, h1, sizeof(double) * (size_t)H);
        }
    }
    free(h0);
    free(h1);
}

double sdo_mlp_loss_grad(const double* W, const double* b, int H, const double* x,
                         const double* target, int nb, double* gW, double* gb) {
    const double* Wl[4];
    const double* bl[4];
    int din[4], dout[4];
    mlp_layers(H, W, b, Wl, bl, din, dout);
    const size_t nw = sdo_mlp_wcount(H), nbias = sdo_mlp_bcount(H);
    memset(gW, 0, sizeof(double) * nw);
    memset(gb, 0, sizeof(double) * nbias);
    double* gWl[4];
    double* gbl[4];
    gWl[0] = gW;
    gWl[1] = gW + (size_t)H * MLP_IN;
    gWl[2] = gWl[1] + (size_t)H * H;
    gWl[3] = gWl[2] + (size_t)H * H;
    gbl[0] = gb;
    gbl[1] = gb + H;
    gbl[2] = gb + 2 * H;
    gbl[3] = gb + 3 * H;
    /* per-sample activations: a[l] = input of layer l, u[l] = pre-activation */
    double* act = (double*)malloc(sizeof(double) * (size_t)(MLP_IN + 3 * H + 3 * H + 3));
    double* dcur = (double*)malloc(sizeof(double) * (size_t)H);
    double* dprev = (double*)malloc(sizeof(double) * (size_t)(H > MLP_IN ? H : MLP_IN));
    double loss = 0.0;
    const double inv = 1.0 / (double)(nb > 0 ? nb : 1);
    for (int s = 0; s < nb; ++s) {
        double* a0 = act;                /* 30 */
        double* u1 = a0 + MLP_IN;        /* H pre-act layer 0 */
        double* u2 = u1 + H;
        double* u3 = u2 + H;
        double* a1 = u3 + H;
        double* a2 = a1 + H;
        double* a3 = a2 + H;
        double* yv = a3 + H;             /* 3 */
        double* ins[4] = {a0, a1, a2, a3};
        double* us[3] = {u1, u2, u3};
        memcpy(a0, x + (size_t)s * MLP_IN, sizeof(double) * MLP_IN);
        for (int l = 0; l < 4; ++l) {
            for (int o = 0; o < dout[l]; ++o) {
                double acc = bl[l][o];
                for (int i = 0; i < din[l]; ++i) acc += Wl[l][(size_t)o * din[l] + i] * ins[l][i];
                if (l == 3) yv[o] = acc;
                else {
                    us[l][o] = acc;
                    ins[l + 1][o] = gelu(acc);
                }
            }
        }
        for (int o = 0; o < MLP_OUT; ++o) {
            const double d = yv[o] - target[(size_t)s * MLP_OUT + o];
            loss += 0.5 * d * d * inv;
            dcur[o] = d * inv;
        }
        for (int l = 3; l >= 0; --l) {
            for (int o = 0; o < dout[l]; ++o) {
                gbl[l][o] += dcur[o];
                for (int i = 0; i < din[l]; ++i) gWl[l][(size_t)o * din[l] + i] += dcur[o] * ins[l][i];
            }
            if (l == 0) break;
            for (int i = 0; i < din[l]; ++i) {
                double acc = 0.0;
                for (int o = 0; o < dout[l]; ++o) acc += Wl[l][(size_t)o * din[l] + i] * dcur[o];
                dprev[i] = acc * gelu_d(us[l - 1][i]);
            }
            memcpy(dcur, dprev, sizeof(double) * (size_t)din[l]);
        }
    }
    free(act);
    free(dcur);
    free(dprev);
    return loss;
}
