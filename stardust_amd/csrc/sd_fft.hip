// sd_fft.hip -- pruned 3D FFT Poisson solve for gfx950.
//
// Replaces the reference's solvePoisson (src/sim/Sim.cu:3586-3603: cuFFT R2C,
// kMulSpec / kPoissonPeriodic, cuFFT C2R) and buildGreen (Sim.cu:3569-3576 +
// kGreen 563-573). Instead of three full S^3 cuFFT transforms plus a separate
// multiply, the solve is five passes over HBM, each an LDS-resident batch of
// one-dimensional Stockham transforms (Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC08: iteration
// v = x[j + r N/R] * W^{r (j mod Ns)}, R-point DFT, store at
// (j/Ns) Ns R + (j mod Ns) + r Ns):
//
//   X  rows     : fixed-point density (G^3, int64) -> A[z][y][kx]   (R2C, input pruned to G)
//   Y  columns  : A -> B[z][ky][kx]                                  (input pruned to G)
//   Z  columns  : B -> B, forward * Green * inverse fused in LDS     (only z < G kept)
//   Y' columns  : B -> A                                             (only y < G kept)
//   X' rows     : A -> pot[z][y][x]                                  (C2R, only x < G kept;
//                                                                    two rows per complex line)
//
// At S=256 the transforms are pruned too: a forward line's first radix stage
// skips the zero-padded inputs, an inverse line's last stage the outputs that
// are not kept.
//
// Complex arrays keep kx (S/2+1 values) contiguous with a row pitch P that is
// a multiple of 16 complex (128 B), so every column tile of 16 is one cache
// line. Packed layout (the one-tile column kernels, S < 512): the real
// kx = S/2 bin rides in the imaginary part of the real kx = 0 bin, so the
// column passes see S/2 columns (whole tiles) instead of S/2 + 1 (a ninth
// tile with one live column). The y pass separates the two real sequences by
// Hermitian symmetry within its line (ky with S-ky): kx = 0 keeps ky <= S/2 in
// place, the kx = S/2 values of ky = 1 .. S/2-1 go to slots S-ky (slots 0 and
// S/2 hold both, as real and imaginary parts); the z pass applies each its own
// multiplier, and the inverse y pass puts the packed line back together. The isolated Green spectrum is real and even in each axis; only its
// octant kx, ky, kz <= S/2 is stored, pre-multiplied by 1/S^3.
#include "sd_device.h"

namespace sd {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---- small DFTs in registers ---------------------------------------------
// UZ: the upper half of the inputs is zero (zero padding), so the innermost
// 4-point transforms drop the additions of zeros.
template <int R, bool INV, bool UZ = false>
struct Dft;

template <bool INV>
struct Dft<2, INV, false> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<3, INV, false> {
    __device__ __forceinline__ static void run(float2* v) {
        const float s3 = 0.86602540378443864676f;
        const float2 t = cadd(v[1], v[2]);
        const float2 d = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        // forward: X1 = m - i s3 d, X2 = m + i s3 d
        const float2 id = make_float2(-s3 * d.y, s3 * d.x);  // i * s3 * d
        v[0] = cadd(v[0], t);
        if (INV) {
            v[1] = cadd(m, id);
            v[2] = csub(m, id);
        } else {
            v[1] = csub(m, id);
            v[2] = cadd(m, id);
        }
    }
};

template <bool INV, bool UZ>
struct Dft<4, INV, UZ> {
    __device__ __forceinline__ static void run(float2* v) {
        const float2 t0 = UZ ? v[0] : cadd(v[0], v[2]), t1 = UZ ? v[0] : csub(v[0], v[2]);
        const float2 t2 = UZ ? v[1] : cadd(v[1], v[3]), t3 = rot90<INV>(UZ ? v[1] : csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV, bool UZ>
struct Dft<8, INV, UZ> {
    __device__ __forceinline__ static void run(float2* v) {
        const float c = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV, UZ>::run(e);
        Dft<4, INV, UZ>::run(o);
        // o[k] *= W8^k, W8 = exp(-+ i pi/4)
        const float sg = INV ? 1.f : -1.f;
        o[1] = cmul(o[1], make_float2(c, sg * c));
        o[2] = rot90<INV>(o[2]);
        o[3] = cmul(o[3], make_float2(-c, sg * c));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <bool INV, bool UZ>
struct Dft<16, INV, UZ> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        Dft<8, INV, UZ>::run(e);
        Dft<8, INV, UZ>::run(o);
        const float sg = INV ? 1.f : -1.f;
        // W16^k = (cos(pi k/8), -+ sin(pi k/8))
        const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
        const float c2 = 0.70710678118654752440f;
        o[1] = cmul(o[1], make_float2(c1, sg * s1));
        o[2] = cmul(o[2], make_float2(c2, sg * c2));
        o[3] = cmul(o[3], make_float2(s1, sg * c1));
        o[4] = rot90<INV>(o[4]);
        o[5] = cmul(o[5], make_float2(-s1, sg * c1));
        o[6] = cmul(o[6], make_float2(-c2, sg * c2));
        o[7] = cmul(o[7], make_float2(-c1, sg * s1));
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 8] = csub(e[k], o[k]);
        }
    }
};

// 32 = 2 x 16: even and odd samples, then the W32^k butterflies.
template <bool INV, bool UZ>
struct Dft<32, INV, UZ> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 e[16], o[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        Dft<16, INV, UZ>::run(e);
        Dft<16, INV, UZ>::run(o);
        // W32^k = (cos(pi k / 16), -+ sin(pi k / 16)), k < 8; k >= 8 from k - 8 times -+ i
        const float sg = INV ? 1.f : -1.f;
        const float cs[8] = {1.0f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
                             0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f,
                             0.19509032201612826785f};
#pragma unroll
        for (int k = 1; k < 16; ++k) {
            const int j = k & 7;
            // W32^j = (cs[j], sg cs[8 - j]) for j < 8 (sin(pi j / 16) = cos(pi (8 - j) / 16))
            const float2 w = make_float2(cs[j], j == 0 ? 0.f : sg * cs[8 - j]);
            float2 ok = j == 0 ? o[k] : cmul(o[k], w);
            if (k >= 8) ok = rot90<INV>(ok);
            o[k] = ok;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 16] = csub(e[k], o[k]);
        }
    }
};

// 12 = 3 x 4 (Cooley-Tukey): n = n1 + 3 n2, k = k2 + 4 k1; four-point
// transforms over n2, twiddles W12^(n1 k2), three-point transforms over n1.
// UZ: inputs n >= 6 are zero (only n2 = 0, 1 carry data).
template <bool INV, bool UZ>
struct Dft<12, INV, UZ> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a[3][4];
#pragma unroll
        for (int n1 = 0; n1 < 3; ++n1) {
#pragma unroll
            for (int n2 = 0; n2 < 4; ++n2) a[n1][n2] = v[n1 + 3 * n2];
            Dft<4, INV, UZ>::run(a[n1]);
        }
        // W12^j = (cos(pi j / 6), -+ sin(pi j / 6)) for j = n1 k2 in {1, 2, 3, 4, 6}
        const float sg = INV ? 1.f : -1.f;
        const float c = 0.86602540378443864676f;
        a[1][1] = cmul(a[1][1], make_float2(c, sg * 0.5f));
        a[1][2] = cmul(a[1][2], make_float2(0.5f, sg * c));
        a[1][3] = rot90<INV>(a[1][3]);
        a[2][1] = cmul(a[2][1], make_float2(0.5f, sg * c));
        a[2][2] = cmul(a[2][2], make_float2(-0.5f, sg * c));
        a[2][3] = make_float2(-a[2][3].x, -a[2][3].y);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            float2 b[3] = {a[0][k2], a[1][k2], a[2][k2]};
            Dft<3, INV, false>::run(b);
#pragma unroll
            for (int k1 = 0; k1 < 3; ++k1) v[k2 + 4 * k1] = b[k1];
        }
    }
};

// LDS line layout: element i lives at PADI(i) = i + i/16. The first Stockham
// stage writes with a stride of R complex values; without the pad a radix-16
// stage would put all 16 lanes of a line in one bank.
__device__ __forceinline__ int PADI(int i) { return i + (i >> 4); }
template <int N>
struct Line {
    static constexpr int LP = N + N / 16 + 1;  // padded line pitch (float2)
};

// Ordering of LDS accesses between the threads of one line. A line's T threads
// are threadIdx / T == line, so when T divides 64 they are lanes of one
// wavefront, whose LDS operations complete in issue order: a wavefront-scope
// fence (a compiler barrier, no wait) is enough. Otherwise the block barrier.
#ifndef SD_FFT_LINE_SYNC
#define SD_FFT_LINE_SYNC 1
#endif
template <int T>
__device__ __forceinline__ void line_sync() {
    if constexpr (SD_FFT_LINE_SYNC && 64 % T == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// ---- one Stockham stage over an LDS line ---------------------------------
// tw[m] = exp(-2 pi i m / N) (FP64-computed table); INV uses the conjugate.
// The last stage of a transform ends with a block barrier: what follows reads
// the lines with another thread mapping.
// ZIN (first stage, NS = 1): elements N/2 .. N-1 of the line are zero padding,
// i.e. inputs r >= R/2 of every butterfly; they are neither read nor added.
// HOUT (last stage, NS = N/R): only elements 0 .. N/2-1 are used afterwards,
// i.e. outputs r < R/2; the others are neither stored nor computed.
template <int N, int R, int NS, int T, bool INV, bool LAST = false, bool ZIN = false, bool HOUT = false>
__device__ __forceinline__ void stage(float2* __restrict__ line, const float2* __restrict__ tw,
                                      int t) {
    constexpr int NBF = N / R;
    constexpr int PER = (NBF + T - 1) / T;
    static_assert(!ZIN || (NS == 1 && R % 2 == 0), "ZIN: first stage only");
    static_assert(!HOUT || (NS * R == N && R % 2 == 0), "HOUT: last stage only");
    float2 v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * T;
        if (PER * T == NBF || j < NBF) {
#pragma unroll
            for (int r = 0; r < (ZIN ? R / 2 : R); ++r) v[q][r] = line[PADI(j + r * NBF)];
#pragma unroll
            for (int r = (ZIN ? R / 2 : R); r < R; ++r) v[q][r] = make_float2(0.f, 0.f);
        }
    }
    line_sync<T>();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + q * T;
        if (PER * T == NBF || j < NBF) {
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[r * k * (N / (NS * R))];
                    if (INV) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            Dft<R, INV, ZIN>::run(v[q]);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < (HOUT ? R / 2 : R); ++r) line[PADI(base + r * NS)] = v[q][r];
        }
    }
    if (LAST)
        __syncthreads();
    else
        line_sync<T>();
}

// Radix plans: T threads per line; every stage is a compile-time (R, Ns).
template <int N>
struct Plan;
template <>
struct Plan<16> {
    static constexpr int T = 4;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<16, 4, 1, T, INV>(l, tw, t);
        stage<16, 4, 4, T, INV, true>(l, tw, t);
    }
};
template <>
struct Plan<32> {
    static constexpr int T = 8;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<32, 4, 1, T, INV>(l, tw, t);
        stage<32, 8, 4, T, INV, true>(l, tw, t);
    }
};
template <>
struct Plan<64> {
    static constexpr int T = 8;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<64, 8, 1, T, INV>(l, tw, t);
        stage<64, 8, 8, T, INV, true>(l, tw, t);
    }
};
template <>
struct Plan<128> {
    static constexpr int T = 16;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<128, 8, 1, T, INV>(l, tw, t);
        stage<128, 16, 8, T, INV, true>(l, tw, t);
    }
};
#ifndef SD_PLAN192_T
#define SD_PLAN192_T 24
#endif
template <>
struct Plan<192> {
    static constexpr int T = SD_PLAN192_T;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<192, 3, 1, T, INV>(l, tw, t);
        stage<192, 8, 3, T, INV>(l, tw, t);
        stage<192, 8, 24, T, INV, true>(l, tw, t);
    }
};
template <>
struct Plan<256> {
    static constexpr int T = 16;
    template <bool INV, bool ZIN = false, bool HOUT = false>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<256, 16, 1, T, INV, false, ZIN>(l, tw, t);
        stage<256, 16, 16, T, INV, true, false, HOUT>(l, tw, t);
    }
};
template <>
struct Plan<512> {
    static constexpr int T = 64;
    template <bool INV>
    __device__ static void run(float2* l, const float2* tw, int t) {
        stage<512, 8, 1, T, INV>(l, tw, t);
        stage<512, 8, 8, T, INV>(l, tw, t);
        stage<512, 8, 64, T, INV, true>(l, tw, t);
    }
};

#ifndef SD_TB
#define SD_TB 16
#endif
constexpr int TB = SD_TB;  // lines per block (columns of a tile / rows of a row batch)

template <int N>
__device__ __forceinline__ void load_twiddles(float2* s_tw, const float2* __restrict__ tw) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = tw[i];
}

// Global <-> LDS staging: every thread first issues all of its tile loads
// (16-byte vectors: two complex values, or two int64 density values) into
// registers and only then writes LDS, so a block has its whole tile in flight
// at once instead of one load per round trip.

// ---- X pass: rows, real input -> half spectrum ----------------------------
// SRC 0: fixed-point density rows (int64, nin values per row, zeroed after the
//        read when clearSrc != 0 so the next deposit starts from an empty grid).
// SRC 1: the isolated Green function g = -1/(|d| + soft) of kGreen
//        (Sim.cu:563-573) generated on the fly for an S^3 grid (rows = S^2).
// SRC 2: float density rows (srcF, nin values per row, nin a multiple of 4),
//        each the value SRC 0 makes of its int64 node: (float)v * 2^-24.
struct RowsFwdArgs {
    long long* src;
    const float* srcF;
    float2* out;
    int rows;      // number of rows
    int nin;       // input values per row (even)
    int P;         // output row pitch (complex, > N/2 + 1)
    int clearSrc;
    int pack;      // packed layout: bin N/2 in the imaginary part of bin 0
    int S;         // Green: grid side
    float soft;    // Green: softening (cells)
    FinalizeJob fin;  // reduction carried by block 0 (fin.partials == nullptr: none)
};

// Rows per LDS line in the x passes. With SD_FFT_ROWS2 two real rows share one
// complex transform (row 2c in the real part, row 2c + 1 in the imaginary
// part); the forward pass splits the two half spectra by Hermitian symmetry,
// A = (Z[k] + conj Z[N-k]) / 2, B = -i (Z[k] - conj Z[N-k]) / 2, and the
// inverse pass transforms Z = A + i B (bins 0 and N/2 taken real) and keeps
// the real part for one row and the imaginary part for the other. A block
// Lines per block and rows per line are set per direction (S=256, two-disk
// step: forward 1 row x 16 lines 14.5 us, 2 x 16: 15.9, 2 x 8: 16.9 -- the
// split reads Z[N-k] across the line; inverse 1 x 16: 11.6 us, 2 x 8: 10.2,
// 2 x 16: 9.5).
#ifndef SD_FFT_RPL_FWD
#define SD_FFT_RPL_FWD 1
#endif
#ifndef SD_FFT_LINES_FWD
#define SD_FFT_LINES_FWD 16
#endif
#ifndef SD_FFT_RPL_INV
#define SD_FFT_RPL_INV 2
#endif
#ifndef SD_FFT_LINES_INV
#define SD_FFT_LINES_INV 16
#endif
template <int RPL_, int LINES_>
struct XPass {
    static constexpr int RPL = RPL_;      // rows per LDS line
    static constexpr int LINES = LINES_;  // lines per block
    static constexpr int ROWS = RPL * LINES;
};
using XFwd = XPass<SD_FFT_RPL_FWD, SD_FFT_LINES_FWD>;
using XInv = XPass<SD_FFT_RPL_INV, SD_FFT_LINES_INV>;

#ifndef SD_FFT_PRUNE
#define SD_FFT_PRUNE 1
#endif
#ifndef SD_FFT_PRUNE_ROWS
#define SD_FFT_PRUNE_ROWS 1
#endif
template <int N, bool INV>
__device__ __forceinline__ void run_line(float2* l, const float2* tw, int t, bool halfIn, bool halfOut) {
    if constexpr (SD_FFT_PRUNE && N == 256) {
        if (halfIn && halfOut) {
            Plan<N>::template run<INV, true, true>(l, tw, t);
            return;
        }
        if (halfIn) {
            Plan<N>::template run<INV, true, false>(l, tw, t);
            return;
        }
        if (halfOut) {
            Plan<N>::template run<INV, false, true>(l, tw, t);
            return;
        }
    }
    Plan<N>::template run<INV>(l, tw, t);
}

template <int N, int SRC>
__device__ __forceinline__ void rows_fwd_body(const RowsFwdArgs& a, const float2* __restrict__ tw, int bx) {
    constexpr int LP = Line<N>::LP;
    constexpr int RPL = XFwd::RPL, TBR = XFwd::LINES, TB = XFwd::ROWS;  // rows per line, lines, rows
    constexpr int THREADS = TBR * Plan<N>::T;
    __shared__ float2 s_line[TBR * LP];
    __shared__ float2 s_tw[N];
    load_twiddles<N>(s_tw, tw);
    float* s_f = reinterpret_cast<float*>(s_line);  // component h of slot e: s_f[2 e + h]
    const int row0 = bx * TB;
    // zero the padded half of every line
    for (int i = threadIdx.x; i < TBR * N; i += THREADS) {
        const int c = i / N, x = i - c * N;
        if (x >= a.nin) s_line[c * LP + PADI(x)] = make_float2(0.f, 0.f);
    }
    if (SRC == 0) {
        constexpr int ITERS = (TB * (N / 2) + THREADS - 1) / THREADS;  // int64 pairs of the TB rows
        const int pairs = a.nin / 2;
        longlong2 buf[ITERS];
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int q = threadIdx.x + it * THREADS;
            const int r = q / pairs, j = q - r * pairs;
            buf[it] = make_longlong2(0, 0);
            if (r < TB && row0 + r < a.rows) {
                longlong2* p = reinterpret_cast<longlong2*>(a.src + (size_t)(row0 + r) * a.nin) + j;
                buf[it] = *p;
            }
        }
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int q = threadIdx.x + it * THREADS;
            const int r = q / pairs, j = q - r * pairs;
            if (r < TB) {
                const float sc = 1.0f / (float)SD_FIX_ONE;
                const int c = r / RPL, h = r % RPL;
                if constexpr (RPL == 1) {
                    s_line[c * LP + PADI(2 * j)] = make_float2((float)buf[it].x * sc, 0.f);
                    s_line[c * LP + PADI(2 * j + 1)] = make_float2((float)buf[it].y * sc, 0.f);
                } else {
                    s_f[2 * (c * LP + PADI(2 * j)) + h] = (float)buf[it].x * sc;
                    s_f[2 * (c * LP + PADI(2 * j + 1)) + h] = (float)buf[it].y * sc;
                }
                if (a.clearSrc && row0 + r < a.rows)
                    reinterpret_cast<longlong2*>(a.src + (size_t)(row0 + r) * a.nin)[j] = make_longlong2(0, 0);
            }
        }
    } else if (SRC == 2) {
        constexpr int ITERS = (TB * (N / 4) + THREADS - 1) / THREADS;  // float quads of the TB rows
        const int quads = a.nin / 4;
        float4 buf[ITERS];
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int q = threadIdx.x + it * THREADS;
            const int r = q / quads, j = q - r * quads;
            buf[it] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < TB && row0 + r < a.rows)
                buf[it] = reinterpret_cast<const float4*>(a.srcF + (size_t)(row0 + r) * a.nin)[j];
        }
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int q = threadIdx.x + it * THREADS;
            const int r = q / quads, j = q - r * quads;
            if (r < TB) {
                const int c = r / RPL, h = r % RPL;
                const float e[4] = {buf[it].x, buf[it].y, buf[it].z, buf[it].w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if constexpr (RPL == 1)
                        s_line[c * LP + PADI(4 * j + k)] = make_float2(e[k], 0.f);
                    else
                        s_f[2 * (c * LP + PADI(4 * j + k)) + h] = e[k];
                }
            }
        }
    } else {
        for (int i = threadIdx.x; i < TB * a.nin; i += THREADS) {
            const int r = i / a.nin, x = i - r * a.nin;
            const int row = row0 + r;
            float val = 0.f;
            if (row < a.rows) {
                const int y = row % a.S, z = row / a.S;
                const int S = a.S, h = S / 2;
                const int dx = x <= h ? x : S - x;
                const int dy = y <= h ? y : S - y;
                const int dz = z <= h ? z : S - z;
                const float rr = sqrtf((float)(dx * dx + dy * dy + dz * dz)) + a.soft;
                val = -1.0f / rr;
            }
            const int c = r / RPL;
            if constexpr (RPL == 1)
                s_line[c * LP + PADI(x)] = make_float2(val, 0.f);
            else
                s_f[2 * (c * LP + PADI(x)) + (r % RPL)] = val;
        }
    }
    __syncthreads();
    const int c = threadIdx.x / Plan<N>::T, t = threadIdx.x % Plan<N>::T;
    run_line<N, false>(s_line + c * LP, s_tw, t, SD_FFT_PRUNE_ROWS && 2 * a.nin <= N, false);  // zero-padded rows
    // store k = 0 .. N/2 as complex pairs (the pair's second half at N/2 + 1
    // lands in the row padding)
    constexpr int KP2 = N / 4 + 1;
    constexpr int OITERS = (TBR * KP2 + THREADS - 1) / THREADS;
#pragma unroll
    for (int it = 0; it < OITERS; ++it) {
        const int q = threadIdx.x + it * THREADS;
        const int cc = q / KP2, j = q - cc * KP2;
        if (cc >= TBR || (a.pack && j == N / 4)) continue;
        const float2* l = s_line + cc * LP;
        if constexpr (RPL == 1) {
            if (row0 + cc < a.rows) {
                float2 u = l[PADI(2 * j)];
                const float2 v = l[PADI(2 * j + 1)];
                if (a.pack && j == 0) u.y = l[PADI(N / 2)].x;  // both bins are real
                reinterpret_cast<float4*>(a.out + (size_t)(row0 + cc) * a.P)[j] = make_float4(u.x, u.y, v.x, v.y);
            }
        } else {
            float2 A[2], B[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int k = 2 * j + e;
                const float2 z = l[PADI(k)], m = l[PADI(k == 0 ? 0 : N - k)];  // k <= N/2 + 1
                A[e] = make_float2(0.5f * (z.x + m.x), 0.5f * (z.y - m.y));
                B[e] = make_float2(0.5f * (z.y + m.y), 0.5f * (m.x - z.x));
            }
            if (a.pack && j == 0) {  // bin N/2 of each row: the real and imaginary parts of Z[N/2]
                const float2 zh = l[PADI(N / 2)];
                A[0].y = zh.x;
                B[0].y = zh.y;
            }
            const int ra = row0 + 2 * cc;
            if (ra < a.rows)
                reinterpret_cast<float4*>(a.out + (size_t)ra * a.P)[j] = make_float4(A[0].x, A[0].y, A[1].x, A[1].y);
            if (ra + 1 < a.rows)
                reinterpret_cast<float4*>(a.out + (size_t)(ra + 1) * a.P)[j] =
                    make_float4(B[0].x, B[0].y, B[1].x, B[1].y);
        }
    }
}

// A pass can carry the previous integrate's partials reduction (finalize_body)
// as one more block, which runs beside the transforms: the reduction must only
// be done before the next integrate, and no mesh pass reads the step scalars.
// The x pass carries it when its blocks have FIN threads; otherwise the fused
// z pass does when its blocks have at least 256 (fewer threads would take
// longer than the pass).
template <int N>
constexpr bool rows_fwd_carries_finalize() {
    return XFwd::LINES * Plan<N>::T == FIN;
}

template <int N, int SRC>
__global__ void __launch_bounds__(XFwd::LINES * Plan<N>::T) k_rows_fwd(RowsFwdArgs a, const float2* __restrict__ tw) {
    if constexpr (rows_fwd_carries_finalize<N>()) {
        if (a.fin.partials) {
            if (blockIdx.x == 0) {
                finalize_body<FIN>(a.fin.partials, a.fin.nparts, a.fin.mode, a.fin.pol, a.fin.sc);
                return;
            }
            rows_fwd_body<N, SRC>(a, tw, blockIdx.x - 1);
            return;
        }
    }
    rows_fwd_body<N, SRC>(a, tw, blockIdx.x);
}

// One launch over the X pass of many meshes of the same FFT side: blockIdx.y
// selects the mesh (every job has the same row count).
template <int N>
__global__ void __launch_bounds__(XFwd::LINES * Plan<N>::T) k_rows_fwd_batch(const RowsFwdArgs* __restrict__ jobs,
                                                                    const float2* __restrict__ tw) {
    const RowsFwdArgs a = jobs[blockIdx.y];
    rows_fwd_body<N, 0>(a, tw, blockIdx.x);
}

// ---- X' pass: rows, half spectrum -> real output --------------------------
struct RowsInvArgs {
    const float2* in;
    float* out;
    int rows;
    int P;
    int nout;  // real values kept per row (row stride of out), multiple of 4
    int pack;  // packed layout: bin N/2 in the imaginary part of bin 0
};

template <int N>
__device__ __forceinline__ void rows_inv_body(const RowsInvArgs& a, const float2* __restrict__ tw, int bx) {
    constexpr int LP = Line<N>::LP;
    constexpr int RPL = XInv::RPL, TBR = XInv::LINES, TB = XInv::ROWS;  // rows per line, lines, rows
    constexpr int THREADS = TBR * Plan<N>::T;
    __shared__ float2 s_line[TBR * LP];
    __shared__ float2 s_tw[N];
    load_twiddles<N>(s_tw, tw);
    const int row0 = bx * TB;
    constexpr int KP2 = N / 4 + 1;
    constexpr int ITERS = (TBR * KP2 + THREADS - 1) / THREADS;
    float4 buf[ITERS][RPL];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int q = threadIdx.x + it * THREADS;
        const int c = q / KP2, j = q - c * KP2;
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            const int row = row0 + RPL * c + h;
            buf[it][h] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (c < TBR && row < a.rows && !(a.pack && j == N / 4))
                buf[it][h] = reinterpret_cast<const float4*>(a.in + (size_t)row * a.P)[j];
        }
    }
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int q = threadIdx.x + it * THREADS;
        const int c = q / KP2, j = q - c * KP2;
        if (c < TBR) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int k = 2 * j + e;
                if (k > N / 2 || (a.pack && k == N / 2)) continue;
                float2 A = e ? make_float2(buf[it][0].z, buf[it][0].w) : make_float2(buf[it][0].x, buf[it][0].y);
                if constexpr (RPL == 1) {
                    if (a.pack && k == 0) {  // bins 0 and N/2, both real
                        s_line[c * LP + PADI(N / 2)] = make_float2(A.y, 0.f);
                        A.y = 0.f;
                    }
                    s_line[c * LP + PADI(k)] = A;
                    if (k > 0 && k < N / 2) s_line[c * LP + PADI(N - k)] = make_float2(A.x, -A.y);
                } else {
                    float2 B = e ? make_float2(buf[it][1].z, buf[it][1].w) : make_float2(buf[it][1].x, buf[it][1].y);
                    if (a.pack && k == 0) s_line[c * LP + PADI(N / 2)] = make_float2(A.y, B.y);
                    if (k == 0 || k == N / 2) A.y = B.y = 0.f;  // the real bins of a real row
                    s_line[c * LP + PADI(k)] = make_float2(A.x - B.y, A.y + B.x);
                    if (k > 0 && k < N / 2) s_line[c * LP + PADI(N - k)] = make_float2(A.x + B.y, B.x - A.y);
                }
            }
        }
    }
    __syncthreads();
    const int c = threadIdx.x / Plan<N>::T, t = threadIdx.x % Plan<N>::T;
    run_line<N, true>(s_line + c * LP, s_tw, t, false, SD_FFT_PRUNE_ROWS && 2 * a.nout <= N);  // only x < nout kept
    const int q4 = a.nout / 4;
    for (int i = threadIdx.x; i < TB * q4; i += THREADS) {
        const int r = i / q4, j = i - r * q4;
        const int row = row0 + r;
        if (row < a.rows) {
            const float* l = reinterpret_cast<const float*>(s_line + (r / RPL) * LP) + (r % RPL);
            reinterpret_cast<float4*>(a.out + (size_t)row * a.nout)[j] =
                make_float4(l[2 * PADI(4 * j)], l[2 * PADI(4 * j + 1)], l[2 * PADI(4 * j + 2)], l[2 * PADI(4 * j + 3)]);
        }
    }
}

template <int N>
__global__ void __launch_bounds__(XInv::LINES * Plan<N>::T) k_rows_inv(RowsInvArgs a, const float2* __restrict__ tw) {
    rows_inv_body<N>(a, tw, blockIdx.x);
}

template <int N>
__global__ void __launch_bounds__(XInv::LINES * Plan<N>::T) k_rows_inv_batch(const RowsInvArgs* __restrict__ jobs,
                                                                    const float2* __restrict__ tw) {
    const RowsInvArgs a = jobs[blockIdx.y];
    rows_inv_body<N>(a, tw, blockIdx.x);
}

// ---- column passes ---------------------------------------------------------
// Column b of slab a: element n at in[a*inA + n*inN + b]; output k at
// out[a*outA + k*outN + b]. A block owns TB adjacent columns of one slab; each
// thread moves column pairs (16 bytes) of TB/2-wide row segments.
enum ColMode { COL_FWD = 0, COL_INV = 1, COL_Z_ISO = 2, COL_Z_PER = 3, COL_GREEN = 4 };

struct ColArgs {
    const float2* in;
    float2* out;
    long long inA, inN, outA, outN;
    int nb;     // columns per slab
    int nin;    // rows loaded (zero beyond)
    int nout;   // rows stored
    // Z passes
    const float* green;  // octant table [kz'][ky'][kx], pitch P, (S/2+1)^2 rows
    float* greenOut;     // COL_GREEN target
    int P;
    int S;               // transform length
    float perScale;      // periodic: -(1/G^3)/(4 pi^2)
    float perSoft;       // periodic: (2 pi soft / G)^2 (0 disables)
    float greenNorm;     // COL_GREEN: 1/S^3
    FinalizeJob fin;     // fused z pass: reduction carried by block (0, 0) (fin.partials == nullptr: none)
    int pack;            // packed layout (nb = S/2; column 0 carries kx = 0 and kx = S/2)
};

// The z pass multiplier at (kx, ky, kz), kx <= N/2: the isolated Green
// octant (folded ky, kz) or kPoissonPeriodic (Sim.cu:528-557).
template <int MODE>
__device__ __forceinline__ float z_mult(const ColArgs& a, int N, int kx, int ky, int kz) {
    const int h = N / 2;
    if constexpr (MODE == COL_Z_ISO) {
        const int kyf = ky <= h ? ky : N - ky, kzf = kz <= h ? kz : N - kz;
        return a.green[((size_t)kzf * (h + 1) + kyf) * a.P + kx];
    } else {
        const int sky = ky <= h ? ky : ky - N, skz = kz <= h ? kz : kz - N;
        const float k2 = (float)(kx * kx + sky * sky + skz * skz);
        if (k2 < 1e-6f) return 0.f;
        float m = a.perScale / k2;
        if (a.perSoft > 0.f) m *= expf(-0.5f * k2 * a.perSoft);
        return m;
    }
}

// One line transform; halfIn: the upper half of the line is zero, halfOut:
// only its lower half is read afterwards (pruned stages where the plan has them).
template <int N, int MODE>
__device__ __forceinline__ void cols_body(const ColArgs& a, const float2* __restrict__ tw, int bx, int slab) {
    constexpr int LP = Line<N>::LP;
    constexpr int THREADS = TB * Plan<N>::T;
    constexpr int PAIRS = TB / 2;
    constexpr int RPI = THREADS / PAIRS;  // rows per staging iteration
    constexpr int ITERS = (N + RPI - 1) / RPI;
    __shared__ float2 s_line[TB * LP];
    __shared__ float2 s_tw[N];
    if (MODE == COL_GREEN && slab > N / 2) return;  // only ky <= S/2 is stored
    load_twiddles<N>(s_tw, tw);
    const int b0 = bx * TB;
    const int pc = threadIdx.x % PAIRS, r0 = threadIdx.x / PAIRS;
    const int col = 2 * pc;
    const bool ok0 = b0 + col < a.nb, ok1 = b0 + col + 1 < a.nb;
    {
        const float2* in = a.in + (size_t)slab * a.inA + b0 + col;
        float4 buf[ITERS];
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int n = r0 + it * RPI;
            buf[it] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (n < a.nin && ok0) {
                if (ok1) {
                    buf[it] = *reinterpret_cast<const float4*>(in + (size_t)n * a.inN);
                } else {
                    const float2 x = in[(size_t)n * a.inN];
                    buf[it] = make_float4(x.x, x.y, 0.f, 0.f);
                }
            }
        }
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int n = r0 + it * RPI;
            if (n < N) {
                s_line[col * LP + PADI(n)] = make_float2(buf[it].x, buf[it].y);
                s_line[(col + 1) * LP + PADI(n)] = make_float2(buf[it].z, buf[it].w);
            }
        }
    }
    __syncthreads();
    const int c = threadIdx.x / Plan<N>::T, t = threadIdx.x % Plan<N>::T;
    float2* line = s_line + c * LP;
    // packed column 0 (tile 0): the inverse y line is U = W0 + i WN, where slot
    // k <= N/2 holds W0(k) and slot N-k holds WN(k) (0 < k < N/2), and
    // W(N-k) = conj W(k) for both real-signal spectra; slots 0 and N/2 already
    // hold W0 + i WN
    if (MODE == COL_INV && a.pack && bx == 0) {
        for (int k = 1 + threadIdx.x; k < N / 2; k += THREADS) {
            const float2 w0 = s_line[PADI(k)], wn = s_line[PADI(N - k)];
            s_line[PADI(k)] = make_float2(w0.x - wn.y, w0.y + wn.x);
            s_line[PADI(N - k)] = make_float2(w0.x + wn.y, wn.x - w0.y);
        }
        __syncthreads();
    }
    // zero-padded lines (isolated meshes: G of S = 2G rows loaded) and
    // half-stored results take the pruned first / last stage
    if (MODE == COL_INV) {
        run_line<N, true>(line, s_tw, t, false, 2 * a.nout <= N);
    } else {
        run_line<N, false>(line, s_tw, t, 2 * a.nin <= N, false);
    }
    // packed column 0 (tile 0) after the forward y transform: Z = F0 + i FN of
    // the two real columns; F0(k) = (Z(k) + conj Z(N-k)) / 2 stays in slot k,
    // FN(k) = (Z(k) - conj Z(N-k)) / 2i goes to slot N-k (0 < k < N/2)
    if (MODE == COL_FWD && a.pack && bx == 0) {
        for (int k = 1 + threadIdx.x; k < N / 2; k += THREADS) {
            const float2 z = s_line[PADI(k)], m = s_line[PADI(N - k)];
            s_line[PADI(k)] = make_float2(0.5f * (z.x + m.x), 0.5f * (z.y - m.y));
            s_line[PADI(N - k)] = make_float2(0.5f * (z.y + m.y), 0.5f * (m.x - z.x));
        }
        __syncthreads();
    }
    if (MODE == COL_GREEN) {
        const int h = N / 2;
        for (int i = threadIdx.x; i < TB * (h + 1); i += THREADS) {
            const int cc = i % TB, kz = i / TB;
            if (b0 + cc < a.nb)
                a.greenOut[((size_t)kz * (h + 1) + slab) * a.P + b0 + cc] =
                    s_line[cc * LP + PADI(kz)].x * a.greenNorm;
        }
        return;
    }
    if (MODE == COL_Z_ISO || MODE == COL_Z_PER) {
        const int h = N / 2;
        const int ky = slab;
        const int kyf = ky <= h ? ky : N - ky;
        const int kx0 = b0 + col;
        // packed column 0: slabs 0 < ky < N/2 hold kx = 0, slabs ky > N/2 hold
        // kx = N/2 at ky' = N - ky (the same folded ky and |sky|), slabs 0 and
        // N/2 hold both (done below)
        const bool col0 = a.pack && kx0 == 0;
        const bool both = col0 && (ky == 0 || ky == h);
        const int kxa = col0 && ky > h ? h : kx0;
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int kz = r0 + it * RPI;
            if (kz >= N) continue;
            float m0, m1;
            if (MODE == COL_Z_ISO) {
                const int kzf = kz <= h ? kz : N - kz;
                // the table row has pitch P > nb, so the pair never leaves it
                const float* grow = a.green + ((size_t)kzf * (h + 1) + kyf) * a.P;
                const float2 g = *reinterpret_cast<const float2*>(grow + kx0);
                m0 = g.x;
                m1 = g.y;
                if (kxa != kx0) m0 = grow[kxa];
            } else {
                // kPoissonPeriodic, Sim.cu:528-557
                const int sky = ky <= h ? ky : ky - N;
                const int skz = kz <= h ? kz : kz - N;
                float mm[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int kx = e ? kx0 + 1 : kxa;
                    const float k2 = (float)(kx * kx + sky * sky + skz * skz);
                    if (k2 < 1e-6f) {
                        mm[e] = 0.f;
                    } else {
                        mm[e] = a.perScale / k2;
                        if (a.perSoft > 0.f) mm[e] *= expf(-0.5f * k2 * a.perSoft);
                    }
                }
                m0 = mm[0];
                m1 = mm[1];
            }
            const float2 u = s_line[col * LP + PADI(kz)], v = s_line[(col + 1) * LP + PADI(kz)];
            if (!both) s_line[col * LP + PADI(kz)] = make_float2(u.x * m0, u.y * m0);
            s_line[(col + 1) * LP + PADI(kz)] = make_float2(v.x * m1, v.y * m1);
        }
        // ky = 0 and N/2 of packed column 0: the line is Z = a + i b of two real
        // z-sequences (kx = 0 and kx = N/2), whose spectra take multipliers m0
        // and mN; with both even in kz, W(kz) = Z(kz) (m0 + mN) / 2 +
        // conj Z(N - kz) (m0 - mN) / 2 (real and imaginary parts apart at kz = 0, N/2)
        if (a.pack && bx == 0 && (ky == 0 || ky == h)) {
            for (int kz = threadIdx.x; kz <= h; kz += THREADS) {
                const float mz0 = z_mult<MODE>(a, N, 0, ky, kz), mzn = z_mult<MODE>(a, N, h, ky, kz);
                const int km = kz == 0 ? 0 : N - kz;
                const float2 z1 = s_line[PADI(kz)], z2 = s_line[PADI(km)];
                if (kz == 0 || kz == h) {
                    s_line[PADI(kz)] = make_float2(z1.x * mz0, z1.y * mzn);
                } else {
                    const float sp = 0.5f * (mz0 + mzn), sm = 0.5f * (mz0 - mzn);
                    s_line[PADI(kz)] = make_float2(z1.x * sp + z2.x * sm, z1.y * sp - z2.y * sm);
                    s_line[PADI(km)] = make_float2(z2.x * sp + z1.x * sm, z2.y * sp - z1.y * sm);
                }
            }
        }
        __syncthreads();
        run_line<N, true>(line, s_tw, t, false, 2 * a.nout <= N);
    }
    float2* out = a.out + (size_t)slab * a.outA + b0 + col;
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int k = r0 + it * RPI;
        if (k < a.nout && ok0) {
            const float2 u = s_line[col * LP + PADI(k)], v = s_line[(col + 1) * LP + PADI(k)];
            if (ok1)
                *reinterpret_cast<float4*>(out + (size_t)k * a.outN) = make_float4(u.x, u.y, v.x, v.y);
            else
                out[(size_t)k * a.outN] = u;
        }
    }
}

template <int N>
constexpr bool cols_carry_finalize() {
    constexpr int TH = TB * Plan<N>::T;
    return FIN % TH == 0 && TH % 64 == 0 && TH >= 256;
}

template <int N, int MODE>
__global__ void __launch_bounds__(TB * Plan<N>::T) k_cols(ColArgs a, const float2* __restrict__ tw) {
    if constexpr ((MODE == COL_Z_ISO || MODE == COL_Z_PER) && cols_carry_finalize<N>()) {
        if (a.fin.partials) {  // one more row of blocks in front: block (0, 0) reduces, the rest exit
            if (blockIdx.y == 0) {
                if (blockIdx.x == 0)
                    finalize_body<TB * Plan<N>::T>(a.fin.partials, a.fin.nparts, a.fin.mode, a.fin.pol, a.fin.sc);
                return;
            }
            cols_body<N, MODE>(a, tw, blockIdx.x, blockIdx.y - 1);
            return;
        }
    }
    cols_body<N, MODE>(a, tw, blockIdx.x, blockIdx.y);
}

// Persistent form of the forward / inverse / fused-z column passes (same
// arithmetic as cols_body, tile by tile): a block walks tiles
// (bx, slab) = tile % colBlocks, tile / colBlocks with stride gridDim.x, and
// issues the next tile's global loads right after staging the current one, so
// they are in flight during the transforms and stores of the current tile.
// The fused-z multipliers of a tile are read before that prefetch (vector
// memory counters retire in order).
template <int N, int MODE>
__global__ void __launch_bounds__(TB * Plan<N>::T) k_cols_stream(ColArgs a, const float2* __restrict__ tw,
                                                                 int colBlocks, int ntiles) {
    static_assert(MODE == COL_FWD || MODE == COL_INV || MODE == COL_Z_ISO || MODE == COL_Z_PER, "mode");
    constexpr int LP = Line<N>::LP;
    constexpr int THREADS = TB * Plan<N>::T;
    constexpr int PAIRS = TB / 2;
    constexpr int RPI = THREADS / PAIRS;
    constexpr int ITERS = (N + RPI - 1) / RPI;
    __shared__ float2 s_line[TB * LP];
    __shared__ float2 s_tw[N];
    load_twiddles<N>(s_tw, tw);
    const int pc = threadIdx.x % PAIRS, r0 = threadIdx.x / PAIRS;
    const int col = 2 * pc;
    const int c = threadIdx.x / Plan<N>::T, t = threadIdx.x % Plan<N>::T;
    float2* line = s_line + c * LP;
    float4 buf[ITERS];
    auto fetch = [&](int tile) {
        const int bx = tile % colBlocks, slab = tile / colBlocks;
        const int b0 = bx * TB;
        const bool ok0 = b0 + col < a.nb, ok1 = b0 + col + 1 < a.nb;
        const float2* in = a.in + (size_t)slab * a.inA + b0 + col;
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int n = r0 + it * RPI;
            buf[it] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (n < a.nin && ok0) {
                if (ok1) {
                    buf[it] = *reinterpret_cast<const float4*>(in + (size_t)n * a.inN);
                } else {
                    const float2 x = in[(size_t)n * a.inN];
                    buf[it] = make_float4(x.x, x.y, 0.f, 0.f);
                }
            }
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int bx = tile % colBlocks, slab = tile / colBlocks;
        const int b0 = bx * TB;
        const bool ok0 = b0 + col < a.nb, ok1 = b0 + col + 1 < a.nb;
        __syncthreads();  // the previous tile's lines have been read out
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int n = r0 + it * RPI;
            if (n < N) {
                s_line[col * LP + PADI(n)] = make_float2(buf[it].x, buf[it].y);
                s_line[(col + 1) * LP + PADI(n)] = make_float2(buf[it].z, buf[it].w);
            }
        }
        __syncthreads();
        // fused-z multipliers of this tile (before the prefetch)
        float m0[ITERS], m1[ITERS];
        if (MODE == COL_Z_ISO || MODE == COL_Z_PER) {
            const int h = N / 2;
            const int ky = slab;
            const int kyf = ky <= h ? ky : N - ky;
            const int kx0 = b0 + col;
#pragma unroll
            for (int it = 0; it < ITERS; ++it) {
                const int kz = r0 + it * RPI;
                m0[it] = m1[it] = 0.f;
                if (kz >= N) continue;
                if (MODE == COL_Z_ISO) {
                    const int kzf = kz <= h ? kz : N - kz;
                    const float2 g = *reinterpret_cast<const float2*>(a.green + ((size_t)kzf * (h + 1) + kyf) * a.P + kx0);
                    m0[it] = g.x;
                    m1[it] = g.y;
                } else {
                    // kPoissonPeriodic, Sim.cu:528-557
                    const int sky = ky <= h ? ky : ky - N;
                    const int skz = kz <= h ? kz : kz - N;
                    float mm[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int kx = kx0 + e;
                        const float k2 = (float)(kx * kx + sky * sky + skz * skz);
                        if (k2 < 1e-6f) {
                            mm[e] = 0.f;
                        } else {
                            mm[e] = a.perScale / k2;
                            if (a.perSoft > 0.f) mm[e] *= expf(-0.5f * k2 * a.perSoft);
                        }
                    }
                    m0[it] = mm[0];
                    m1[it] = mm[1];
                }
            }
        }
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        if (MODE == COL_INV) {
            Plan<N>::template run<true>(line, s_tw, t);
        } else {
            Plan<N>::template run<false>(line, s_tw, t);
        }
        if (MODE == COL_Z_ISO || MODE == COL_Z_PER) {
#pragma unroll
            for (int it = 0; it < ITERS; ++it) {
                const int kz = r0 + it * RPI;
                if (kz >= N) continue;
                const float2 u = s_line[col * LP + PADI(kz)], v = s_line[(col + 1) * LP + PADI(kz)];
                s_line[col * LP + PADI(kz)] = make_float2(u.x * m0[it], u.y * m0[it]);
                s_line[(col + 1) * LP + PADI(kz)] = make_float2(v.x * m1[it], v.y * m1[it]);
            }
            __syncthreads();
            Plan<N>::template run<true>(line, s_tw, t);
        }
        float2* out = a.out + (size_t)slab * a.outA + b0 + col;
#pragma unroll
        for (int it = 0; it < ITERS; ++it) {
            const int k = r0 + it * RPI;
            if (k < a.nout && ok0) {
                const float2 u = s_line[col * LP + PADI(k)], v = s_line[(col + 1) * LP + PADI(k)];
                if (ok1)
                    *reinterpret_cast<float4*>(out + (size_t)k * a.outN) = make_float4(u.x, u.y, v.x, v.y);
                else
                    out[(size_t)k * a.outN] = u;
            }
        }
    }
}

// Column pass of many meshes of the same FFT side and mesh size: blockIdx.z
// selects the mesh.
template <int N, int MODE>
__global__ void __launch_bounds__(TB * Plan<N>::T) k_cols_batch(const ColArgs* __restrict__ jobs,
                                                                const float2* __restrict__ tw) {
    const ColArgs a = jobs[blockIdx.z];
    cols_body<N, MODE>(a, tw, blockIdx.x, blockIdx.y);
}

// ---- register-resident column passes at N = 256 --------------------------------
// 256 = 16 x 16 four-step transform (Bailey, "FFTs in external or hierarchical
// memory", J. Supercomputing 4, 1990) held in registers. Thread (c, t) of a
// 256-thread block owns column c = threadIdx % 16 of the tile and elements
// n = t + 16 r (t = threadIdx / 16, r = 0 .. 15) of that column's line, loaded
// straight from HBM (the 16 lanes of one row read one 128-B line). It
// transforms over r in registers, multiplies by W^(t k2), and exchanges
// through LDS once (element (t, k2) to thread k2), then transforms over t:
//   X[k2 + 16 k1] = sum_t W16^(t k1) W^(t k2) sum_r W16^(r k2) x[t + 16 r].
// Thread t then holds outputs k = t + 16 k1, the distribution its inputs had,
// so the fused z pass multiplies and transforms back in place and every result
// is stored straight from registers: a line crosses LDS once per transform
// instead of once per radix stage plus the staging in and out.
// HIN: inputs n >= 128 are zero padding (the first transform over r skips
// them); HOUT: only outputs < 128 are stored (the last transform over t of the
// inverse computes only those).
// Element access by a 32-bit element offset from a uniform base (the load or
// store adds the offset to the scalar base itself).
template <typename T>
__device__ __forceinline__ T ld_c(const T* base, unsigned i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_c(T* base, unsigned i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T)) = v;
}
constexpr int RLP = 257;  // LDS pitch of one line (float2): odd, so the 16 lines of a row fall in distinct banks

template <bool INV, bool ZIN>
__device__ __forceinline__ void fft256_regs(float2 v[16], float2* __restrict__ s_x, const float2* __restrict__ s_tw,
                                            int c, int t, bool syncFirst) {
    Dft<16, INV, ZIN>::run(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        float2 w = s_tw[t * k];
        if (INV) w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
    if (syncFirst) __syncthreads();  // earlier readers of s_x are done
    float2* line = s_x + c * RLP;
#pragma unroll
    for (int k = 0; k < 16; ++k) line[k * 16 + t] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = line[t * 16 + k];
    Dft<16, INV, false>::run(v);
}

template <int MODE, bool HIN, bool HOUT>
__device__ __forceinline__ void cols_reg_body(const ColArgs& a, const float2* __restrict__ tw, int bx, int slab) {
    static_assert(MODE == COL_FWD || MODE == COL_INV || MODE == COL_Z_ISO || MODE == COL_Z_PER, "mode");
    constexpr int N = 256, h = N / 2;
    constexpr bool ZP = MODE == COL_Z_ISO || MODE == COL_Z_PER;
    __shared__ float2 s_x[16 * RLP];
    __shared__ float2 s_tw[N];
    s_tw[threadIdx.x] = tw[threadIdx.x];
    const int c = threadIdx.x & 15, t = threadIdx.x >> 4;
    const int kx0 = bx * 16 + c;
    const bool okc = kx0 < a.nb;
    // element offsets (< 2^24 at N = 256: 24-bit multiplies) from the uniform
    // base; a partial tile's dead columns load a live one and are dropped
    const int kxl = okc ? kx0 : a.nb - 1;
    float2 v[16];
    {
        const unsigned off = __umul24((unsigned)slab, (unsigned)a.inA) + (unsigned)kxl +
                             __umul24((unsigned)t, (unsigned)a.inN),
                       step = (unsigned)a.inN << 4;
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = r < (HIN ? 8 : 16) ? ld_c(a.in, off + r * step) : make_float2(0.f, 0.f);
        if (!okc) {
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = make_float2(0.f, 0.f);
        }
    }
    // fused z pass: the multiplier of output kz = t + 16 k1 (isolated: the
    // Green octant at folded ky, kz; periodic: kPoissonPeriodic, Sim.cu:528-557).
    // Packed column 0: slabs ky > N/2 hold kx = N/2.
    float m[ZP ? 16 : 1];
    const bool col0 = a.pack && kx0 == 0;
    if constexpr (ZP) {
        const int ky = slab, kyf = ky <= h ? ky : N - ky;
        const int kxa = col0 && ky > h ? h : kxl;
        const unsigned P = (unsigned)a.P, gbase = __umul24((unsigned)kyf, P) + (unsigned)kxa,
                       gplane = __umul24((unsigned)(h + 1), P);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const int kz = t + 16 * k1;
            if (MODE == COL_Z_ISO) {
                const int kzf = kz <= h ? kz : N - kz;
                m[k1] = ld_c(a.green, gbase + __umul24((unsigned)kzf, gplane));
            } else {
                m[k1] = z_mult<COL_Z_PER>(a, N, kxa, ky, kz);
            }
        }
    }
    __syncthreads();  // s_tw
    const bool packTile = a.pack && bx == 0;  // block-uniform
    // packed column 0 of the inverse y pass: rebuild U = W0 + i WN (slot k < N/2
    // holds W0(k), slot N-k holds WN(k); W(N-k) = conj W(k) for both)
    if (MODE == COL_INV && packTile) {
        if (c == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) s_x[t + 16 * r] = v[r];
        }
        __syncthreads();
        if (c == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int k = t + 16 * r;
                if (k == 0 || k == h) continue;
                const float2 o = v[r], p = s_x[N - k];
                v[r] = k < h ? make_float2(o.x - p.y, o.y + p.x) : make_float2(p.x + o.y, o.x - p.y);
            }
        }
    }
    fft256_regs<MODE == COL_INV, HIN && MODE != COL_INV>(v, s_x, s_tw, c, t, MODE == COL_INV && packTile);
    // packed column 0 after the forward y transform: Z = F0 + i FN of two real
    // columns; F0(k) = (Z(k) + conj Z(N-k)) / 2 stays in slot k, FN(k) =
    // (Z(k) - conj Z(N-k)) / 2i goes to slot N-k (0 < k < N/2)
    if (MODE == COL_FWD && packTile) {
        __syncthreads();
        if (c == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) s_x[t + 16 * r] = v[r];
        }
        __syncthreads();
        if (c == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int k = t + 16 * r;
                if (k == 0 || k == h) continue;
                const float2 o = v[r], p = s_x[N - k];
                v[r] = k < h ? make_float2(0.5f * (o.x + p.x), 0.5f * (o.y - p.y))
                             : make_float2(0.5f * (p.y + o.y), 0.5f * (o.x - p.x));
            }
        }
    }
    if constexpr (ZP) {
        // ky = 0 and N/2 of packed column 0: the line is Z = a + i b of two real
        // z-sequences (kx = 0 and kx = N/2) with multipliers m0 and mN, both even
        // in kz: W(kz) = Z(kz) (m0 + mN) / 2 + conj Z(N - kz) (m0 - mN) / 2
        // (real and imaginary parts apart at kz = 0, N/2)
        const bool both = packTile && (slab == 0 || slab == h);  // block-uniform
        if (!(both && c == 0)) {
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) v[k1] = make_float2(v[k1].x * m[k1], v[k1].y * m[k1]);
        }
        if (both) {
            __syncthreads();
            if (c == 0) {
#pragma unroll
                for (int k1 = 0; k1 < 16; ++k1) s_x[t + 16 * k1] = v[k1];
            }
            __syncthreads();
            if (c == 0) {
#pragma unroll
                for (int k1 = 0; k1 < 16; ++k1) {
                    const int kz = t + 16 * k1;
                    const float mz0 = z_mult<MODE>(a, N, 0, slab, kz), mzn = z_mult<MODE>(a, N, h, slab, kz);
                    const float2 o = v[k1];
                    if (kz == 0 || kz == h) {
                        v[k1] = make_float2(o.x * mz0, o.y * mzn);
                    } else {
                        const float2 p = s_x[N - kz];
                        const float sp = 0.5f * (mz0 + mzn), sm = 0.5f * (mz0 - mzn);
                        v[k1] = make_float2(o.x * sp + p.x * sm, o.y * sp - p.y * sm);
                    }
                }
            }
        }
        fft256_regs<true, false>(v, s_x, s_tw, c, t, true);
    }
    if (okc) {
        const unsigned off = __umul24((unsigned)slab, (unsigned)a.outA) + (unsigned)kx0 +
                             __umul24((unsigned)t, (unsigned)a.outN),
                       step = (unsigned)a.outN << 4;
#pragma unroll
        for (int r = 0; r < (HOUT ? 8 : 16); ++r) st_c(a.out, off + r * step, v[r]);
    }
}

template <int MODE, bool HIN, bool HOUT>
__global__ void __launch_bounds__(256) k_cols_reg(ColArgs a, const float2* __restrict__ tw) {
    if constexpr (MODE == COL_Z_ISO || MODE == COL_Z_PER) {
        if (a.fin.partials) {  // one more row of blocks in front: block (0, 0) reduces, the rest exit
            if (blockIdx.y == 0) {
                if (blockIdx.x == 0) finalize_body<256>(a.fin.partials, a.fin.nparts, a.fin.mode, a.fin.pol, a.fin.sc);
                return;
            }
            cols_reg_body<MODE, HIN, HOUT>(a, tw, blockIdx.x, blockIdx.y - 1);
            return;
        }
    }
    cols_reg_body<MODE, HIN, HOUT>(a, tw, blockIdx.x, blockIdx.y);
}

template <int MODE, bool HIN, bool HOUT>
__global__ void __launch_bounds__(256) k_cols_reg_batch(const ColArgs* __restrict__ jobs,
                                                        const float2* __restrict__ tw) {
    const ColArgs a = jobs[blockIdx.z];
    cols_reg_body<MODE, HIN, HOUT>(a, tw, blockIdx.x, blockIdx.y);
}

// ---- register-resident fused z pass at N = 192 -----------------------------------
// The four-step form with 192 = 16 x 12. Thread t of a column's 16 holds
// n = t + 16 r (r < 12): the forward transform takes the 12-point transform
// over r, the twiddles W192^(t k2), one LDS exchange, and 16-point transforms
// over t; threads u < 12 then hold X[u + 12 k1] (k1 < 16), multiply, and the
// inverse runs the same steps backwards (16-point transforms over k1,
// twiddles W192^(-t u), one exchange, 12-point transforms over u), which
// leaves thread t with n = t + 16 r again, so loads and stores both go
// straight from registers. The middle steps run on threads u < 12: the
// fourth wavefront of the block (t = 12..15) waits at the barriers.
// Isolated meshes only (a periodic mesh has a power-of-two side): the
// zero-padded inputs (n >= 96, r >= 6) are not loaded, and only outputs
// n < 96 are stored.
constexpr int LP192 = 193;  // LDS pitch of one line (float2)

template <bool ZIN>
__device__ __forceinline__ void fft192_fwd_regs(float2 v[12], float2 y[16], float2* __restrict__ line,
                                                const float2* __restrict__ s_tw, int t) {
    Dft<12, false, ZIN>::run(v);
#pragma unroll
    for (int k = 1; k < 12; ++k) v[k] = cmul(v[k], s_tw[t * k]);
#pragma unroll
    for (int k = 0; k < 12; ++k) line[k * 16 + t] = v[k];
    __syncthreads();
    if (t < 12) {  // wavefront-uniform
#pragma unroll
        for (int j = 0; j < 16; ++j) y[j] = line[t * 16 + j];
        Dft<16, false, false>::run(y);
    }
}

// y: thread u < 12 holds X[u + 12 j]; returns x[t + 16 r] in v.
__device__ __forceinline__ void fft192_inv_regs(float2 y[16], float2 v[12], float2* __restrict__ line,
                                                const float2* __restrict__ s_tw, int t) {
    __syncthreads();  // earlier readers of the line are done
    if (t < 12) {
        Dft<16, true, false>::run(y);
#pragma unroll
        for (int k = 1; k < 16; ++k) {
            float2 w = s_tw[k * t];
            w.y = -w.y;
            y[k] = cmul(y[k], w);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) line[k * 12 + t] = y[k];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 12; ++u) v[u] = line[t * 12 + u];
    Dft<12, true, false>::run(v);
}

__device__ __forceinline__ void cols192_z_body(const ColArgs& a, const float2* __restrict__ tw, int bx, int slab) {
    constexpr int N = 192, h = N / 2;
    __shared__ float2 s_x[16 * LP192];
    __shared__ float2 s_tw[N];
    if (threadIdx.x < N) s_tw[threadIdx.x] = tw[threadIdx.x];
    const int c = threadIdx.x & 15, t = threadIdx.x >> 4;
    const int kx0 = bx * 16 + c;
    const bool okc = kx0 < a.nb;
    const int kxl = okc ? kx0 : a.nb - 1;  // a partial tile's dead columns load a live one
    float2 v[12];
    {
        const unsigned off = __umul24((unsigned)slab, (unsigned)a.inA) + (unsigned)kxl +
                             __umul24((unsigned)t, (unsigned)a.inN),
                       step = (unsigned)a.inN << 4;
#pragma unroll
        for (int r = 0; r < 12; ++r) v[r] = r < 6 ? ld_c(a.in, off + r * step) : make_float2(0.f, 0.f);
        if (!okc) {
#pragma unroll
            for (int r = 0; r < 12; ++r) v[r] = make_float2(0.f, 0.f);
        }
    }
    // the multiplier of output kz = t + 12 k1 (threads t < 12)
    float m[16];
    const bool col0 = a.pack && kx0 == 0;
    {
        const int ky = slab, kyf = ky <= h ? ky : N - ky;
        const int kxa = col0 && ky > h ? h : kxl;
        const unsigned P = (unsigned)a.P, gbase = __umul24((unsigned)kyf, P) + (unsigned)kxa,
                       gplane = __umul24((unsigned)(h + 1), P);
        const int tu = t < 12 ? t : 0;
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const int kz = tu + 12 * k1, kzf = kz <= h ? kz : N - kz;
            m[k1] = ld_c(a.green, gbase + __umul24((unsigned)kzf, gplane));
        }
    }
    __syncthreads();  // s_tw
    float2 y[16];
    float2* line = s_x + c * LP192;
    fft192_fwd_regs<true>(v, y, line, s_tw, t);
    // ky = 0 and N/2 of packed column 0: the line is Z = a + i b of two real
    // z-sequences (kx = 0 and kx = N/2) with multipliers m0 and mN, both even in
    // kz: W(kz) = Z(kz) (m0 + mN) / 2 + conj Z(N - kz) (m0 - mN) / 2 (real and
    // imaginary parts apart at kz = 0, N/2)
    const bool both = a.pack && bx == 0 && (slab == 0 || slab == h);  // block-uniform
    if (t < 12 && !(both && c == 0)) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) y[k1] = make_float2(y[k1].x * m[k1], y[k1].y * m[k1]);
    }
    if (both) {
        __syncthreads();
        if (c == 0 && t < 12) {
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) s_x[t + 12 * k1] = y[k1];
        }
        __syncthreads();
        if (c == 0 && t < 12) {
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) {
                const int kz = t + 12 * k1;
                const float mz0 = z_mult<COL_Z_ISO>(a, N, 0, slab, kz), mzn = z_mult<COL_Z_ISO>(a, N, h, slab, kz);
                const float2 o = y[k1];
                if (kz == 0 || kz == h) {
                    y[k1] = make_float2(o.x * mz0, o.y * mzn);
                } else {
                    const float2 p = s_x[N - kz];
                    const float sp = 0.5f * (mz0 + mzn), sm = 0.5f * (mz0 - mzn);
                    y[k1] = make_float2(o.x * sp + p.x * sm, o.y * sp - p.y * sm);
                }
            }
        }
    }
    fft192_inv_regs(y, v, line, s_tw, t);
    if (okc) {
        const unsigned off = __umul24((unsigned)slab, (unsigned)a.outA) + (unsigned)kx0 +
                             __umul24((unsigned)t, (unsigned)a.outN),
                       step = (unsigned)a.outN << 4;
#pragma unroll
        for (int r = 0; r < 6; ++r) st_c(a.out, off + r * step, v[r]);
    }
}

__global__ void __launch_bounds__(256) k_cols192_z(ColArgs a, const float2* __restrict__ tw) {
    if (a.fin.partials) {  // one more row of blocks in front: block (0, 0) reduces, the rest exit
        if (blockIdx.y == 0) {
            if (blockIdx.x == 0) finalize_body<256>(a.fin.partials, a.fin.nparts, a.fin.mode, a.fin.pol, a.fin.sc);
            return;
        }
        cols192_z_body(a, tw, blockIdx.x, blockIdx.y - 1);
        return;
    }
    cols192_z_body(a, tw, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(256) k_cols192_z_batch(const ColArgs* __restrict__ jobs,
                                                         const float2* __restrict__ tw) {
    const ColArgs a = jobs[blockIdx.z];
    cols192_z_body(a, tw, blockIdx.x, blockIdx.y);
}

// ---- register-resident column passes at N = 512 -----------------------------------
// 512 = 16 x 32. Thread t of a column's 16 holds n = t + 16 r (r < 32): the
// forward transform takes the 32-point transform over r, the twiddles
// W512^(t k2), one LDS exchange, and the 16-point transforms over t of
// k2 = u and u + 16; thread u then holds X[u + 32 k1] and X[u + 16 + 32 k1]
// (k1 < 16), multiplies, and the inverse runs the same steps backwards
// (16-point transforms, conjugate twiddles, one exchange, 32-point
// transforms), which leaves thread t with n = t + 16 r again: loads and stores
// go straight from registers. HIO: the isolated mesh's zero-padded inputs
// (n >= 256) are not loaded, and only outputs n < 256 are stored. The S = 512
// solve has no packed column (nb = 257, the last tile holds kx = 256 alone).
#ifndef SD_Z512_LINES
#define SD_Z512_LINES 16
#endif
constexpr int Z512_LINES = SD_Z512_LINES;  // columns per block
constexpr int LP512 = 513;                 // LDS pitch of one line (float2)

template <bool ZIN>
__device__ __forceinline__ void fft512_fwd_regs(float2 v[32], float2 y0[16], float2 y1[16],
                                                float2* __restrict__ line, const float2* __restrict__ s_tw, int t) {
    Dft<32, false, ZIN>::run(v);
#pragma unroll
    for (int k = 1; k < 32; ++k) v[k] = cmul(v[k], s_tw[t * k]);
#pragma unroll
    for (int k = 0; k < 32; ++k) line[k * 16 + t] = v[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        y0[j] = line[t * 16 + j];
        y1[j] = line[(t + 16) * 16 + j];
    }
    Dft<16, false, false>::run(y0);
    Dft<16, false, false>::run(y1);
}

__device__ __forceinline__ void fft512_inv_regs(float2 y0[16], float2 y1[16], float2 v[32],
                                                float2* __restrict__ line, const float2* __restrict__ s_tw, int t) {
    Dft<16, true, false>::run(y0);
    Dft<16, true, false>::run(y1);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        float2 w0 = s_tw[k * t], w1 = s_tw[k * (t + 16)];
        w0.y = -w0.y;
        w1.y = -w1.y;
        y0[k] = cmul(y0[k], w0);
        y1[k] = cmul(y1[k], w1);
    }
    __syncthreads();  // the forward exchange's readers are done
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        line[k * 32 + t] = y0[k];
        line[k * 32 + t + 16] = y1[k];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 32; ++u) v[u] = line[t * 32 + u];
    Dft<32, true, false>::run(v);
}

// MODE: the fused z pass, or the y forward / inverse pass alone (the forward
// stores from the middle distribution, the inverse loads into it). HIO: the
// pruned side of the pass (forward input, inverse output, or both for z).
template <int MODE, bool HIO>
__device__ __forceinline__ void cols512_body(const ColArgs& a, const float2* __restrict__ tw, int bx, int slab) {
    constexpr int N = 512, h = N / 2;
    constexpr bool ZP = MODE == COL_Z_ISO || MODE == COL_Z_PER;
    __shared__ float2 s_x[Z512_LINES * LP512];
    __shared__ float2 s_tw[N];
    for (int i = threadIdx.x; i < N; i += Z512_LINES * 16) s_tw[i] = tw[i];
    const int c = threadIdx.x % Z512_LINES, t = threadIdx.x / Z512_LINES;
    const int kx0 = bx * Z512_LINES + c;
    const bool okc = kx0 < a.nb;
    const int kxl = okc ? kx0 : a.nb - 1;  // a partial tile's dead columns load a live one
    const unsigned inOff = (unsigned)slab * (unsigned)a.inA + (unsigned)kxl, inN = (unsigned)a.inN;
    float2 v[32];
    float2 y0[16], y1[16];
    if constexpr (MODE != COL_INV) {
        const unsigned off = inOff + (unsigned)t * inN, step = inN << 4;
#pragma unroll
        for (int r = 0; r < 32; ++r) v[r] = r < (HIO ? 16 : 32) ? ld_c(a.in, off + r * step) : make_float2(0.f, 0.f);
    } else {
        const unsigned off = inOff + (unsigned)t * inN, step = inN << 5;
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            y0[k1] = ld_c(a.in, off + k1 * step);
            y1[k1] = ld_c(a.in, off + 16u * inN + k1 * step);
        }
    }
    if (!okc) {
#pragma unroll
        for (int r = 0; r < 32; ++r) v[r] = make_float2(0.f, 0.f);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) y0[k1] = y1[k1] = make_float2(0.f, 0.f);
    }
    // the z pass's multipliers of outputs kz = t + 32 k1 and t + 16 + 32 k1
    float m0[16], m1[16];
    const int ky = slab;
    if constexpr (MODE == COL_Z_ISO) {
        const int kyf = ky <= h ? ky : N - ky;
        const unsigned P = (unsigned)a.P, gbase = (unsigned)kyf * P + (unsigned)kxl, gplane = (unsigned)(h + 1) * P;
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            const int kz0 = t + 32 * k1, kz1 = kz0 + 16;
            m0[k1] = ld_c(a.green, gbase + (unsigned)(kz0 <= h ? kz0 : N - kz0) * gplane);
            m1[k1] = ld_c(a.green, gbase + (unsigned)(kz1 <= h ? kz1 : N - kz1) * gplane);
        }
    }
    __syncthreads();  // s_tw
    float2* line = s_x + c * LP512;
    if constexpr (MODE != COL_INV) fft512_fwd_regs<HIO && MODE != COL_INV>(v, y0, y1, line, s_tw, t);
    if constexpr (MODE == COL_FWD) {
        if (okc) {
            const unsigned off = (unsigned)slab * (unsigned)a.outA + (unsigned)kx0 + (unsigned)t * (unsigned)a.outN,
                           step = (unsigned)a.outN << 5;
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) {
                st_c(a.out, off + k1 * step, y0[k1]);
                st_c(a.out, off + 16u * (unsigned)a.outN + k1 * step, y1[k1]);
            }
        }
        return;
    }
    if constexpr (MODE == COL_Z_PER) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            m0[k1] = z_mult<COL_Z_PER>(a, N, kxl, ky, t + 32 * k1);
            m1[k1] = z_mult<COL_Z_PER>(a, N, kxl, ky, t + 16 + 32 * k1);
        }
    }
    if constexpr (ZP) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) {
            y0[k1] = make_float2(y0[k1].x * m0[k1], y0[k1].y * m0[k1]);
            y1[k1] = make_float2(y1[k1].x * m1[k1], y1[k1].y * m1[k1]);
        }
    }
    fft512_inv_regs(y0, y1, v, line, s_tw, t);
    if (okc) {
        const unsigned off = (unsigned)slab * (unsigned)a.outA + (unsigned)kx0 + (unsigned)t * (unsigned)a.outN,
                       step = (unsigned)a.outN << 4;
#pragma unroll
        for (int r = 0; r < (HIO ? 16 : 32); ++r) st_c(a.out, off + r * step, v[r]);
    }
}

template <int MODE, bool HIO>
__global__ void __launch_bounds__(Z512_LINES * 16) k_cols512_reg(ColArgs a, const float2* __restrict__ tw) {
    cols512_body<MODE, HIO>(a, tw, blockIdx.x, blockIdx.y);
}

// ---- register-resident row passes at N = 256 -----------------------------------
// The same four-step transform with lines along x: the 16 threads of a line are
// 16 consecutive lanes (t = lane % 16), which read one row's consecutive values
// and exchange through LDS inside their wavefront, so every exchange is ordered
// by wavefront-scope fences (no block barrier). Element (k2, t) of a line sits
// at k2 * 17 + t: the writes of one k2 and the reads of one t both fall in 16
// distinct banks. Two real rows share a line (Z = a + i b): the forward pass
// splits the two half spectra by Hermitian symmetry (one exchange of Z(N - k)),
// the inverse pass builds Z(k) = A(k) + i B(k) from the stored half spectra
// (Z(N - k) = conj A(k) + i conj B(k)) and keeps the real part for one row and
// the imaginary part for the other.
constexpr int XRLP = 16 * 17;  // LDS floats2 per line
__device__ __forceinline__ int xr_pos(int k) { return (k >> 4) * 17 + (k & 15); }

#ifndef SD_XREG_LINES_FWD
#define SD_XREG_LINES_FWD 16
#endif
#ifndef SD_XREG_RPL_FWD
#define SD_XREG_RPL_FWD 1
#endif
#ifndef SD_XREG_LINES_INV
#define SD_XREG_LINES_INV 8
#endif

template <bool INV, bool ZIN>
__device__ __forceinline__ void fft256_regs_wave(float2 v[16], float2* __restrict__ line,
                                                 const float2* __restrict__ s_tw, int t) {
    Dft<16, INV, ZIN>::run(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        float2 w = s_tw[t * k];
        if (INV) w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
    line_sync<16>();  // earlier reads of this line are done
#pragma unroll
    for (int k = 0; k < 16; ++k) line[k * 17 + t] = v[k];
    line_sync<16>();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = line[t * 17 + k];
    Dft<16, INV, false>::run(v);
}

// x pass, fixed-point density rows -> packed half spectra (SRC 0 of rows_fwd_body).
// RPL rows per line: 1 (imaginary part zero; Z(k), k <= N/2, is the row's
// half spectrum) or 2 (split by Hermitian symmetry).
template <int LINES, int RPL, bool HIN>
__device__ __forceinline__ void rows_fwd_reg_body(const RowsFwdArgs& a, const float2* __restrict__ tw, int bx) {
    constexpr int N = 256, h = N / 2, NR = HIN ? 8 : 16;
    static_assert(RPL == 1 || RPL == 2, "rows per line");
    __shared__ float2 s_x[LINES * XRLP];
    __shared__ float2 s_tw[N];
    for (int i = threadIdx.x; i < N; i += LINES * 16) s_tw[i] = tw[i];
    const int t = threadIdx.x & 15, c = threadIdx.x >> 4;
    const int ra = bx * (RPL * LINES) + RPL * c;
    // whole blocks of rows (the launcher checks), each nin = N/2 (HIN) or N
    // values; element offsets < 2^24
    const unsigned ia = __umul24((unsigned)ra, (unsigned)a.nin) + (unsigned)t, ib = ia + (unsigned)a.nin;
    long long qa[NR], qb[RPL == 2 ? NR : 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        qa[r] = ld_c(a.src, ia + 16u * r);
        if constexpr (RPL == 2) qb[r] = ld_c(a.src, ib + 16u * r);
    }
    if (a.clearSrc) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            st_c(a.src, ia + 16u * r, 0ll);
            if constexpr (RPL == 2) st_c(a.src, ib + 16u * r, 0ll);
        }
    }
    float2 v[16];
    const float sc = 1.0f / (float)SD_FIX_ONE;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        v[r] = make_float2(0.f, 0.f);
        if (r < NR) v[r] = make_float2((float)qa[r] * sc, RPL == 2 ? (float)qb[RPL == 2 ? r : 0] * sc : 0.f);
    }
    __syncthreads();  // s_tw
    float2* line = s_x + c * XRLP;
    fft256_regs_wave<false, HIN>(v, line, s_tw, t);
    const unsigned oa = __umul24((unsigned)ra, (unsigned)a.P), ob = oa + (unsigned)a.P;
    const float2 zh = v[8];  // Z(N/2) on thread 0
    if constexpr (RPL == 1) {
#pragma unroll
        for (int k1 = 0; k1 <= 8; ++k1) {
            const int k = t + 16 * k1;
            if (k > h || (a.pack && k == h)) continue;
            float2 z = v[k1];
            if (a.pack && k == 0) z.y = zh.x;  // bins 0 and N/2, both real
            st_c(a.out, oa + k, z);
        }
    } else {
        // Z(k) = A(k) + i B(k): A = (Z(k) + conj Z(N-k)) / 2, B = -i (Z(k) - conj Z(N-k)) / 2
        line_sync<16>();
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) line[xr_pos(t + 16 * k1)] = v[k1];
        line_sync<16>();
#pragma unroll
        for (int k1 = 0; k1 <= 8; ++k1) {
            const int k = t + 16 * k1;
            if (k > h || (a.pack && k == h)) continue;
            const float2 z = v[k1], m = line[xr_pos((N - k) & (N - 1))];
            float2 A = make_float2(0.5f * (z.x + m.x), 0.5f * (z.y - m.y));
            float2 B = make_float2(0.5f * (z.y + m.y), 0.5f * (m.x - z.x));
            if (a.pack && k == 0) {  // bin N/2 of each row: the real and imaginary parts of Z(N/2)
                A.y = zh.x;
                B.y = zh.y;
            }
            st_c(a.out, oa + k, A);
            st_c(a.out, ob + k, B);
        }
    }
}

template <int LINES, int RPL, bool HIN>
__global__ void __launch_bounds__(LINES * 16) k_rows_fwd_reg(RowsFwdArgs a, const float2* __restrict__ tw) {
    rows_fwd_reg_body<LINES, RPL, HIN>(a, tw, blockIdx.x);
}
template <int LINES, int RPL, bool HIN>
__global__ void __launch_bounds__(LINES * 16) k_rows_fwd_reg_batch(const RowsFwdArgs* __restrict__ jobs,
                                                                   const float2* __restrict__ tw) {
    const RowsFwdArgs a = jobs[blockIdx.y];
    rows_fwd_reg_body<LINES, RPL, HIN>(a, tw, blockIdx.x);
}

// x' pass, packed half spectra -> real rows (x < nout kept).
template <int LINES, bool HOUT>
__device__ __forceinline__ void rows_inv_reg_body(const RowsInvArgs& a, const float2* __restrict__ tw, int bx) {
    constexpr int N = 256, h = N / 2;
    __shared__ float2 s_x[LINES * XRLP];
    __shared__ float2 s_tw[N];
    for (int i = threadIdx.x; i < N; i += LINES * 16) s_tw[i] = tw[i];
    const int t = threadIdx.x & 15, c = threadIdx.x >> 4;
    const int ra = bx * (2 * LINES) + 2 * c;
    // whole blocks of rows (the launcher checks); element offsets < 2^24
    const unsigned ia = __umul24((unsigned)ra, (unsigned)a.P), ib = ia + (unsigned)a.P;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int k = t + 16 * r;
        const int kk = k <= h ? k : N - k;
        const unsigned src = a.pack && kk == h ? 0u : (unsigned)kk;
        float2 A = ld_c(a.in, ia + src), B = ld_c(a.in, ib + src);
        if (kk == 0 || kk == h) {  // the real bins of a real row (packed: bin N/2 in the imaginary part of bin 0)
            if (a.pack && kk == h) {
                A.x = A.y;
                B.x = B.y;
            }
            A.y = B.y = 0.f;
        }
        v[r] = k <= h ? make_float2(A.x - B.y, A.y + B.x) : make_float2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();  // s_tw
    fft256_regs_wave<true, false>(v, s_x + c * XRLP, s_tw, t);
    const unsigned oa = __umul24((unsigned)ra, (unsigned)a.nout) + (unsigned)t, ob = oa + (unsigned)a.nout;
#pragma unroll
    for (int r = 0; r < (HOUT ? 8 : 16); ++r) {
        st_c(a.out, oa + 16u * r, v[r].x);
        st_c(a.out, ob + 16u * r, v[r].y);
    }
}

template <int LINES, bool HOUT>
__global__ void __launch_bounds__(LINES * 16) k_rows_inv_reg(RowsInvArgs a, const float2* __restrict__ tw) {
    rows_inv_reg_body<LINES, HOUT>(a, tw, blockIdx.x);
}
template <int LINES, bool HOUT>
__global__ void __launch_bounds__(LINES * 16) k_rows_inv_reg_batch(const RowsInvArgs* __restrict__ jobs,
                                                                   const float2* __restrict__ tw) {
    const RowsInvArgs a = jobs[blockIdx.y];
    rows_inv_reg_body<LINES, HOUT>(a, tw, blockIdx.x);
}

struct GradJob {
    const float* pot;
    float4* acc;
    int G;
    float h;
    int periodic;
    int pad;
};

// ---- gradient: kGridAccel, Sim.cu:589-611 (grid_accel_body in sd_device.h) ---
__global__ void __launch_bounds__(256) k_grid_accel(const float* __restrict__ pot, float4* __restrict__ acc,
                                                    int G, float h, int periodic) {
    grid_accel_body(pot, acc, G, h, periodic, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Gradient of many meshes of the same size: blockIdx.y selects the mesh.
__global__ void __launch_bounds__(256) k_grid_accel_batch(const GradJob* __restrict__ jobs) {
    const GradJob j = jobs[blockIdx.y];
    grid_accel_body(j.pot, j.acc, j.G, j.h, j.periodic, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

}  // namespace sd

// ---------------------------------------------------------------------------
// Host launchers (C linkage inside the library; declared in sd_internal.h)
// ---------------------------------------------------------------------------
#include <vector>

#include "sd_internal.h"

namespace sd {

// Register-resident x passes at N = 256 (k_rows_fwd_reg, k_rows_inv_reg): the
// density rows (not the Green build) and the inverse rows, pruned when the rows
// hold G = S/2 values.
#ifndef SD_ROWS_REG_FWD
#define SD_ROWS_REG_FWD 1
#endif
#ifndef SD_ROWS_REG_INV
#define SD_ROWS_REG_INV 1
#endif
constexpr int XREG_ROWS_FWD = SD_XREG_RPL_FWD * SD_XREG_LINES_FWD, XREG_ROWS_INV = 2 * SD_XREG_LINES_INV;
// The register x passes take whole blocks of rows of N/2 (pruned) or N values.
static bool rows_fwd_reg_fits(int rows, int nin) {
    return rows % XREG_ROWS_FWD == 0 && (nin == 128 || nin == 256);
}
static bool rows_inv_reg_fits(int rows, int nout) {
    return rows % XREG_ROWS_INV == 0 && (nout == 128 || nout == 256);
}
static void launch_rows_fwd_reg(const RowsFwdArgs& a, const float2* tw, hipStream_t st) {
    const dim3 grid((a.rows + XREG_ROWS_FWD - 1) / XREG_ROWS_FWD), block(SD_XREG_LINES_FWD * 16);
    if (2 * a.nin <= 256)
        hipLaunchKernelGGL((k_rows_fwd_reg<SD_XREG_LINES_FWD, SD_XREG_RPL_FWD, true>), grid, block, 0, st, a, tw);
    else
        hipLaunchKernelGGL((k_rows_fwd_reg<SD_XREG_LINES_FWD, SD_XREG_RPL_FWD, false>), grid, block, 0, st, a, tw);
}
static void launch_rows_inv_reg(const RowsInvArgs& a, const float2* tw, hipStream_t st) {
    const dim3 grid((a.rows + XREG_ROWS_INV - 1) / XREG_ROWS_INV), block(SD_XREG_LINES_INV * 16);
    if (2 * a.nout <= 256)
        hipLaunchKernelGGL((k_rows_inv_reg<SD_XREG_LINES_INV, true>), grid, block, 0, st, a, tw);
    else
        hipLaunchKernelGGL((k_rows_inv_reg<SD_XREG_LINES_INV, false>), grid, block, 0, st, a, tw);
}

template <int N>
static void launch_rows_fwd(const RowsFwdArgs& a, int src, const float2* tw, hipStream_t st) {
    if constexpr (N == 256 && SD_ROWS_REG_FWD) {
        static_assert(!rows_fwd_carries_finalize<N>(), "the register x pass carries no reduction");
        if (src == 0 && rows_fwd_reg_fits(a.rows, a.nin)) {
            launch_rows_fwd_reg(a, tw, st);
            return;
        }
    }
    const int carry = a.fin.partials ? 1 : 0;  // only where rows_fwd_carries_finalize<N>()
    dim3 grid((a.rows + XFwd::ROWS - 1) / XFwd::ROWS + carry), block(XFwd::LINES * Plan<N>::T);
    if (src == 0)
        hipLaunchKernelGGL((k_rows_fwd<N, 0>), grid, block, 0, st, a, tw);
    else if (src == 2)
        hipLaunchKernelGGL((k_rows_fwd<N, 2>), grid, block, 0, st, a, tw);
    else
        hipLaunchKernelGGL((k_rows_fwd<N, 1>), grid, block, 0, st, a, tw);
}
template <int N>
static void launch_rows_inv(const RowsInvArgs& a, const float2* tw, hipStream_t st) {
    if constexpr (N == 256 && SD_ROWS_REG_INV) {
        if (rows_inv_reg_fits(a.rows, a.nout)) {
            launch_rows_inv_reg(a, tw, st);
            return;
        }
    }
    dim3 grid((a.rows + XInv::ROWS - 1) / XInv::ROWS), block(XInv::LINES * Plan<N>::T);
    hipLaunchKernelGGL((k_rows_inv<N>), grid, block, 0, st, a, tw);
}
// Resident blocks of a kernel on the whole device.
template <typename K>
static int resident_blocks(K kernel, int threads) {
    int per = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    return per * cus;
}

// Transform size from which the persistent, prefetching column kernels are
// used: measured faster at N = 512 (z pass -13%), slower at N = 256, where
// the prefetch registers cost more occupancy than the overlap gains.
#ifndef SD_COLS_STREAM_MIN_N
#define SD_COLS_STREAM_MIN_N 512
#endif
template <int N, int MODE>
static void launch_cols_stream(const ColArgs& a, int slabs, const float2* tw, hipStream_t st) {
    const int threads = TB * Plan<N>::T;
    const int colBlocks = (a.nb + TB - 1) / TB;
    const int ntiles = colBlocks * slabs;
    static const int resident = resident_blocks(k_cols_stream<N, MODE>, threads);  // per instantiation
    int grid = resident;
    if (grid > ntiles) grid = ntiles;
    hipLaunchKernelGGL((k_cols_stream<N, MODE>), dim3(grid), dim3(threads), 0, st, a, tw, colBlocks, ntiles);
}

// Register-resident column passes at N = 256 (k_cols_reg), with the pruned
// forms the zero padding allows: forward transforms of G = S/2 input rows
// (y forward, fused z) and inverse transforms of which only G output rows are
// stored (fused z, y inverse). The unpruned form serves every other shape.
#ifndef SD_COLS_REG
#define SD_COLS_REG 1
#endif
// Which pruned form a pass takes: forward transforms prune their input
// (y forward, fused z), inverse ones their output (fused z, y inverse); the
// fused z pass is instantiated pruned on both sides or on neither.
template <int MODE>
static bool cols_reg_pruned(int nin, int nout) {
    if (MODE == COL_FWD) return 2 * nin <= 256;
    if (MODE == COL_INV) return 2 * nout <= 256;
    return 2 * nin <= 256 && 2 * nout <= 256;
}
template <int MODE, bool P>
struct ColsRegForm {
    static constexpr bool HIN = P && MODE != COL_INV;
    static constexpr bool HOUT = P && MODE != COL_FWD;
};
// The register forms load whole lines of nin = N/2 (pruned) or N rows and
// store nout = N/2 (pruned) or N rows, with no per-row bounds.
template <int MODE>
static bool cols_reg_fits(int nin, int nout) {
    const bool pr = cols_reg_pruned<MODE>(nin, nout);
    const bool hin = pr && MODE != COL_INV, hout = pr && MODE != COL_FWD;
    return nin == (hin ? 128 : 256) && nout == (hout ? 128 : 256);
}
template <int MODE>
static void launch_cols_reg(const ColArgs& a, dim3 grid, const float2* tw, hipStream_t st) {
    using Pr = ColsRegForm<MODE, true>;
    if (cols_reg_pruned<MODE>(a.nin, a.nout))
        hipLaunchKernelGGL((k_cols_reg<MODE, Pr::HIN, Pr::HOUT>), grid, dim3(256), 0, st, a, tw);
    else
        hipLaunchKernelGGL((k_cols_reg<MODE, false, false>), grid, dim3(256), 0, st, a, tw);
}
template <int MODE>
static void launch_cols_reg_batch(const ColArgs* jobs, int nin, int nout, dim3 grid, const float2* tw,
                                  hipStream_t st) {
    using Pr = ColsRegForm<MODE, true>;
    if (cols_reg_pruned<MODE>(nin, nout))
        hipLaunchKernelGGL((k_cols_reg_batch<MODE, Pr::HIN, Pr::HOUT>), grid, dim3(256), 0, st, jobs, tw);
    else
        hipLaunchKernelGGL((k_cols_reg_batch<MODE, false, false>), grid, dim3(256), 0, st, jobs, tw);
}

// The register-resident z pass at N = 192 takes whole lines of the isolated
// G = 96 mesh: 96 rows in and out.
static bool cols192_z_fits(int nin, int nout) { return nin == 96 && nout == 96; }

template <int N>
static void launch_cols(const ColArgs& a, int mode, int slabs, const float2* tw, hipStream_t st) {
    // a carried reduction (fused z pass, no persistent form) adds a row of blocks
    dim3 grid((a.nb + TB - 1) / TB, slabs + (a.fin.partials ? 1 : 0)), block(TB * Plan<N>::T);
    if constexpr (N == 256 && SD_COLS_REG) {
        static_assert(TB == 16, "register-resident tiles are 16 columns wide");
        switch (mode) {
#define SD_REG_CASE(M)                                                      \
    case M:                                                                 \
        if (cols_reg_fits<M>(a.nin, a.nout)) {                              \
            launch_cols_reg<M>(a, grid, tw, st);                            \
            return;                                                         \
        }                                                                   \
        break;
            SD_REG_CASE(COL_FWD)
            SD_REG_CASE(COL_INV)
            SD_REG_CASE(COL_Z_ISO)
            SD_REG_CASE(COL_Z_PER)
#undef SD_REG_CASE
            default: break;  // the Green build keeps the staged kernel
        }
    }
    if constexpr (N == 192 && SD_COLS_REG) {
        static_assert(TB == 16, "register-resident tiles are 16 columns wide");
        if (mode == COL_Z_ISO && cols192_z_fits(a.nin, a.nout)) {
            hipLaunchKernelGGL(k_cols192_z, grid, dim3(256), 0, st, a, tw);
            return;
        }
    }
    if constexpr (N == 512 && SD_COLS_REG) {
        // register-resident passes: whole lines, no packed column, no carried
        // reduction; isolated meshes are padded (G = 256), periodic ones full (G = 512)
        if (!a.pack && !a.fin.partials) {
            const dim3 g((a.nb + Z512_LINES - 1) / Z512_LINES, slabs), b(Z512_LINES * 16);
            const bool inH = a.nin == 256, inF = a.nin == 512, outH = a.nout == 256, outF = a.nout == 512;
            switch (mode) {
                case COL_Z_ISO:
                    if (inH && outH) { hipLaunchKernelGGL((k_cols512_reg<COL_Z_ISO, true>), g, b, 0, st, a, tw); return; }
                    break;
                case COL_Z_PER:
                    if (inF && outF) { hipLaunchKernelGGL((k_cols512_reg<COL_Z_PER, false>), g, b, 0, st, a, tw); return; }
                    break;
                case COL_FWD:
                    if (inH && outF) { hipLaunchKernelGGL((k_cols512_reg<COL_FWD, true>), g, b, 0, st, a, tw); return; }
                    if (inF && outF) { hipLaunchKernelGGL((k_cols512_reg<COL_FWD, false>), g, b, 0, st, a, tw); return; }
                    break;
                case COL_INV:
                    if (inF && outH) { hipLaunchKernelGGL((k_cols512_reg<COL_INV, true>), g, b, 0, st, a, tw); return; }
                    if (inF && outF) { hipLaunchKernelGGL((k_cols512_reg<COL_INV, false>), g, b, 0, st, a, tw); return; }
                    break;
                default: break;
            }
        }
    }
    if constexpr (N >= SD_COLS_STREAM_MIN_N) {
        switch (mode) {
            case COL_FWD: launch_cols_stream<N, COL_FWD>(a, slabs, tw, st); return;
            case COL_INV: launch_cols_stream<N, COL_INV>(a, slabs, tw, st); return;
            case COL_Z_ISO: launch_cols_stream<N, COL_Z_ISO>(a, slabs, tw, st); return;
            case COL_Z_PER: launch_cols_stream<N, COL_Z_PER>(a, slabs, tw, st); return;
            default: break;  // the Green build keeps the one-tile kernel
        }
    }
    switch (mode) {
        case COL_FWD: hipLaunchKernelGGL((k_cols<N, COL_FWD>), grid, block, 0, st, a, tw); break;
        case COL_INV: hipLaunchKernelGGL((k_cols<N, COL_INV>), grid, block, 0, st, a, tw); break;
        case COL_Z_ISO: hipLaunchKernelGGL((k_cols<N, COL_Z_ISO>), grid, block, 0, st, a, tw); break;
        case COL_Z_PER: hipLaunchKernelGGL((k_cols<N, COL_Z_PER>), grid, block, 0, st, a, tw); break;
        default: hipLaunchKernelGGL((k_cols<N, COL_GREEN>), grid, block, 0, st, a, tw); break;
    }
}

#define SD_DISPATCH_N(N_, CALL)                 \
    switch (N_) {                               \
        case 16: CALL(16); break;               \
        case 32: CALL(32); break;               \
        case 64: CALL(64); break;               \
        case 128: CALL(128); break;             \
        case 192: CALL(192); break;             \
        case 256: CALL(256); break;             \
        case 512: CALL(512); break;             \
        default: return -1;                     \
    }

bool fft_length_supported(int n) {
    return n == 16 || n == 32 || n == 64 || n == 128 || n == 192 || n == 256 || n == 512;
}

int fft_pitch(int S) { return ((S / 2 + 1) + 15) / 16 * 16; }

// Whether the solve at FFT side S uses the packed half-spectrum layout: the
// one-tile column kernels implement it, the persistent ones (S >= 512) not.
static bool fft_packs(int S) { return S < SD_COLS_STREAM_MIN_N; }

// The pass that carries a partials reduction at FFT side N: 1 the x pass, 2
// the fused z pass (one-tile form), 0 none (it is launched before the passes).
template <int N>
static int carrier_of() {
    if constexpr (rows_fwd_carries_finalize<N>()) return 1;
    if constexpr (cols_carry_finalize<N>() && N < SD_COLS_STREAM_MIN_N) return 2;
    return 0;
}
static int finalize_carrier(int S) {
#define CALL(NN) return carrier_of<NN>()
    SD_DISPATCH_N(S, CALL)
#undef CALL
    return 0;
}

int mesh_rows_fwd(const MeshPlan& m, long long* rhoFix, float2* A, int clearSrc, const FinalizeJob* fin,
                  hipStream_t st) {
    RowsFwdArgs a{};
    if (fin) {
        const int c = finalize_carrier(m.S);
        if (c == 1) a.fin = *fin;
        if (c == 0) launch_finalize(fin->partials, fin->nparts, fin->mode, fin->pol, fin->sc, st);
    }
    a.src = rhoFix;
    a.out = A;
    a.rows = m.G * m.G;
    a.nin = m.G;
    a.P = m.P;
    a.clearSrc = clearSrc;
    a.pack = fft_packs(m.S);
#define CALL(NN) launch_rows_fwd<NN>(a, 0, m.tw, st)
    SD_DISPATCH_N(m.S, CALL)
#undef CALL
    return 0;
}

// The same x pass from a float density grid (read only). Not at S = 256,
// whose register x pass reads the int64 grid.
int mesh_rows_fwd_f32(const MeshPlan& m, const float* rho, float2* A, const FinalizeJob* fin, hipStream_t st) {
    if (m.S == 256 || m.G % 4) return -1;
    RowsFwdArgs a{};
    if (fin) {
        const int c = finalize_carrier(m.S);
        if (c == 1) a.fin = *fin;
        if (c == 0) launch_finalize(fin->partials, fin->nparts, fin->mode, fin->pol, fin->sc, st);
    }
    a.srcF = rho;
    a.out = A;
    a.rows = m.G * m.G;
    a.nin = m.G;
    a.P = m.P;
    a.pack = fft_packs(m.S);
#define CALL(NN) launch_rows_fwd<NN>(a, 2, m.tw, st)
    SD_DISPATCH_N(m.S, CALL)
#undef CALL
    return 0;
}

int mesh_solve_rest(const MeshPlan& m, float2* A, float2* B, const float* green, float* pot,
                    const FinalizeJob* fin, hipStream_t st) {
    const int G = m.G, S = m.S, P = m.P, pack = fft_packs(S) ? 1 : 0, NK = S / 2 + 1 - pack;
    // Y forward: A[z][y][kx] (y < G) -> B[z][ky][kx]
    {
        ColArgs a{};
        a.in = A; a.out = B;
        a.inA = (long long)G * P; a.inN = P;
        a.outA = (long long)S * P; a.outN = P;
        a.nb = NK; a.nin = G; a.nout = S; a.P = P; a.S = S; a.pack = pack;
#define CALL(NN) launch_cols<NN>(a, COL_FWD, G, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    // Z fused: columns over z of B for each ky
    {
        ColArgs a{};
        a.in = B; a.out = B;
        a.inA = P; a.inN = (long long)S * P;
        a.outA = P; a.outN = (long long)S * P;
        a.nb = NK; a.nin = G; a.nout = G; a.P = P; a.S = S; a.pack = pack;
        a.green = green;
        a.perScale = m.perScale;
        a.perSoft = m.perSoft;
        if (fin && finalize_carrier(S) == 2) a.fin = *fin;
        const int mode = m.periodic ? COL_Z_PER : COL_Z_ISO;
#define CALL(NN) launch_cols<NN>(a, mode, S, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    // Y inverse: B -> A (y < G)
    {
        ColArgs a{};
        a.in = B; a.out = A;
        a.inA = (long long)S * P; a.inN = P;
        a.outA = (long long)G * P; a.outN = P;
        a.nb = NK; a.nin = S; a.nout = G; a.P = P; a.S = S; a.pack = pack;
#define CALL(NN) launch_cols<NN>(a, COL_INV, G, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    // X inverse: A rows -> pot (x < G)
    {
        RowsInvArgs a{};
        a.in = A; a.out = pot; a.rows = G * G; a.P = P; a.nout = G; a.pack = pack;
#define CALL(NN) launch_rows_inv<NN>(a, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    return 0;
}

int mesh_build_green(const MeshPlan& m, float softCells, float2* tmpA, float2* tmpB, float* greenOut,
                     hipStream_t st) {
    const int S = m.S, P = m.P, NK = S / 2 + 1;
    {
        RowsFwdArgs a{};
        a.out = tmpA; a.rows = S * S; a.nin = S; a.P = P; a.S = S; a.soft = softCells;
#define CALL(NN) launch_rows_fwd<NN>(a, 1, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    {
        ColArgs a{};
        a.in = tmpA; a.out = tmpB;
        a.inA = (long long)S * P; a.inN = P;
        a.outA = (long long)S * P; a.outN = P;
        a.nb = NK; a.nin = S; a.nout = S; a.P = P; a.S = S;
#define CALL(NN) launch_cols<NN>(a, COL_FWD, S, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    {
        ColArgs a{};
        a.in = tmpB;
        a.inA = P; a.inN = (long long)S * P;
        a.nb = NK; a.nin = S; a.nout = 0; a.P = P; a.S = S;
        a.greenOut = greenOut;
        a.greenNorm = (float)(1.0 / ((double)S * S * S));
#define CALL(NN) launch_cols<NN>(a, COL_GREEN, S, m.tw, st)
        SD_DISPATCH_N(S, CALL)
#undef CALL
    }
    return 0;
}

void mesh_grid_accel(const float* pot, float4* acc, int G, float potScale, int periodic,
                     hipStream_t st) {
    const long long cells = (long long)G * G * G;
    const float h = 0.5f * (float)G * potScale;
    hipLaunchKernelGGL(k_grid_accel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, pot,
                       acc, G, h, periodic);
}

// ---------------------------------------------------------------------------
// Mesh groups: meshes of one size (G, S, boundary) solved together, one launch
// per pass for the whole group. Each job is exactly the single-mesh solve
// (same kernels, same arguments), so every mesh's result is bit-identical to
// its own mesh_rows_fwd + mesh_solve_rest + mesh_grid_accel.
// ---------------------------------------------------------------------------
struct MeshGroup {
    int G, S, P, periodic, n;
    const float2* tw;
    RowsFwdArgs* xf;
    ColArgs *yf, *z, *yi;
    RowsInvArgs* xi;
    GradJob* grad;
    void* block;  // one device allocation holding every table
};

MeshGroup* mesh_group_create(const MeshGroupJob* jobs, int n) {
    if (n <= 0) return nullptr;
    const MeshPlan& m0 = jobs[0].plan;
    for (int j = 1; j < n; ++j) {
        const MeshPlan& m = jobs[j].plan;
        if (m.G != m0.G || m.S != m0.S || m.P != m0.P || m.periodic != m0.periodic) return nullptr;
    }
    if (!fft_length_supported(m0.S)) return nullptr;
    const int G = m0.G, S = m0.S, P = m0.P, pack = fft_packs(S) ? 1 : 0, NK = S / 2 + 1 - pack;
    std::vector<RowsFwdArgs> xf(n);
    std::vector<ColArgs> yf(n), z(n), yi(n);
    std::vector<RowsInvArgs> xi(n);
    std::vector<GradJob> gr(n);
    for (int j = 0; j < n; ++j) {
        const MeshGroupJob& J = jobs[j];
        const MeshPlan& m = J.plan;
        RowsFwdArgs& a = xf[j];
        a = RowsFwdArgs{};
        a.src = J.rhoFix; a.out = J.A; a.rows = G * G; a.nin = G; a.P = P; a.clearSrc = 1; a.pack = pack;
        ColArgs& b = yf[j];
        b = ColArgs{};
        b.in = J.A; b.out = J.B;
        b.inA = (long long)G * P; b.inN = P; b.outA = (long long)S * P; b.outN = P;
        b.nb = NK; b.nin = G; b.nout = S; b.P = P; b.S = S; b.pack = pack;
        ColArgs& c = z[j];
        c = ColArgs{};
        c.in = J.B; c.out = J.B;
        c.inA = P; c.inN = (long long)S * P; c.outA = P; c.outN = (long long)S * P;
        c.nb = NK; c.nin = G; c.nout = G; c.P = P; c.S = S; c.pack = pack;
        c.green = J.green; c.perScale = m.perScale; c.perSoft = m.perSoft;
        ColArgs& d = yi[j];
        d = ColArgs{};
        d.in = J.B; d.out = J.A;
        d.inA = (long long)S * P; d.inN = P; d.outA = (long long)G * P; d.outN = P;
        d.nb = NK; d.nin = S; d.nout = G; d.P = P; d.S = S; d.pack = pack;
        RowsInvArgs& e = xi[j];
        e = RowsInvArgs{};
        e.in = J.A; e.out = J.pot; e.rows = G * G; e.P = P; e.nout = G; e.pack = pack;
        gr[j] = GradJob{J.pot, J.accG, G, 0.5f * (float)G * J.potScale, m.periodic, 0};
    }
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t oXf = 0, oYf = oXf + al(sizeof(RowsFwdArgs) * n), oZ = oYf + al(sizeof(ColArgs) * n),
                 oYi = oZ + al(sizeof(ColArgs) * n), oXi = oYi + al(sizeof(ColArgs) * n),
                 oGr = oXi + al(sizeof(RowsInvArgs) * n), total = oGr + al(sizeof(GradJob) * n);
    char* d = nullptr;
    if (hipMalloc((void**)&d, total) != hipSuccess) return nullptr;
    bool ok = hipMemcpy(d + oXf, xf.data(), sizeof(RowsFwdArgs) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d + oYf, yf.data(), sizeof(ColArgs) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d + oZ, z.data(), sizeof(ColArgs) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d + oYi, yi.data(), sizeof(ColArgs) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d + oXi, xi.data(), sizeof(RowsInvArgs) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d + oGr, gr.data(), sizeof(GradJob) * n, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        (void)hipFree(d);
        return nullptr;
    }
    MeshGroup* g = new MeshGroup();
    g->G = G; g->S = S; g->P = P; g->periodic = m0.periodic; g->n = n; g->tw = m0.tw;
    g->xf = (RowsFwdArgs*)(d + oXf);
    g->yf = (ColArgs*)(d + oYf);
    g->z = (ColArgs*)(d + oZ);
    g->yi = (ColArgs*)(d + oYi);
    g->xi = (RowsInvArgs*)(d + oXi);
    g->grad = (GradJob*)(d + oGr);
    g->block = d;
    return g;
}

void mesh_group_destroy(MeshGroup* g) {
    if (!g) return;
    (void)hipFree(g->block);
    delete g;
}

template <int N>
static void group_solve(const MeshGroup& g, bool grad, hipStream_t st) {
    const dim3 blk(TB * Plan<N>::T), fblk(XFwd::LINES * Plan<N>::T), iblk(XInv::LINES * Plan<N>::T);
    const int NK = g.S / 2 + 1 - (fft_packs(g.S) ? 1 : 0), colBlocks = (NK + TB - 1) / TB;
    const int fBlocks = (g.G * g.G + XFwd::ROWS - 1) / XFwd::ROWS, iBlocks = (g.G * g.G + XInv::ROWS - 1) / XInv::ROWS;
    if (N == 256 && SD_ROWS_REG_FWD && rows_fwd_reg_fits(g.G * g.G, g.G)) {
        const dim3 grid((g.G * g.G + XREG_ROWS_FWD - 1) / XREG_ROWS_FWD, g.n), block(SD_XREG_LINES_FWD * 16);
        if (2 * g.G <= 256)
            hipLaunchKernelGGL((k_rows_fwd_reg_batch<SD_XREG_LINES_FWD, SD_XREG_RPL_FWD, true>), grid, block, 0, st, g.xf, g.tw);
        else
            hipLaunchKernelGGL((k_rows_fwd_reg_batch<SD_XREG_LINES_FWD, SD_XREG_RPL_FWD, false>), grid, block, 0, st, g.xf, g.tw);
    } else {
        hipLaunchKernelGGL((k_rows_fwd_batch<N>), dim3(fBlocks, g.n), fblk, 0, st, g.xf, g.tw);
    }
    bool regFits = true;
    if constexpr (N == 256 && SD_COLS_REG)
        regFits = cols_reg_fits<COL_FWD>(g.G, g.S) && cols_reg_fits<COL_Z_ISO>(g.G, g.G) &&
                  cols_reg_fits<COL_INV>(g.S, g.G);
    if (N == 256 && SD_COLS_REG && regFits) {
        // the single-mesh solve's kernels and pruned forms (nin / nout as there)
        launch_cols_reg_batch<COL_FWD>(g.yf, g.G, g.S, dim3(colBlocks, g.G, g.n), g.tw, st);
        if (g.periodic)
            launch_cols_reg_batch<COL_Z_PER>(g.z, g.G, g.G, dim3(colBlocks, g.S, g.n), g.tw, st);
        else
            launch_cols_reg_batch<COL_Z_ISO>(g.z, g.G, g.G, dim3(colBlocks, g.S, g.n), g.tw, st);
        launch_cols_reg_batch<COL_INV>(g.yi, g.S, g.G, dim3(colBlocks, g.G, g.n), g.tw, st);
    } else {
        hipLaunchKernelGGL((k_cols_batch<N, COL_FWD>), dim3(colBlocks, g.G, g.n), blk, 0, st, g.yf, g.tw);
        if (N == 192 && SD_COLS_REG && !g.periodic && cols192_z_fits(g.G, g.G)) {
            // z lines of G rows in and out (the single-mesh solve's nin / nout)
            hipLaunchKernelGGL(k_cols192_z_batch, dim3(colBlocks, g.S, g.n), dim3(256), 0, st, g.z, g.tw);
        } else if (g.periodic)
            hipLaunchKernelGGL((k_cols_batch<N, COL_Z_PER>), dim3(colBlocks, g.S, g.n), blk, 0, st, g.z, g.tw);
        else
            hipLaunchKernelGGL((k_cols_batch<N, COL_Z_ISO>), dim3(colBlocks, g.S, g.n), blk, 0, st, g.z, g.tw);
        hipLaunchKernelGGL((k_cols_batch<N, COL_INV>), dim3(colBlocks, g.G, g.n), blk, 0, st, g.yi, g.tw);
    }
    if (N == 256 && SD_ROWS_REG_INV && rows_inv_reg_fits(g.G * g.G, g.G)) {
        const dim3 grid((g.G * g.G + XREG_ROWS_INV - 1) / XREG_ROWS_INV, g.n), block(SD_XREG_LINES_INV * 16);
        if (2 * g.G <= 256)
            hipLaunchKernelGGL((k_rows_inv_reg_batch<SD_XREG_LINES_INV, true>), grid, block, 0, st, g.xi, g.tw);
        else
            hipLaunchKernelGGL((k_rows_inv_reg_batch<SD_XREG_LINES_INV, false>), grid, block, 0, st, g.xi, g.tw);
    } else {
        hipLaunchKernelGGL((k_rows_inv_batch<N>), dim3(iBlocks, g.n), iblk, 0, st, g.xi, g.tw);
    }
    if (!grad) return;
    const long long cells = (long long)g.G * g.G * g.G;
    hipLaunchKernelGGL(k_grid_accel_batch, dim3((unsigned)((cells + 255) / 256), g.n), dim3(256), 0, st, g.grad);
}

int mesh_group_solve(const MeshGroup* g, bool grad, hipStream_t st) {
    if (!g) return -1;
#define CALL(NN) group_solve<NN>(*g, grad, st)
    SD_DISPATCH_N(g->S, CALL)
#undef CALL
    return 0;
}

}  // namespace sd
