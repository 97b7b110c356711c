// surrogate.hip -- the learned field surrogate: a 4-layer 3D convolutional
// operator that advances the temperature and nutrient fields by 16 FTCS
// substeps per call, with BF16 MFMA (gfx950 v_mfma_f32_32x32x16_bf16), FP32
// accumulation, hand-written training (loss, input and weight gradients,
// Adam) and inference, and the plain FTCS as residual judge and fallback.
// The reference has no surrogate; its FTCS (src/fields/diffuse.cu:19-44) is
// the operator being learned, run here by the same kernels as the world's.
//
// Model. Eight input channels (channels-last BF16): normalized T and
// nutrient, the heat source of one call, cell occupancy, a constant 1 (zero
// in the padding, so the net sees the walls), and the call's first-order
// increments computed in FP32 (see k_pack). Three hidden layers of 32
// channels with leaky ReLU, then 2 outputs: the normalized increments of T
// and nutrient over the call. Every layer is a 3x3x3 convolution with zero
// padding; tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
//
// The dynamics a call replaces, per substep: u <- FTCS(u) (Neumann walls),
// then T += dt * S (S = occupancy * per-cell heating, K/s) and
// C -= dt * k * occupancy * C (nutrient uptake).
//
// MFMA mapping (verified on the device: D row = (r/4)*8 + (lane/32)*4 + r%4,
// column = lane%32, for accumulator register r; the k index is free as long
// as A and B agree). A wave computes a tile of 32 voxels along x (rows) by 32
// output channels (columns); lane (i, h) = (lane%32, lane/32) feeds voxel or
// output channel i and the 8 k values h*8..h*8+7 of each 16-wide k step.
// The inference kernels (a net call) run the same network on a halo-padded
// activation layout, the output layer on v_mfma_f32_16x16x32_bf16 (16 output
// columns for the 2 real ones instead of 32); the training kernels keep the
// unpadded layout their gradients read.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "astro_hot.h"
#include "kernels.h"

using namespace astro_dev;

namespace astro_sur {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CH = 32;        // hidden channels (and the MFMA tile width)
constexpr int CIN0 = 8;       // input channels, padded
constexpr int NOUT = 2;       // outputs: dT, dC
constexpr int NLAYER = 4;
constexpr int TAPS = 27;
constexpr int WG_BLOCKS = 256;   // blocks of a weight-gradient pass

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }

// ---------------------------------------------------------------------------
// Convolution, one 32-voxel x 32-channel tile per wave. MODE:
//   0  hidden layer: + bias, leaky ReLU of slope `leak` (1 = linear), BF16
//      [V][32];
//   1  output layer: + bias, FP32 [V][NOUT] (output channels >= NOUT unused);
//   2  input gradient: the convolution of the next layer's dZ with its
//      flipped, transposed weights, times the activation's slope at this
//      layer's output A.
// W is [27][32 out][CINP in] BF16. nx must be a multiple of 32.
// ---------------------------------------------------------------------------
template <int CINP, int MODE>
__global__ __launch_bounds__(256) void k_conv(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                              const float* __restrict__ bias, const bf16* __restrict__ aprev,
                                              bf16* __restrict__ yb, float* __restrict__ yf, int nx, int ny, int nz,
                                              float leak) {
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int tiles_x = nx / 32;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    if (tile >= static_cast<int64_t>(tiles_x) * ny * nz) return;   // whole wave; no block barriers below
    const int tx = static_cast<int>(tile % tiles_x);
    const int64_t row = tile / tiles_x;
    const int y = static_cast<int>(row % ny), z = static_cast<int>(row / ny);
    const int x0 = tx * 32;
    const int i = lane & 31, h = lane >> 5;
    const bf16x8 zero8 = {};
    f32x16 acc = {};
    if constexpr (CINP == CH) {
        for (int t = 0; t < TAPS; ++t) {
            const int zz = z + tap_dz(t), yy = y + tap_dy(t);
            if (zz < 0 || zz >= nz || yy < 0 || yy >= ny) continue;   // wave-uniform
            const int xx = x0 + i + tap_dx(t);
            const bool ok = xx >= 0 && xx < nx;
            const int64_t vox = (static_cast<int64_t>(zz) * ny + yy) * nx + (ok ? xx : 0);
            const bf16x8* xp = reinterpret_cast<const bf16x8*>(X + vox * CH + h * 8);
            const bf16x8* wp = reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CH + h * 8);
            const bf16x8 a0 = ok ? xp[0] : zero8, a1 = ok ? xp[2] : zero8;
            acc = mfma(a0, wp[0], acc);
            acc = mfma(a1, wp[2], acc);
        }
    } else if (x0 > 0 && x0 + 32 < nx && y > 0 && y + 1 < ny && z > 0 && z + 1 < nz) {
        // 8 input channels, a tile with every operand inside the volume: the
        // same taps and MFMAs as below at a base address plus a per-tap
        // offset, with no bounds tests.
        const bf16* base = X + ((static_cast<int64_t>(z) * ny + y) * nx + x0 + i) * CIN0;
        for (int m = 0; m < (TAPS + 1) / 2; ++m) {
            const int t = 2 * m + h;
            bf16x8 a = zero8, b = zero8;
            if (t < TAPS) {
                const int64_t toff = ((static_cast<int64_t>(tap_dz(t)) * ny + tap_dy(t)) * nx + tap_dx(t)) * CIN0;
                a = *reinterpret_cast<const bf16x8*>(base + toff);
                b = *reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CIN0);
            }
            acc = mfma(a, b, acc);
        }
    } else {
        // 8 input channels: the two lane halves take two taps per k step.
        for (int m = 0; m < (TAPS + 1) / 2; ++m) {
            const int t = 2 * m + h;
            bf16x8 a = zero8, b = zero8;
            if (t < TAPS) {
                const int zz = z + tap_dz(t), yy = y + tap_dy(t), xx = x0 + i + tap_dx(t);
                if (zz >= 0 && zz < nz && yy >= 0 && yy < ny && xx >= 0 && xx < nx)
                    a = *reinterpret_cast<const bf16x8*>(X + ((static_cast<int64_t>(zz) * ny + yy) * nx + xx) * CIN0);
                b = *reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CIN0);
            }
            acc = mfma(a, b, acc);
        }
    }
    const int64_t vrow = (static_cast<int64_t>(z) * ny + y) * nx + x0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t v = vrow + (r >> 2) * 8 + h * 4 + (r & 3);
        float val = acc[r];
        if constexpr (MODE == 0) {
            val += bias[i];
            val = val > 0.0f ? val : leak * val;
            yb[v * CH + i] = static_cast<bf16>(val);
        } else if constexpr (MODE == 1) {
            if (i < NOUT) yf[v * NOUT + i] = val + bias[i];
        } else {
            const float ap = static_cast<float>(aprev[v * CH + i]);
            yb[v * CH + i] = static_cast<bf16>(ap > 0.0f ? val : leak * val);
        }
    }
}

// ---------------------------------------------------------------------------
// 32 input channels, operands from the caches, R output rows per wave: tile
// (x0..x0+31, y..y+R-1, z), so each weight operand feeds R MFMAs and the R
// accumulators are independent MFMA chains.
// ---------------------------------------------------------------------------
constexpr int CONV_ROWS = 4;

template <int MODE, int R>
__global__ __launch_bounds__(256) void k_conv_rows(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                                   const float* __restrict__ bias, const bf16* __restrict__ aprev,
                                                   bf16* __restrict__ yb, float* __restrict__ yf, int nx, int ny,
                                                   int nz, float leak) {
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int tiles_x = nx / 32, ygroups = (ny + R - 1) / R;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    if (tile >= static_cast<int64_t>(tiles_x) * ygroups * nz) return;   // whole wave
    const int tx = static_cast<int>(tile % tiles_x);
    const int64_t row = tile / tiles_x;
    const int y = static_cast<int>(row % ygroups) * R, z = static_cast<int>(row / ygroups);
    const int x0 = tx * 32;
    const int i = lane & 31, h = lane >> 5;
    const bf16x8 zero8 = {};
    f32x16 acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = f32x16{};
    // A tile whose every operand lies inside the volume (rows y-1 .. y+R,
    // planes z-1 .. z+1, voxels x0-1 .. x0+32) takes each operand at one base
    // address plus a wave-uniform offset per tap and row, with no bounds
    // tests or zero selects; the same operands meet the same MFMAs in the
    // same order as below.
    const bool inner = x0 > 0 && x0 + 32 < nx && y > 0 && y + R < ny && z > 0 && z + 1 < nz;   // wave-uniform
    if (inner) {
        const bf16* base = X + ((static_cast<int64_t>(z) * ny + y) * nx + x0 + i) * CH + h * 8;
        for (int t = 0; t < TAPS; ++t) {
            const int64_t toff = ((static_cast<int64_t>(tap_dz(t)) * ny + tap_dy(t)) * nx + tap_dx(t)) * CH;
            const bf16x8* wp = reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CH + h * 8);
            const bf16x8 b0 = wp[0], b1 = wp[2];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const bf16x8* pa =
                    reinterpret_cast<const bf16x8*>(base + toff + static_cast<int64_t>(k) * nx * CH);
                acc[k] = mfma(pa[0], b0, acc[k]);
                acc[k] = mfma(pa[2], b1, acc[k]);
            }
        }
    }
    for (int t = 0; t < TAPS && !inner; ++t) {
        const int zz = z + tap_dz(t);
        if (zz < 0 || zz >= nz) continue;   // wave-uniform
        const int xx = x0 + i + tap_dx(t);
        const bool okx = xx >= 0 && xx < nx;
        const int xs = okx ? xx : 0;
        const bf16x8* wp = reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CH + h * 8);
        const bf16x8 b0 = wp[0], b1 = wp[2];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int yy = y + k + tap_dy(t);
            const bool ok = okx && yy >= 0 && yy < ny;
            const int64_t v = (static_cast<int64_t>(zz) * ny + (ok ? yy : 0)) * nx + xs;
            const bf16x8* pa = reinterpret_cast<const bf16x8*>(X + v * CH + h * 8);
            const bf16x8 a0 = ok ? pa[0] : zero8, a1 = ok ? pa[2] : zero8;
            acc[k] = mfma(a0, b0, acc[k]);
            acc[k] = mfma(a1, b1, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int yo = y + k;
        if (yo >= ny) break;
        const int64_t vrow = (static_cast<int64_t>(z) * ny + yo) * nx + x0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t v = vrow + (q >> 2) * 8 + h * 4 + (q & 3);
            float val = acc[k][q];
            if constexpr (MODE == 0) {
                val += bias[i];
                val = val > 0.0f ? val : leak * val;
                yb[v * CH + i] = static_cast<bf16>(val);
            } else if constexpr (MODE == 1) {
                if (i < NOUT) yf[v * NOUT + i] = val + bias[i];
            } else {
                const float ap = static_cast<float>(aprev[v * CH + i]);
                yb[v * CH + i] = static_cast<bf16>(ap > 0.0f ? val : leak * val);
            }
        }
    }
}

inline unsigned conv_rows_grid(int nx, int ny, int nz) {
    return static_cast<unsigned>((static_cast<int64_t>(nx / 32) * ((ny + CONV_ROWS - 1) / CONV_ROWS) * nz + 3) / 4);
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[t][o][c] = sum_v dZ[v][o] * X[v + off_t][c], and
// db[o] = sum_v dZ[v][o]. Each block takes a contiguous run of (y, z) rows
// and, for each of the 9 (dz, dy) offsets, stages every row pair transposed
// in LDS (input channels x voxels, dZ channels x voxels) and accumulates the
// three dx taps; the 4 waves split the row's voxels and meet in LDS. The
// block's sums go to partial[block] (no float atomics across blocks: the
// Adam pass adds the partials in a fixed order).
// partial layout per block: [27][32 out][CINP in] then [32] bias.
// ---------------------------------------------------------------------------
constexpr int WG_PAD = 8;   // LDS row: x = -1 at index 7, x = 0 at index 8

template <int CINP>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16* __restrict__ X, const bf16* __restrict__ dZ,
                                                    float* __restrict__ partial, int nx, int ny, int nz) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int rowlen = nx + 2 * WG_PAD;
    bf16* xs = reinterpret_cast<bf16*>(smem);                    // [CINP][rowlen]
    bf16* zs = xs + CINP * rowlen;                               // [CH][nx]
    float* red = reinterpret_cast<float*>(zs + CH * nx);         // [3][32][32] block sums of 3 taps
    float* dbias = red + 3 * CH * CH;                            // [32]
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 31, h = lane >> 5;
    const int rows = ny * nz;
    const int per = (rows + gridDim.x - 1) / gridDim.x;
    const int r0 = min(rows, static_cast<int>(blockIdx.x) * per), r1 = min(rows, r0 + per);
    const int P = TAPS * CH * CINP + CH;
    float* out = partial + static_cast<int64_t>(blockIdx.x) * P;
    for (int q = tid; q < CINP * rowlen; q += 256) xs[q] = static_cast<bf16>(0.0f);
    for (int q = tid; q < CH; q += 256) dbias[q] = 0.0f;
    float bsum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int rr = 0; rr < 9; ++rr) {
        const int dz = rr / 3 - 1, dy = rr % 3 - 1;
        f32x16 acc[3] = {{}, {}, {}};
        for (int q = tid; q < 3 * CH * CH; q += 256) red[q] = 0.0f;
        for (int row = r0; row < r1; ++row) {
            const int y = row % ny, z = row / ny;
            const int yy = y + dy, zz = z + dz;
            const bool live = yy >= 0 && yy < ny && zz >= 0 && zz < nz;
            if (!live && rr != 4) continue;   // block-uniform
            __syncthreads();
            // stage dZ of the row and X of the shifted row, transposed
            const bf16* zrow = dZ + (static_cast<int64_t>(z) * ny + y) * nx * CH;
            for (int q = tid; q < nx * (CH / 8); q += 256) {
                const int v = q / (CH / 8), c0 = (q % (CH / 8)) * 8;
                const bf16x8 d = *reinterpret_cast<const bf16x8*>(zrow + static_cast<int64_t>(v) * CH + c0);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    zs[(c0 + e) * nx + v] = d[e];
                    if (rr == 4) bsum[e] += static_cast<float>(d[e]);
                }
            }
            if (live) {
                const bf16* xrow = X + (static_cast<int64_t>(zz) * ny + yy) * nx * CINP;
                for (int q = tid; q < nx * (CINP / 8); q += 256) {
                    const int v = q / (CINP / 8), c0 = (q % (CINP / 8)) * 8;
                    const bf16x8 d = *reinterpret_cast<const bf16x8*>(xrow + static_cast<int64_t>(v) * CINP + c0);
#pragma unroll
                    for (int e = 0; e < 8; ++e) xs[(c0 + e) * rowlen + WG_PAD + v] = d[e];
                }
            }
            __syncthreads();
            // 16 voxels per k step, the waves taking every fourth step; lane
            // (i, h) holds voxels v0 + h*8 .. v0 + h*8 + 7 of input channel i
            // (A) and output channel i (B)
            for (int v0 = wave * 16; v0 < nx; v0 += 64) {
                const int vb = v0 + h * 8;
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(zs + i * nx + vb);
                const bf16* xr = xs + (i < CINP ? i : 0) * rowlen + WG_PAD + vb;   // 16-byte aligned
                const uint4 wm = *reinterpret_cast<const uint4*>(xr - 8);
                const uint4 w0 = *reinterpret_cast<const uint4*>(xr);
                const uint4 wp = *reinterpret_cast<const uint4*>(xr + 8);
                // x - 1: elements vb-1 .. vb+6; x + 1: vb+1 .. vb+8
                uint4 lo, hi;
                lo.x = __builtin_amdgcn_alignbyte(w0.x, wm.w, 2);
                lo.y = __builtin_amdgcn_alignbyte(w0.y, w0.x, 2);
                lo.z = __builtin_amdgcn_alignbyte(w0.z, w0.y, 2);
                lo.w = __builtin_amdgcn_alignbyte(w0.w, w0.z, 2);
                hi.x = __builtin_amdgcn_alignbyte(w0.y, w0.x, 2);
                hi.y = __builtin_amdgcn_alignbyte(w0.z, w0.y, 2);
                hi.z = __builtin_amdgcn_alignbyte(w0.w, w0.z, 2);
                hi.w = __builtin_amdgcn_alignbyte(wp.x, w0.w, 2);
                bf16x8 am, a0, ap;
                __builtin_memcpy(&am, &lo, 16);
                __builtin_memcpy(&a0, &w0, 16);
                __builtin_memcpy(&ap, &hi, 16);
                if (i >= CINP) am = a0 = ap = bf16x8{};
                acc[0] = mfma(am, b, acc[0]);
                acc[1] = mfma(a0, b, acc[1]);
                acc[2] = mfma(ap, b, acc[2]);
            }
        }
        // the waves' sums meet in LDS; D row = input channel, column = output
        // channel
        __syncthreads();
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int c = (r >> 2) * 8 + h * 4 + (r & 3);
                atomicAdd(&red[(d * CH + i) * CH + c], acc[d][r]);
            }
        __syncthreads();
        for (int q = tid; q < 3 * CH * CINP; q += 256) {
            const int d = q / (CH * CINP), o = (q / CINP) % CH, c = q % CINP;
            out[((rr * 3 + d) * CH + o) * CINP + c] = red[(d * CH + o) * CH + c];
        }
        __syncthreads();   // before the next offset clears `red`
    }
    // bias: thread tid staged channels (tid % 4) * 8 .. + 7 of its voxels
#pragma unroll
    for (int e = 0; e < 8; ++e) atomicAdd(&dbias[(tid % (CH / 8)) * 8 + e], bsum[e]);
    __syncthreads();
    if (tid < CH) out[TAPS * CH * CINP + tid] = dbias[tid];
}

template <int CINP>
size_t wgrad_lds(int nx) {
    return sizeof(bf16) * (CINP * (nx + 2 * WG_PAD) + CH * nx) + sizeof(float) * (3 * CH * CH + CH);
}

// ---------------------------------------------------------------------------
// Adam on one layer: sums the weight-gradient partials in block order,
// updates the FP32 master weights and moments, writes the BF16 weights and,
// for 32-channel inputs, the flipped transposed copy the input-gradient pass
// uses (wd[26 - t][c][o] = w[t][o][c]). Output channels >= nout stay zero.
// ---------------------------------------------------------------------------
struct AdamArgs {
    const float* partial;
    int nblocks;
    int cinp, nout;
    float* master;   // [27*32*cinp + 32]
    float* m;
    float* v;
    bf16* w;
    bf16* wd;        // null for the input layer
    float* bias;
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 - b1^t, c2 = 1 - b2^t
    float gscale;
};

__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    const int P = TAPS * CH * a.cinp + CH;
    const int p = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (p >= P) return;
    const bool is_bias = p >= TAPS * CH * a.cinp;
    const int o = is_bias ? p - TAPS * CH * a.cinp : (p / a.cinp) % CH;
    if (o >= a.nout) return;
    float g = 0.0f;
    for (int b = 0; b < a.nblocks; ++b) g += a.partial[static_cast<int64_t>(b) * P + p];
    g *= a.gscale;
    const float mm = a.b1 * a.m[p] + (1.0f - a.b1) * g;
    const float vv = a.b2 * a.v[p] + (1.0f - a.b2) * g * g;
    a.m[p] = mm;
    a.v[p] = vv;
    const float wnew = a.master[p] - a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
    a.master[p] = wnew;
    if (is_bias) {
        a.bias[o] = wnew;
        return;
    }
    a.w[p] = static_cast<bf16>(wnew);
    if (a.wd) {
        const int t = p / (CH * a.cinp), c = p % a.cinp;
        a.wd[(static_cast<int64_t>(TAPS - 1 - t) * CH + c) * CH + o] = static_cast<bf16>(wnew);
    }
}

// ---------------------------------------------------------------------------
// Elementwise passes
// ---------------------------------------------------------------------------
struct Norm {
    float t_ref, t_amp;     // input T = (T - t_ref) / t_amp
    float c_ref, c_amp;
    float occ_amp;
    float dt_out, dc_out;   // outputs: dT = dt_out * y0, dC = dc_out * y1
    float dt_call;
};

// Input channels: T, C, the call's heat source, occupancy, the constant 1,
// and three first-order increments computed in FP32 from the fields' own
// neighbours (the call's Euler diffusion increment of T and of C, and its
// Euler uptake of C), each in output units. BF16 cannot hold the small
// differences between neighbouring values that diffusion acts on, so the
// network gets them pre-formed and learns the rest of the 16-substep operator.
struct PackArgs {
    const float *T, *C, *S, *occ;
    bf16* x;
    int nx, ny, nz;
    float ctx, cty, ctz;   // call-integrated diffusion numbers of T (16 x c)
    float ccx, ccy, ccz;   // of C
    float uptake_call;     // dt_call * uptake
    int px, py;            // padded destination: row length and rows per plane (0: unpadded [V][8])
};

__device__ __forceinline__ float lap_neumann(const float* u, int64_t v, int x, int y, int z, int nx, int ny, int nz,
                                             float cx, float cy, float cz) {
    const int64_t sx = 1, sy = nx, sz = static_cast<int64_t>(nx) * ny;
    const float c = u[v];
    const float l = x > 0 ? u[v - sx] : c, r = x < nx - 1 ? u[v + sx] : c;
    const float d = y > 0 ? u[v - sy] : c, f = y < ny - 1 ? u[v + sy] : c;
    const float b = z > 0 ? u[v - sz] : c, t = z < nz - 1 ? u[v + sz] : c;
    return cx * ((l - c) + (r - c)) + cy * ((d - c) + (f - c)) + cz * ((b - c) + (t - c));
}

__global__ __launch_bounds__(256) void k_pack(PackArgs a, Norm nm) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= n) return;
    const int x = static_cast<int>(v % a.nx), y = static_cast<int>((v / a.nx) % a.ny);
    const int z = static_cast<int>(v / (static_cast<int64_t>(a.nx) * a.ny));
    const float tc = a.T[v], cc = a.C[v], oc = a.occ[v];
    bf16x8 o = {};
    o[0] = static_cast<bf16>((tc - nm.t_ref) / nm.t_amp);
    o[1] = static_cast<bf16>((cc - nm.c_ref) / nm.c_amp);
    o[2] = static_cast<bf16>(a.S[v] * nm.dt_call / nm.dt_out);
    o[3] = static_cast<bf16>(oc / nm.occ_amp);
    o[4] = static_cast<bf16>(1.0f);
    o[5] = static_cast<bf16>(lap_neumann(a.T, v, x, y, z, a.nx, a.ny, a.nz, a.ctx, a.cty, a.ctz) / nm.dt_out);
    o[6] = static_cast<bf16>(lap_neumann(a.C, v, x, y, z, a.nx, a.ny, a.nz, a.ccx, a.ccy, a.ccz) / nm.dc_out);
    o[7] = static_cast<bf16>(-a.uptake_call * oc * cc / nm.dc_out);
    const int64_t dst = a.px ? (static_cast<int64_t>(z + 1) * a.py + (y + 1)) * a.px + (x + 1) : v;
    *reinterpret_cast<bf16x8*>(a.x + dst * CIN0) = o;
}

// target = normalized increments; also squared-norm partials of the raw
// increments (for the scales)
__global__ __launch_bounds__(256) void k_target(const float* T0, const float* C0, const float* T1, const float* C1,
                                                float* tgt, int64_t n, Norm nm) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= n) return;
    tgt[v * NOUT + 0] = (T1[v] - T0[v]) / nm.dt_out;
    tgt[v * NOUT + 1] = (C1[v] - C0[v]) / nm.dc_out;
}

// MSE over voxels and both outputs; dZ of the output layer (BF16, 32
// channels, >= NOUT zero) and per-block loss partials. The gradient is scaled
// by gscale_in (to keep BF16 gradients in range); Adam divides it back out.
__global__ __launch_bounds__(256) void k_loss(const float* pred, const float* tgt, bf16* dz, float* lpart,
                                              int64_t n, float gscale_in) {
    __shared__ float s[256];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    float l = 0.0f;
    if (v < n) {
        bf16x8 o0 = {}, o1 = {}, o2 = {}, o3 = {};
        const float d0 = pred[v * NOUT] - tgt[v * NOUT], d1 = pred[v * NOUT + 1] - tgt[v * NOUT + 1];
        l = d0 * d0 + d1 * d1;
        o0[0] = static_cast<bf16>(2.0f * d0 * gscale_in);
        o0[1] = static_cast<bf16>(2.0f * d1 * gscale_in);
        bf16x8* p = reinterpret_cast<bf16x8*>(dz + v * CH);
        p[0] = o0;
        p[1] = o1;
        p[2] = o2;
        p[3] = o3;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (static_cast<int>(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// The plain path's substep in one pass over both fields: the FTCS of T and of
// C (Neumann walls: a missing neighbour is the cell's own value), then the
// sources, T' = T* + dt S and C' = C* - (dt k) occ C*. A thread takes 4
// consecutive x points over PZC planes, so each plane it loads serves as the
// centre of one output and the z neighbour of two others. 24 compulsory bytes
// per point: T, C, S, occupancy in, T', C' out.
// ---------------------------------------------------------------------------
constexpr int PZC = 4;

struct PlainArgs {
    const float *T, *C, *S, *occ;
    float *To, *Co;
    int nx, ny, nz;
    float tcx, tcy, tcz;   // diffusion numbers of T per substep
    float ccx, ccy, ccz;   // of C
    float dt, du;          // dt_sub and dt_sub * uptake
};

__device__ __forceinline__ float4 ld4f(const float* p) { return *reinterpret_cast<const float4*>(p); }

// One field's 4 x points at planes k0 .. k0+PZC-1 (those below nz).
__device__ __forceinline__ void ftcs4(const float* __restrict__ u, int x, int y, int k0, int nx, int ny, int nz,
                                      float cx, float cy, float cz, float4 (&out)[PZC]) {
    const int64_t pl = static_cast<int64_t>(nx) * ny;
    const int64_t row = static_cast<int64_t>(y) * nx + x;
    float4 pz[PZC + 2];
#pragma unroll
    for (int m = 0; m < PZC + 2; ++m) {
        const int k = k0 - 1 + m;
        pz[m] = (k >= 0 && k < nz) ? ld4f(u + k * pl + row) : float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int m = 0; m < PZC; ++m) {
        const int k = k0 + m;
        if (k >= nz) break;
        const int64_t id = k * pl + row;
        const float4 c4 = pz[m + 1];
        const float4 b4 = k > 0 ? pz[m] : c4, f4 = k < nz - 1 ? pz[m + 2] : c4;
        const float4 d4 = y > 0 ? ld4f(u + id - nx) : c4, p4 = y < ny - 1 ? ld4f(u + id + nx) : c4;
        const float l0 = x > 0 ? u[id - 1] : c4.x, r3 = x + 4 < nx ? u[id + 4] : c4.w;
        const float c[4] = {c4.x, c4.y, c4.z, c4.w}, b[4] = {b4.x, b4.y, b4.z, b4.w};
        const float f[4] = {f4.x, f4.y, f4.z, f4.w}, dn[4] = {d4.x, d4.y, d4.z, d4.w};
        const float up[4] = {p4.x, p4.y, p4.z, p4.w};
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float l = e > 0 ? c[e - 1] : l0, r = e < 3 ? c[e + 1] : r3, cc = c[e];
            o[e] = cc + (cx * ((l + r) - 2.0f * cc) + cy * ((dn[e] + up[e]) - 2.0f * cc) +
                         cz * ((b[e] + f[e]) - 2.0f * cc));
        }
        out[m] = float4{o[0], o[1], o[2], o[3]};
    }
}

__global__ __launch_bounds__(256) void k_plain_substep(PlainArgs a) {
    const int nx4 = a.nx / 4;
    const int zchunks = (a.nz + PZC - 1) / PZC;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= static_cast<int64_t>(nx4) * a.ny * zchunks) return;
    const int x = static_cast<int>(t % nx4) * 4;
    const int64_t r = t / nx4;
    const int y = static_cast<int>(r % a.ny), k0 = static_cast<int>(r / a.ny) * PZC;
    float4 tv[PZC], cv[PZC];
    ftcs4(a.T, x, y, k0, a.nx, a.ny, a.nz, a.tcx, a.tcy, a.tcz, tv);
    ftcs4(a.C, x, y, k0, a.nx, a.ny, a.nz, a.ccx, a.ccy, a.ccz, cv);
    const int64_t pl = static_cast<int64_t>(a.nx) * a.ny;
#pragma unroll
    for (int m = 0; m < PZC; ++m) {
        const int k = k0 + m;
        if (k >= a.nz) break;
        const int64_t id = k * pl + static_cast<int64_t>(y) * a.nx + x;
        const float4 s4 = ld4f(a.S + id), q4 = ld4f(a.occ + id);
        const float4 t4 = tv[m], c4 = cv[m];
        *reinterpret_cast<float4*>(a.To + id) =
            float4{t4.x + a.dt * s4.x, t4.y + a.dt * s4.y, t4.z + a.dt * s4.z, t4.w + a.dt * s4.w};
        *reinterpret_cast<float4*>(a.Co + id) =
            float4{c4.x - a.du * q4.x * c4.x, c4.y - a.du * q4.y * c4.y, c4.z - a.du * q4.z * c4.z,
                   c4.w - a.du * q4.w * c4.w};
    }
}

inline unsigned plain_grid(int nx, int ny, int nz) {
    return static_cast<unsigned>((static_cast<int64_t>(nx / 4) * ny * ((nz + PZC - 1) / PZC) + 255) / 256);
}

// ---------------------------------------------------------------------------
// Inference on a halo-padded layout. Activations live in [PZ][PY][PX][ch]
// with PX = nx + 2, PY = ny rounded up to a multiple of RROWS, + 2, and
// PZ = nz rounded up to a multiple of ZPLANES, + 2; voxel (x, y, z) sits at
// (x + 1, y + 1, z + 1) and everything else stays zero (written once at
// allocation, never again), which is exactly the zero padding of the
// convolution. Every tile then reads its operands at a base address plus
// wave-uniform offsets, with no bounds tests or selects.
// ---------------------------------------------------------------------------
constexpr int RROWS = 8;

struct Pad {
    int px, py, pz;
    __host__ __device__ int64_t at(int x, int y, int z) const {
        return (static_cast<int64_t>(z + 1) * py + (y + 1)) * px + (x + 1);
    }
};

// Output planes a z-marching wave owns (k_conv_hid_z); the padding covers
// whole groups of them.
constexpr int ZPLANES = 16;

inline Pad pad_of(int nx, int ny, int nz) {
    return Pad{nx + 2, (ny + RROWS - 1) / RROWS * RROWS + 2, (nz + ZPLANES - 1) / ZPLANES * ZPLANES + 2};
}

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Input layer (8 channels), padded in, padded out: the two lane halves take
// two taps per 16-wide k step, as in k_conv; the weight operands are loaded
// once per wave and reused over its RROWS rows.
__global__ __launch_bounds__(256) void k_conv_in_p(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                                   const float* __restrict__ bias, bf16* __restrict__ Y, int nx,
                                                   int ny, int nz, Pad p, float leak) {
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int tiles_x = nx / 32, yg = (ny + RROWS - 1) / RROWS;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    if (tile >= static_cast<int64_t>(tiles_x) * yg * nz) return;   // whole wave
    const int x0 = static_cast<int>(tile % tiles_x) * 32;
    const int64_t row = tile / tiles_x;
    const int y0 = static_cast<int>(row % yg) * RROWS, z = static_cast<int>(row / yg);
    const int i = lane & 31, h = lane >> 5;
    constexpr int STEPS = (TAPS + 1) / 2;
    bf16x8 b[STEPS];
    int64_t toff[STEPS];
#pragma unroll
    for (int m = 0; m < STEPS; ++m) {
        const int t = 2 * m + h;
        b[m] = t < TAPS ? *reinterpret_cast<const bf16x8*>(W + (static_cast<int64_t>(t) * CH + i) * CIN0) : bf16x8{};
        toff[m] = t < TAPS ? ((static_cast<int64_t>(tap_dz(t)) * p.py + tap_dy(t)) * p.px + tap_dx(t)) * CIN0 : 0;
    }
    const float bi = bias[i];
    for (int k = 0; k < RROWS; ++k) {
        const int y = y0 + k;
        if (y >= ny) break;   // wave-uniform
        const bf16* base = X + p.at(x0 + i, y, z) * CIN0;
        f32x16 acc = {};
#pragma unroll
        for (int m = 0; m < STEPS; ++m) {
            const bf16x8 a = 2 * m + h < TAPS ? *reinterpret_cast<const bf16x8*>(base + toff[m]) : bf16x8{};
            acc = mfma(a, b[m], acc);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int x = x0 + (q >> 2) * 8 + h * 4 + (q & 3);
            float val = acc[q] + bi;
            val = val > 0.0f ? val : leak * val;
            Y[p.at(x, y, z) * CH + i] = static_cast<bf16>(val);
        }
    }
}

// A hidden layer (32 -> 32, + bias, leaky ReLU), padded in, padded out,
// marching along z. A wave owns 32 voxels along x by R rows
// by Z planes. It walks the input planes z0-1 .. z0+Z once; each input plane
// feeds three output planes (dz = -1, 0, +1), whose accumulators it keeps in
// a ring of three, so every input row is loaded once per dx instead of once
// per (dz, dx). The block's four waves share the layer's weights, staged once
// in LDS (27 x 32 x 32 BF16, 54 KB). Per output row the taps are summed in
// (dz, dx, dy) order. Each input row loaded for a dx feeds the (up to) three
// output rows whose dy tap reads it. The ring slot of output plane zo is
// (zo - z0 + 1) mod 3, fixed at compile time by unrolling the march by three.
template <int R, int Z>
__global__ __launch_bounds__(256) void k_conv_hid_z(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                                    const float* __restrict__ bias, bf16* __restrict__ Y, int nx,
                                                    int ny, int nz, Pad p, float leak) {
    __shared__ __align__(16) bf16 sw[TAPS * CH * CH];
    for (int q = static_cast<int>(threadIdx.x); q < TAPS * CH * CH / 8; q += 256)
        reinterpret_cast<bf16x8*>(sw)[q] = reinterpret_cast<const bf16x8*>(W)[q];
    __syncthreads();   // the last block barrier: waves without a tile leave below
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int tiles_x = nx / 32, yg = (ny + R - 1) / R, zg = (nz + Z - 1) / Z;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    if (tile >= static_cast<int64_t>(tiles_x) * yg * zg) return;   // whole wave
    const int x0 = static_cast<int>(tile % tiles_x) * 32;
    const int64_t rest = tile / tiles_x;
    const int y0 = static_cast<int>(rest % yg) * R, z0 = static_cast<int>(rest / yg) * Z;
    const int zend = z0 + Z < nz ? z0 + Z : nz;   // output planes [z0, zend)
    const int i = lane & 31, h = lane >> 5;
    const int64_t rstride = static_cast<int64_t>(p.px) * CH;
    const float bi = bias[i];
    f32x16 acc[3][R];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int k = 0; k < R; ++k) acc[s][k] = f32x16{};
    // One input plane zi whose ring position is J = (zi - z0 + 1) mod 3:
    // output plane zi - dz sits in slot (J - dz) mod 3.
    const auto plane = [&](const int zi, auto J) {
        constexpr int j = decltype(J)::value;
#pragma unroll 1
        for (int dx = -1; dx <= 1; ++dx) {
            const bf16* rb = X + p.at(x0 + i + dx, y0 - 1, zi) * CH + h * 8;
            bf16x8 a[R + 2][2];
#pragma unroll
            for (int m = 0; m < R + 2; ++m) {
                const bf16x8* pa = reinterpret_cast<const bf16x8*>(rb + m * rstride);
                a[m][0] = pa[0];
                a[m][1] = pa[2];
            }
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) {
                const int zo = zi - dz;
                if (zo < z0 || zo >= zend) continue;   // wave-uniform
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy) {
                    const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                    const bf16x8* wp = reinterpret_cast<const bf16x8*>(sw + (t * CH + i) * CH + h * 8);
                    const bf16x8 b0 = wp[0], b1 = wp[2];
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        // slot of output plane zi - dz
                        f32x16& c = acc[(j - dz + 3) % 3][k];
                        c = mfma(a[k + 1 + dy][0], b0, c);
                        c = mfma(a[k + 1 + dy][1], b1, c);
                    }
                }
            }
        }
        // Output plane zi - 1 has had its last input plane: store and clear.
        constexpr int done = (j + 2) % 3;
        const int zo = zi - 1;
        if (zo >= z0 && zo < zend) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int y = y0 + k;
                if (y >= ny) break;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int x = x0 + (q >> 2) * 8 + h * 4 + (q & 3);
                    float val = acc[done][k][q] + bi;
                    val = val > 0.0f ? val : leak * val;
                    Y[p.at(x, y, zo) * CH + i] = static_cast<bf16>(val);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) acc[done][k] = f32x16{};
    };
    for (int zi = z0 - 1; zi <= zend; zi += 3) {
        plane(zi, std::integral_constant<int, 0>{});
        if (zi + 1 <= zend) plane(zi + 1, std::integral_constant<int, 1>{});
        if (zi + 2 <= zend) plane(zi + 2, std::integral_constant<int, 2>{});
    }
}

// The output layer (32 -> 2), marching along z as k_conv_hid_z, on
// 16 x 16 x 32 MFMAs: 16 voxels by 16 output columns (2 of them real) per
// instruction, all 32 input channels of one tap in its k. Lane l feeds voxel
// l % 16 with channels (l / 16) * 8 .. + 7, and weight column l % 16 with the
// same channels; it receives column l % 16 of rows (l / 16) * 4 .. + 3. The
// LDS holds the 16 weight columns the MFMAs read. With `apply`, the outputs go
// straight into the fields (T += dt_out y0, C += dc_out y1); else into pred
// [V][2].
template <int R, int Z>
__global__ __launch_bounds__(256) void k_conv_out_z(const bf16* __restrict__ X, const bf16* __restrict__ W,
                                                    const float* __restrict__ bias, float* __restrict__ pred,
                                                    float* __restrict__ T, float* __restrict__ C, int nx, int ny,
                                                    int nz, Pad p, float dt_out, float dc_out, int apply) {
    __shared__ __align__(16) bf16 sw[TAPS * 16 * CH];   // [tap][column < 16][input channel]
    for (int q = static_cast<int>(threadIdx.x); q < TAPS * 16 * CH / 8; q += 256) {
        const int t = q / (16 * CH / 8), rem = q % (16 * CH / 8);
        reinterpret_cast<bf16x8*>(sw)[q] = reinterpret_cast<const bf16x8*>(W + static_cast<int64_t>(t) * CH * CH)[rem];
    }
    __syncthreads();   // the last block barrier: waves without a tile leave below
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    const int tiles_x = nx / 32, yg = (ny + R - 1) / R, zg = (nz + Z - 1) / Z;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    if (tile >= static_cast<int64_t>(tiles_x) * yg * zg) return;   // whole wave
    const int x0 = static_cast<int>(tile % tiles_x) * 32;
    const int64_t rest = tile / tiles_x;
    const int y0 = static_cast<int>(rest % yg) * R, z0 = static_cast<int>(rest / yg) * Z;
    const int zend = z0 + Z < nz ? z0 + Z : nz;
    const int n = lane & 15, g = lane >> 4;
    const int64_t rstride = static_cast<int64_t>(p.px) * CH;
    const float bi = n < NOUT ? bias[n] : 0.0f;
    f32x4 acc[3][R][2];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int k = 0; k < R; ++k) acc[s][k][0] = acc[s][k][1] = f32x4{};
    const auto plane = [&](const int zi, auto J) {
        constexpr int j = decltype(J)::value;
#pragma unroll 1
        for (int dx = -1; dx <= 1; ++dx) {
            const bf16* rb = X + p.at(x0 + n + dx, y0 - 1, zi) * CH + g * 8;
            bf16x8 a[R + 2][2];
#pragma unroll
            for (int m = 0; m < R + 2; ++m) {
                a[m][0] = *reinterpret_cast<const bf16x8*>(rb + m * rstride);
                a[m][1] = *reinterpret_cast<const bf16x8*>(rb + m * rstride + 16 * CH);
            }
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) {
                const int zo = zi - dz;
                if (zo < z0 || zo >= zend) continue;   // wave-uniform
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy) {
                    const int t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(sw + (t * 16 + n) * CH + g * 8);
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        f32x4* c = acc[(j - dz + 3) % 3][k];
                        c[0] = mfma16(a[k + 1 + dy][0], b, c[0]);
                        c[1] = mfma16(a[k + 1 + dy][1], b, c[1]);
                    }
                }
            }
        }
        constexpr int done = (j + 2) % 3;
        const int zo = zi - 1;
        if (n < NOUT && zo >= z0 && zo < zend) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int y = y0 + k;
                if (y >= ny) break;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int x = x0 + half * 16 + g * 4 + r;
                        const int64_t v = (static_cast<int64_t>(zo) * ny + y) * nx + x;
                        const float val = acc[done][k][half][r] + bi;
                        if (!apply) pred[v * NOUT + n] = val;
                        else if (n == 0) T[v] += dt_out * val;
                        else C[v] += dc_out * val;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) acc[done][k][0] = acc[done][k][1] = f32x4{};
    };
    for (int zi = z0 - 1; zi <= zend; zi += 3) {
        plane(zi, std::integral_constant<int, 0>{});
        if (zi + 1 <= zend) plane(zi + 1, std::integral_constant<int, 1>{});
        if (zi + 2 <= zend) plane(zi + 2, std::integral_constant<int, 2>{});
    }
}

inline unsigned padded_grid(int nx, int ny, int nz, int rows = RROWS) {
    return static_cast<unsigned>((static_cast<int64_t>(nx / 32) * ((ny + rows - 1) / rows) * nz + 3) / 4);
}

// Rows and planes per wave, measured at 128^3: the z-marching hidden layer at
// 2 rows x 16 planes takes 131 us (133 at 8 planes; one plane per wave with 4
// rows took 216); the output layer at 2 rows x 16 planes 182 us (207 at 8
// planes; 4 rows leave one wave per SIMD and take 255).
constexpr int HID_ROWS = 2;
constexpr int OUT_ROWS = 2;
static_assert(RROWS % HID_ROWS == 0 && RROWS % OUT_ROWS == 0, "the padding must cover every row group");

// Unpadded inputs [V][8] into the padded layout's interior.
__global__ __launch_bounds__(256) void k_pad_input(const bf16* __restrict__ x, bf16* __restrict__ xp, int nx, int ny,
                                                   int nz, Pad p) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= static_cast<int64_t>(nx) * ny * nz) return;
    const int xx = static_cast<int>(v % nx), yy = static_cast<int>((v / nx) % ny);
    const int zz = static_cast<int>(v / (static_cast<int64_t>(nx) * ny));
    *reinterpret_cast<bf16x8*>(xp + p.at(xx, yy, zz) * CIN0) = *reinterpret_cast<const bf16x8*>(x + v * CIN0);
}

// Per block: the squared difference of two increments from the same start
// (candidate a1 - start, reference b1 - start) and the reference's squared
// norm, in double.
__global__ __launch_bounds__(256) void k_compare(const float* a1, const float* b1, const float* start, double* out,
                                                 int64_t n) {
    __shared__ double s[2][256];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    double e = 0.0, r = 0.0;
    if (v < n) {
        const double da = static_cast<double>(a1[v]) - start[v], db = static_cast<double>(b1[v]) - start[v];
        e = (da - db) * (da - db);
        r = db * db;
    }
    s[0][threadIdx.x] = e;
    s[1][threadIdx.x] = r;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (static_cast<int>(threadIdx.x) < d) {
            s[0][threadIdx.x] += s[0][threadIdx.x + d];
            s[1][threadIdx.x] += s[1][threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s[0][0];
        out[2 * blockIdx.x + 1] = s[1][0];
    }
}

__global__ __launch_bounds__(256) void k_sum(const float* a, double* out, int64_t n) {
    __shared__ double s[256];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    s[threadIdx.x] = v < n ? static_cast<double>(a[v]) : 0.0;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (static_cast<int>(threadIdx.x) < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// Random states of the training distribution: Gaussian warm and cool spots
// in T, nutrient gradients, and a 25k-cell culture of up to four clusters
// (occupancy counts; heating per cell drawn per sample).
// ---------------------------------------------------------------------------
constexpr int MAX_BLOBS = 6;
struct Blob {
    float cx, cy, cz, inv2s2, amp_t, amp_c;
};
struct GenArgs {
    Blob blob[MAX_BLOBS];
    int nblob;
    int nx, ny, nz;
    float dx, dy, dz;
    float t_amb, c_amb;
    float gcx, gcy;   // nutrient gradient per metre (relative)
};

__global__ __launch_bounds__(256) void k_gen_fields(float* T, float* C, GenArgs g) {
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= n) return;
    const int x = static_cast<int>(v % g.nx), y = static_cast<int>((v / g.nx) % g.ny), z = static_cast<int>(v / (static_cast<int64_t>(g.nx) * g.ny));
    const float px = (x + 0.5f) * g.dx, py = (y + 0.5f) * g.dy, pz = (z + 0.5f) * g.dz;
    float t = g.t_amb, c = 1.0f + g.gcx * (px - 0.5f * g.nx * g.dx) + g.gcy * (py - 0.5f * g.ny * g.dy);
    for (int b = 0; b < g.nblob; ++b) {
        const Blob& q = g.blob[b];
        const float r2 = (px - q.cx) * (px - q.cx) + (py - q.cy) * (py - q.cy) + (pz - q.cz) * (pz - q.cz);
        const float w = __expf(-r2 * q.inv2s2);
        t += q.amp_t * w;
        c += q.amp_c * w;
    }
    T[v] = t;
    C[v] = g.c_amb * fmaxf(c, 0.05f);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float unit(uint64_t s) { return static_cast<float>(s >> 40) * (1.0f / 16777216.0f); }

struct CellArgs {
    int ncells, nclusters;
    float ccx[4], ccy[4], ccz[4], cr[4];   // cluster centres and radii (m); radius 0 = uniform
    uint64_t seed;
    int nx, ny, nz;
    float dx, dy, dz;
};

__global__ __launch_bounds__(256) void k_gen_cells(unsigned int* count, CellArgs a) {
    const int k = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (k >= a.ncells) return;
    const uint64_t s0 = mix64(a.seed ^ (static_cast<uint64_t>(k) * 0xD1B54A32D192ED03ull));
    const uint64_t s1 = mix64(s0), s2 = mix64(s1), s3 = mix64(s2);
    const int cl = static_cast<int>(unit(s3) * a.nclusters) % a.nclusters;
    float px, py, pz;
    const float W = a.nx * a.dx, H = a.ny * a.dy, D = a.nz * a.dz;
    if (a.cr[cl] <= 0.0f) {
        px = unit(s0) * W;
        py = unit(s1) * H;
        pz = unit(s2) * D;
    } else {
        // a rough Gaussian cluster from sums of uniforms
        const uint64_t s4 = mix64(s3), s5 = mix64(s4);
        px = a.ccx[cl] + a.cr[cl] * (unit(s0) + unit(s4) - 1.0f) * 1.7f;
        py = a.ccy[cl] + a.cr[cl] * (unit(s1) + unit(s5) - 1.0f) * 1.7f;
        pz = a.ccz[cl] + 0.3f * a.cr[cl] * (unit(s2) * 2.0f - 1.0f);
    }
    const int x = min(a.nx - 1, max(0, static_cast<int>(px / a.dx)));
    const int y = min(a.ny - 1, max(0, static_cast<int>(py / a.dy)));
    const int z = min(a.nz - 1, max(0, static_cast<int>(pz / a.dz)));
    atomicAdd(&count[(static_cast<int64_t>(z) * a.ny + y) * a.nx + x], 1u);
}

__global__ __launch_bounds__(256) void k_counts(const unsigned int* count, float* occ, float* S, int64_t n,
                                                float heat) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= n) return;
    const float c = static_cast<float>(count[v]);
    occ[v] = c;
    S[v] = c * heat;
}

inline unsigned grid1(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace astro_sur

using namespace astro_sur;

// ===========================================================================
// Host side
// ===========================================================================
struct astro_surrogate {
    astro_surrogate_desc d{};
    int64_t n = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Norm nm{};
    float heat_max = 0.0f;   // K/s per cell, upper end of the training draw

    // fields: T, C and their ping-pong scratch, sources, a saved start
    float *T = nullptr, *C = nullptr, *Ts = nullptr, *Cs = nullptr, *S = nullptr, *occ = nullptr;
    float *T0 = nullptr, *C0 = nullptr, *Tf = nullptr, *Cf = nullptr;
    unsigned int* count = nullptr;

    // network
    bf16* x0 = nullptr;           // [V][8]
    bf16* act[3] = {};            // hidden activations [V][32]
    // inference: the same network on the halo-padded layout (see Pad)
    Pad pad{};
    bf16* x0p = nullptr;          // [PZ][PY][PX][8]
    bf16* actp[2] = {};           // [PZ][PY][PX][32], ping-pong
    bf16* dz[2] = {};             // gradient ping-pong [V][32]
    float* pred = nullptr;        // [V][2]
    float* tgt = nullptr;         // [V][2]
    float* lpart = nullptr;
    double* dpart = nullptr;
    float* partial = nullptr;     // weight-gradient partials of one layer
    struct Layer {
        int cinp, nout;
        bf16* w = nullptr;
        bf16* wd = nullptr;
        float* bias = nullptr;
        float *master = nullptr, *m = nullptr, *v = nullptr;
        int params() const { return TAPS * CH * cinp + CH; }
    } L[NLAYER];
    int adam_t = 0;

    // gate
    int64_t calls = 0, fallbacks = 0, judged = 0;
    double last_judge_err = -1.0, last_budget_res = -1.0;
    int cooldown = 0;

    std::vector<void*> allocs;
};

namespace {

int32_t sfail(astro_surrogate* s, int32_t st, const std::string& what) {
    s->err = what;
    return st;
}

int32_t scheck(astro_surrogate* s, hipError_t e, const char* what) {
    if (e == hipSuccess) return ASTRO_OK;
    return sfail(s, e == hipErrorOutOfMemory ? ASTRO_OUT_OF_MEMORY : ASTRO_DEVICE_ERROR,
                 std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
int32_t salloc(astro_surrogate* s, T*& p, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return scheck(s, e, "surrogate allocation");
    s->allocs.push_back(q);
    p = static_cast<T*>(q);
    return hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), s->stream) == hipSuccess
               ? ASTRO_OK
               : sfail(s, ASTRO_DEVICE_ERROR, "memset");
}

Ftcs ftcs_of(const astro_surrogate* s, double D) {
    Ftcs f{};
    f.square2d = false;
    f.cx = static_cast<float>(D * s->d.dt_sub / (s->d.dx * s->d.dx));
    f.cy = static_cast<float>(D * s->d.dt_sub / (s->d.dy * s->d.dy));
    f.cz = static_cast<float>(D * s->d.dt_sub / (s->d.dz * s->d.dz));
    f.bc = ASTRO_BC_NEUMANN;
    return f;
}

// The plain path: steps_per_call x (FTCS of T and C, then the sources), on
// (T, C) in place (ping-pong through the scratch buffers).
void plain_advance(astro_surrogate* s, float*& T, float*& C, float*& Ts, float*& Cs) {
    const Ftcs ft = ftcs_of(s, s->d.d_temp), fc = ftcs_of(s, s->d.d_nutrient);
    PlainArgs a{};
    a.S = s->S;
    a.occ = s->occ;
    a.nx = s->d.nx;
    a.ny = s->d.ny;
    a.nz = s->d.nz;
    a.tcx = ft.cx;
    a.tcy = ft.cy;
    a.tcz = ft.cz;
    a.ccx = fc.cx;
    a.ccy = fc.cy;
    a.ccz = fc.cz;
    a.dt = static_cast<float>(s->d.dt_sub);
    a.du = a.dt * static_cast<float>(s->d.uptake);
    const unsigned g = plain_grid(a.nx, a.ny, a.nz);
    for (int k = 0; k < s->d.steps_per_call; ++k) {
        a.T = T;
        a.C = C;
        a.To = Ts;
        a.Co = Cs;
        k_plain_substep<<<g, 256, 0, s->stream>>>(a);
        std::swap(T, Ts);
        std::swap(C, Cs);
    }
}

PackArgs pack_args(const astro_surrogate* s, const float* T, const float* C) {
    PackArgs a{};
    a.T = T;
    a.C = C;
    a.S = s->S;
    a.occ = s->occ;
    a.x = s->x0;
    a.nx = s->d.nx;
    a.ny = s->d.ny;
    a.nz = s->d.nz;
    const double k = s->d.steps_per_call;
    const Ftcs ft = ftcs_of(s, s->d.d_temp), fc = ftcs_of(s, s->d.d_nutrient);
    a.ctx = static_cast<float>(k * ft.cx);
    a.cty = static_cast<float>(k * ft.cy);
    a.ctz = static_cast<float>(k * ft.cz);
    a.ccx = static_cast<float>(k * fc.cx);
    a.ccy = static_cast<float>(k * fc.cy);
    a.ccz = static_cast<float>(k * fc.cz);
    a.uptake_call = static_cast<float>(k * s->d.dt_sub * s->d.uptake);
    return a;
}

// The inputs of a call straight into the padded layout.
PackArgs pack_args_padded(const astro_surrogate* s, const float* T, const float* C) {
    PackArgs a = pack_args(s, T, C);
    a.x = s->x0p;
    a.px = s->pad.px;
    a.py = s->pad.py;
    return a;
}

// Inference from x0p: with `apply`, the increments go into (T, C); else the
// normalized outputs into pred.
void forward_padded(astro_surrogate* s, bool apply) {
    const int nx = s->d.nx, ny = s->d.ny, nz = s->d.nz;
    const unsigned g = padded_grid(nx, ny, nz);
    hipStream_t st = s->stream;
    const float lk = static_cast<float>(s->d.leak);
    const Pad p = s->pad;
    k_conv_in_p<<<g, 256, 0, st>>>(s->x0p, s->L[0].w, s->L[0].bias, s->actp[0], nx, ny, nz, p, lk);
    const unsigned gh = static_cast<unsigned>(
        (static_cast<int64_t>(nx / 32) * ((ny + HID_ROWS - 1) / HID_ROWS) * ((nz + ZPLANES - 1) / ZPLANES) + 3) / 4);
    k_conv_hid_z<HID_ROWS, ZPLANES><<<gh, 256, 0, st>>>(s->actp[0], s->L[1].w, s->L[1].bias, s->actp[1], nx, ny, nz,
                                                         p, lk);
    k_conv_hid_z<HID_ROWS, ZPLANES><<<gh, 256, 0, st>>>(s->actp[1], s->L[2].w, s->L[2].bias, s->actp[0], nx, ny, nz,
                                                         p, lk);
    const unsigned go = static_cast<unsigned>(
        (static_cast<int64_t>(nx / 32) * ((ny + OUT_ROWS - 1) / OUT_ROWS) * ((nz + ZPLANES - 1) / ZPLANES) + 3) / 4);
    k_conv_out_z<OUT_ROWS, ZPLANES><<<go, 256, 0, st>>>(s->actp[0], s->L[3].w, s->L[3].bias, s->pred, s->T, s->C, nx,
                                                         ny, nz, p, s->nm.dt_out, s->nm.dc_out, apply ? 1 : 0);
}

void forward(astro_surrogate* s) {
    const int nx = s->d.nx, ny = s->d.ny, nz = s->d.nz;
    const unsigned g = static_cast<unsigned>((static_cast<int64_t>(nx / 32) * ny * nz + 3) / 4);
    hipStream_t st = s->stream;
    const float lk = static_cast<float>(s->d.leak);
    k_conv<CIN0, 0><<<g, 256, 0, st>>>(s->x0, s->L[0].w, s->L[0].bias, nullptr, s->act[0], nullptr, nx, ny, nz, lk);
    const unsigned g2 = conv_rows_grid(nx, ny, nz);
    k_conv_rows<0, CONV_ROWS><<<g2, 256, 0, st>>>(s->act[0], s->L[1].w, s->L[1].bias, nullptr, s->act[1], nullptr,
                                                   nx, ny, nz, lk);
    k_conv_rows<0, CONV_ROWS><<<g2, 256, 0, st>>>(s->act[1], s->L[2].w, s->L[2].bias, nullptr, s->act[2], nullptr,
                                                   nx, ny, nz, lk);
    k_conv_rows<1, CONV_ROWS><<<g2, 256, 0, st>>>(s->act[2], s->L[3].w, s->L[3].bias, nullptr, nullptr, s->pred,
                                                   nx, ny, nz, lk);
}

constexpr float GRAD_SCALE = 1024.0f;   // applied to dLoss/dpred before BF16, removed in Adam

// Loss and the weight gradients of every layer into their partials; `apply`
// runs Adam after each layer's gradient. Returns the mean loss via the host.
int32_t backward(astro_surrogate* s, bool apply, float lr, std::vector<std::vector<float>>* grads_out) {
    const int nx = s->d.nx, ny = s->d.ny, nz = s->d.nz;
    const unsigned g = static_cast<unsigned>((static_cast<int64_t>(nx / 32) * ny * nz + 3) / 4);
    hipStream_t st = s->stream;
    k_loss<<<grid1(s->n), 256, 0, st>>>(s->pred, s->tgt, s->dz[0], s->lpart, s->n, GRAD_SCALE);
    const bf16* inputs[NLAYER] = {s->x0, s->act[0], s->act[1], s->act[2]};
    int cur = 0;
    if (apply) ++s->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    // mean over voxels and outputs
    const float gscale = 1.0f / (GRAD_SCALE * static_cast<float>(s->n) * NOUT);
    for (int l = NLAYER - 1; l >= 0; --l) {
        astro_surrogate::Layer& Ly = s->L[l];
        if (Ly.cinp == CH)
            k_conv_wgrad<CH><<<WG_BLOCKS, 256, wgrad_lds<CH>(nx), st>>>(inputs[l], s->dz[cur], s->partial, nx, ny, nz);
        else
            k_conv_wgrad<CIN0><<<WG_BLOCKS, 256, wgrad_lds<CIN0>(nx), st>>>(inputs[l], s->dz[cur], s->partial, nx, ny,
                                                                           nz);
        if (grads_out) {
            const int P = Ly.params();
            std::vector<float> part(static_cast<size_t>(P) * WG_BLOCKS);
            if (hipMemcpyAsync(part.data(), s->partial, sizeof(float) * part.size(), hipMemcpyDeviceToHost, st) !=
                    hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return sfail(s, ASTRO_DEVICE_ERROR, "gradient download");
            std::vector<float> gsum(P, 0.0f);
            for (int b = 0; b < WG_BLOCKS; ++b)
                for (int p = 0; p < P; ++p) gsum[p] += part[static_cast<size_t>(b) * P + p];
            for (float& x : gsum) x *= gscale;
            (*grads_out)[l] = gsum;
        }
        if (l > 0) {
            // dZ of layer l-1 = conv(dZ_l, flipped W_l) * leaky'(act[l-1])
            k_conv_rows<2, CONV_ROWS><<<conv_rows_grid(nx, ny, nz), 256, 0, st>>>(
                s->dz[cur], Ly.wd, nullptr, s->act[l - 1], s->dz[cur ^ 1], nullptr, nx, ny, nz,
                static_cast<float>(s->d.leak));
        }
        if (apply) {
            AdamArgs a{};
            a.partial = s->partial;
            a.nblocks = WG_BLOCKS;
            a.cinp = Ly.cinp;
            a.nout = Ly.nout;
            a.master = Ly.master;
            a.m = Ly.m;
            a.v = Ly.v;
            a.w = Ly.w;
            a.wd = Ly.wd;
            a.bias = Ly.bias;
            a.lr = lr;
            a.b1 = b1;
            a.b2 = b2;
            a.eps = 1e-8f;
            a.c1 = 1.0f - std::pow(b1, static_cast<float>(s->adam_t));
            a.c2 = 1.0f - std::pow(b2, static_cast<float>(s->adam_t));
            a.gscale = gscale;
            k_adam<<<grid1(Ly.params()), 256, 0, st>>>(a);
        }
        cur ^= 1;
    }
    return ASTRO_OK;
}

double mean_loss(astro_surrogate* s) {
    std::vector<float> lp(grid1(s->n));
    if (hipMemcpyAsync(lp.data(), s->lpart, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return -1.0;
    double sum = 0.0;
    for (float x : lp) sum += x;
    return sum / (static_cast<double>(s->n) * NOUT);
}

// Writes the master weights of layer l (BF16 copies included) from host.
int32_t upload_layer(astro_surrogate* s, int l, const float* w) {
    astro_surrogate::Layer& Ly = s->L[l];
    const int P = Ly.params();
    std::vector<float> m(w, w + P);
    for (int p = 0; p < P; ++p) {
        const bool is_bias = p >= TAPS * CH * Ly.cinp;
        const int o = is_bias ? p - TAPS * CH * Ly.cinp : (p / Ly.cinp) % CH;
        if (o >= Ly.nout) m[p] = 0.0f;
    }
    std::vector<bf16> wb(TAPS * CH * Ly.cinp), wdb(Ly.wd ? TAPS * CH * CH : 0, static_cast<bf16>(0.0f));
    for (int p = 0; p < TAPS * CH * Ly.cinp; ++p) {
        wb[p] = static_cast<bf16>(m[p]);
        if (Ly.wd) {
            const int t = p / (CH * Ly.cinp), o = (p / Ly.cinp) % CH, c = p % Ly.cinp;
            wdb[(static_cast<size_t>(TAPS - 1 - t) * CH + c) * CH + o] = wb[p];
        }
    }
    hipStream_t st = s->stream;
    int32_t rc = scheck(s, hipMemcpyAsync(Ly.master, m.data(), sizeof(float) * P, hipMemcpyHostToDevice, st), "w");
    if (rc == ASTRO_OK) rc = scheck(s, hipMemcpyAsync(Ly.w, wb.data(), sizeof(bf16) * wb.size(), hipMemcpyHostToDevice, st), "w");
    if (rc == ASTRO_OK && Ly.wd)
        rc = scheck(s, hipMemcpyAsync(Ly.wd, wdb.data(), sizeof(bf16) * wdb.size(), hipMemcpyHostToDevice, st), "w");
    if (rc == ASTRO_OK)
        rc = scheck(s, hipMemcpyAsync(Ly.bias, m.data() + TAPS * CH * Ly.cinp, sizeof(float) * CH,
                                      hipMemcpyHostToDevice, st), "w");
    if (rc == ASTRO_OK) rc = scheck(s, hipMemsetAsync(Ly.m, 0, sizeof(float) * P, st), "w");
    if (rc == ASTRO_OK) rc = scheck(s, hipMemsetAsync(Ly.v, 0, sizeof(float) * P, st), "w");
    if (rc == ASTRO_OK) rc = scheck(s, hipStreamSynchronize(st), "w");
    return rc;
}

// A random state of the training distribution into (T, C, S, occ).
void gen_state(astro_surrogate* s, uint64_t seed, float* T, float* C) {
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + 17);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double W = s->d.nx * s->d.dx, H = s->d.ny * s->d.dy, D = s->d.nz * s->d.dz;
    GenArgs g{};
    g.nblob = 2 + static_cast<int>(U(rng) * 4.0);
    if (g.nblob > MAX_BLOBS) g.nblob = MAX_BLOBS;
    for (int b = 0; b < g.nblob; ++b) {
        const double sig = (0.06 + 0.14 * U(rng)) * std::min(W, H);
        g.blob[b].cx = static_cast<float>(U(rng) * W);
        g.blob[b].cy = static_cast<float>(U(rng) * H);
        g.blob[b].cz = static_cast<float>((0.2 + 0.6 * U(rng)) * D);
        g.blob[b].inv2s2 = static_cast<float>(1.0 / (2.0 * sig * sig));
        g.blob[b].amp_t = static_cast<float>(-2.0 + 7.0 * U(rng));
        g.blob[b].amp_c = static_cast<float>(-0.4 + 0.8 * U(rng));
    }
    g.nx = s->d.nx;
    g.ny = s->d.ny;
    g.nz = s->d.nz;
    g.dx = static_cast<float>(s->d.dx);
    g.dy = static_cast<float>(s->d.dy);
    g.dz = static_cast<float>(s->d.dz);
    g.t_amb = static_cast<float>(s->d.temp_ref);
    g.c_amb = static_cast<float>(s->d.nutrient_ref);
    g.gcx = static_cast<float>((U(rng) - 0.5) * 0.6 / W);
    g.gcy = static_cast<float>((U(rng) - 0.5) * 0.6 / H);
    k_gen_fields<<<grid1(s->n), 256, 0, s->stream>>>(T, C, g);
    CellArgs c{};
    c.ncells = s->d.cells;
    c.nclusters = 1 + static_cast<int>(U(rng) * 4.0);
    if (c.nclusters > 4) c.nclusters = 4;
    for (int k = 0; k < c.nclusters; ++k) {
        c.ccx[k] = static_cast<float>((0.15 + 0.7 * U(rng)) * W);
        c.ccy[k] = static_cast<float>((0.15 + 0.7 * U(rng)) * H);
        c.ccz[k] = static_cast<float>((0.2 + 0.6 * U(rng)) * D);
        c.cr[k] = U(rng) < 0.25 ? 0.0f : static_cast<float>((0.04 + 0.12 * U(rng)) * std::min(W, H));
    }
    c.seed = rng();
    c.nx = s->d.nx;
    c.ny = s->d.ny;
    c.nz = s->d.nz;
    c.dx = g.dx;
    c.dy = g.dy;
    c.dz = g.dz;
    hipMemsetAsync(s->count, 0, sizeof(unsigned int) * s->n, s->stream);
    k_gen_cells<<<grid1(c.ncells), 256, 0, s->stream>>>(s->count, c);
    const float heat = static_cast<float>(U(rng)) * s->heat_max;
    k_counts<<<grid1(s->n), 256, 0, s->stream>>>(s->count, s->occ, s->S, s->n, heat);
}

// One training sample: a random state, its plain trajectory, the target.
void make_sample(astro_surrogate* s, uint64_t seed) {
    gen_state(s, seed, s->T0, s->C0);
    hipMemcpyAsync(s->T, s->T0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
    hipMemcpyAsync(s->C, s->C0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
    plain_advance(s, s->T, s->C, s->Ts, s->Cs);
    k_pack<<<grid1(s->n), 256, 0, s->stream>>>(pack_args(s, s->T0, s->C0), s->nm);
    k_target<<<grid1(s->n), 256, 0, s->stream>>>(s->T0, s->C0, s->T, s->C, s->tgt, s->n, s->nm);
}

// Relative L2 of increments (a - start) against (b - start) over the field.
double rel_increment_err(astro_surrogate* s, const float* a, const float* b, const float* start) {
    k_compare<<<grid1(s->n), 256, 0, s->stream>>>(a, b, start, s->dpart, s->n);
    std::vector<double> p(2 * grid1(s->n));
    if (hipMemcpyAsync(p.data(), s->dpart, sizeof(double) * p.size(), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return -1.0;
    double e = 0.0, r = 0.0;
    for (size_t k = 0; k < p.size(); k += 2) {
        e += p[k];
        r += p[k + 1];
    }
    return r > 0.0 ? std::sqrt(e / r) : std::sqrt(e);
}

double field_sum(astro_surrogate* s, const float* a) {
    k_sum<<<grid1(s->n), 256, 0, s->stream>>>(a, s->dpart, s->n);
    std::vector<double> p(grid1(s->n));
    if (hipMemcpyAsync(p.data(), s->dpart, sizeof(double) * p.size(), hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return 0.0;
    double sum = 0.0;
    for (double x : p) sum += x;
    return sum;
}

}  // namespace

extern "C" {

void astro_surrogate_desc_default(astro_surrogate_desc* d) {
    if (!d) return;
    *d = astro_surrogate_desc{};
    d->abi_version = ASTRO_ABI_VERSION;
    d->nx = d->ny = d->nz = 128;
    d->dx = d->dy = 550e-6 / 128.0;
    d->dz = 200e-6 / 128.0;
    d->d_temp = ASTRO_WATER_THERMAL_DIFFUSIVITY;
    d->d_nutrient = ASTRO_CO2_DIFFUSIVITY_WATER;
    d->dt_sub = 1e-4 / 16.0;
    d->steps_per_call = 16;
    d->uptake = 200.0;
    d->temp_ref = 369.0;
    d->nutrient_ref = 1e-3;
    d->cells = 25000;
    d->seed = 1;
    d->leak = 1.0;
    d->judge_every = 8;
    d->judge_tol = 2e-2;
    d->budget_tol = 2e-2;
    d->cooldown = 4;
}

int32_t astro_surrogate_create(const astro_surrogate_desc* din, astro_surrogate** out) {
    if (!din || !out) return ASTRO_INVALID_ARGUMENT;
    *out = nullptr;
    if (din->abi_version != ASTRO_ABI_VERSION) return ASTRO_CONTRACT_VERSION_MISMATCH;
    const astro_surrogate_desc& d = *din;
    if (d.nx < 32 || d.nx % 32 != 0 || d.ny < 2 || d.nz < 2 || d.steps_per_call < 1 || !(d.dt_sub > 0.0) ||
        !(d.dx > 0 && d.dy > 0 && d.dz > 0) || d.cells < 0 || !(d.leak >= 0.0 && d.leak <= 1.0))
        return ASTRO_INVALID_ARGUMENT;
    if (static_cast<int64_t>(d.nx) * d.ny * d.nz >= (1ll << 31)) return ASTRO_INVALID_ARGUMENT;
    // the FTCS substep must be stable for both fields
    const double inv = 1.0 / (d.dx * d.dx) + 1.0 / (d.dy * d.dy) + 1.0 / (d.dz * d.dz);
    if (d.d_temp * d.dt_sub * inv > 0.5 || d.d_nutrient * d.dt_sub * inv > 0.5) return ASTRO_INVALID_ARGUMENT;
    auto s = std::make_unique<astro_surrogate>();
    s->d = d;
    s->n = static_cast<int64_t>(d.nx) * d.ny * d.nz;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return ASTRO_DEVICE_ERROR;
    const int64_t n = s->n;
    int32_t st = ASTRO_OK;
    for (float** p : {&s->T, &s->C, &s->Ts, &s->Cs, &s->S, &s->occ, &s->T0, &s->C0, &s->Tf, &s->Cf})
        if (st == ASTRO_OK) st = salloc(s.get(), *p, n);
    if (st == ASTRO_OK) st = salloc(s.get(), s->count, n);
    if (st == ASTRO_OK) st = salloc(s.get(), s->x0, n * CIN0);
    for (int k = 0; k < 3 && st == ASTRO_OK; ++k) st = salloc(s.get(), s->act[k], n * CH);
    // padded buffers: zeroed here, and only their interiors are ever written
    s->pad = pad_of(d.nx, d.ny, d.nz);
    const int64_t np = static_cast<int64_t>(s->pad.px) * s->pad.py * s->pad.pz;
    if (st == ASTRO_OK) st = salloc(s.get(), s->x0p, np * CIN0);
    for (int k = 0; k < 2 && st == ASTRO_OK; ++k) st = salloc(s.get(), s->actp[k], np * CH);
    for (int k = 0; k < 2 && st == ASTRO_OK; ++k) st = salloc(s.get(), s->dz[k], n * CH);
    if (st == ASTRO_OK) st = salloc(s.get(), s->pred, n * NOUT);
    if (st == ASTRO_OK) st = salloc(s.get(), s->tgt, n * NOUT);
    if (st == ASTRO_OK) st = salloc(s.get(), s->lpart, grid1(n));
    if (st == ASTRO_OK) st = salloc(s.get(), s->dpart, 2 * grid1(n));
    if (st == ASTRO_OK) st = salloc(s.get(), s->partial, static_cast<size_t>(TAPS * CH * CH + CH) * WG_BLOCKS);
    const int cinp[NLAYER] = {CIN0, CH, CH, CH}, nout[NLAYER] = {CH, CH, CH, NOUT};
    for (int l = 0; l < NLAYER && st == ASTRO_OK; ++l) {
        astro_surrogate::Layer& Ly = s->L[l];
        Ly.cinp = cinp[l];
        Ly.nout = nout[l];
        const int P = Ly.params();
        st = salloc(s.get(), Ly.w, TAPS * CH * Ly.cinp);
        if (st == ASTRO_OK && l > 0) st = salloc(s.get(), Ly.wd, TAPS * CH * CH);
        if (st == ASTRO_OK) st = salloc(s.get(), Ly.bias, CH);
        if (st == ASTRO_OK) st = salloc(s.get(), Ly.master, P);
        if (st == ASTRO_OK) st = salloc(s.get(), Ly.m, P);
        if (st == ASTRO_OK) st = salloc(s.get(), Ly.v, P);
    }
    if (st != ASTRO_OK) {
        astro_surrogate_destroy(s.release());
        return st;
    }
    // He-style initialization from the seed; the output layer starts small.
    std::mt19937_64 rng(d.seed);
    std::normal_distribution<double> N01(0.0, 1.0);
    for (int l = 0; l < NLAYER && st == ASTRO_OK; ++l) {
        astro_surrogate::Layer& Ly = s->L[l];
        const int fan_in = TAPS * (l == 0 ? CIN0 : CH);
        const double sd = (l == NLAYER - 1 ? 0.1 : 1.0) * std::sqrt(2.0 / fan_in);
        std::vector<float> w(Ly.params(), 0.0f);
        for (int p = 0; p < TAPS * CH * Ly.cinp; ++p) {
            w[p] = static_cast<float>(sd * N01(rng));
        }
        st = upload_layer(s.get(), l, w.data());
    }
    // normalization: per-cell heating such that a dense voxel warms ~1 K per
    // call; output scales from the plain increments of a few random states
    s->heat_max = static_cast<float>(1.0 / (d.steps_per_call * d.dt_sub) / 4.0);
    s->nm.t_ref = static_cast<float>(d.temp_ref);
    s->nm.t_amp = 5.0f;
    s->nm.c_ref = static_cast<float>(d.nutrient_ref);
    s->nm.c_amp = static_cast<float>(d.nutrient_ref);
    s->nm.dt_call = static_cast<float>(d.dt_sub * d.steps_per_call);
    s->nm.occ_amp = 4.0f;
    s->nm.dt_out = 1.0f;
    s->nm.dc_out = 1.0f;
    if (st == ASTRO_OK && d.cells >= 0) {
        double et = 0.0, ec = 0.0;
        const int probes = 4;
        for (int k = 0; k < probes; ++k) {
            make_sample(s.get(), d.seed * 7919 + 1000003ull * (k + 1));
            std::vector<float> tg(static_cast<size_t>(n) * NOUT);
            if (hipMemcpyAsync(tg.data(), s->tgt, sizeof(float) * tg.size(), hipMemcpyDeviceToHost, s->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(s->stream) != hipSuccess) {
                st = sfail(s.get(), ASTRO_DEVICE_ERROR, "scale probe");
                break;
            }
            for (int64_t v = 0; v < n; ++v) {
                et += static_cast<double>(tg[v * NOUT]) * tg[v * NOUT];
                ec += static_cast<double>(tg[v * NOUT + 1]) * tg[v * NOUT + 1];
            }
        }
        const double rt = std::sqrt(et / (probes * n)), rc = std::sqrt(ec / (probes * n));
        s->nm.dt_out = static_cast<float>(rt > 0 ? rt : 1.0);
        s->nm.dc_out = static_cast<float>(rc > 0 ? rc : 1.0);
    }
    if (st != ASTRO_OK) {
        astro_surrogate_destroy(s.release());
        return st;
    }
    *out = s.release();
    return ASTRO_OK;
}

void astro_surrogate_destroy(astro_surrogate* s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    for (void* p : s->allocs) hipFree(p);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

const char* astro_surrogate_last_error(const astro_surrogate* s) { return s ? s->err.c_str() : ""; }

int32_t astro_surrogate_train(astro_surrogate* s, int32_t trajectories, int32_t epochs, float lr, uint64_t seed,
                              double* loss_out) {
    if (!s || trajectories <= 0 || epochs <= 0 || !(lr > 0.0f)) return ASTRO_INVALID_ARGUMENT;
    double last = 0.0;
    int counted = 0;
    for (int e = 0; e < epochs; ++e) {
        double sum = 0.0;
        for (int k = 0; k < trajectories; ++k) {
            // a fixed set of trajectories (seed, k); cosine-decayed rate
            make_sample(s, seed * 1000003ull + static_cast<uint64_t>(k));
            forward(s);
            const float prog = (e * trajectories + k) / static_cast<float>(epochs * trajectories);
            const float rate = lr * 0.5f * (1.0f + std::cos(3.14159265f * prog));
            const int32_t st = backward(s, true, rate, nullptr);
            if (st != ASTRO_OK) return st;
            if (e == epochs - 1) {
                sum += mean_loss(s);
                ++counted;
            }
        }
        last = counted ? sum / counted : 0.0;
    }
    if (loss_out) *loss_out = last;
    return scheck(s, hipStreamSynchronize(s->stream), "train");
}

int32_t astro_surrogate_sample(astro_surrogate* s, uint64_t seed) {
    if (!s) return ASTRO_INVALID_ARGUMENT;
    gen_state(s, seed, s->T, s->C);
    return scheck(s, hipStreamSynchronize(s->stream), "sample");
}

int32_t astro_surrogate_set_state(astro_surrogate* s, const float* T, const float* C, const float* src,
                                  const float* occ, size_t n) {
    if (!s || n != static_cast<size_t>(s->n)) return ASTRO_INVALID_ARGUMENT;
    const size_t b = sizeof(float) * n;
    int32_t st = ASTRO_OK;
    if (T) st = scheck(s, hipMemcpy(s->T, T, b, hipMemcpyHostToDevice), "upload");
    if (st == ASTRO_OK && C) st = scheck(s, hipMemcpy(s->C, C, b, hipMemcpyHostToDevice), "upload");
    if (st == ASTRO_OK && src) st = scheck(s, hipMemcpy(s->S, src, b, hipMemcpyHostToDevice), "upload");
    if (st == ASTRO_OK && occ) st = scheck(s, hipMemcpy(s->occ, occ, b, hipMemcpyHostToDevice), "upload");
    return st;
}

int32_t astro_surrogate_get_state(astro_surrogate* s, float* T, float* C, float* src, float* occ, size_t n) {
    if (!s || n != static_cast<size_t>(s->n)) return ASTRO_INVALID_ARGUMENT;
    int32_t st = scheck(s, hipStreamSynchronize(s->stream), "synchronize");
    const size_t b = sizeof(float) * n;
    if (st == ASTRO_OK && T) st = scheck(s, hipMemcpy(T, s->T, b, hipMemcpyDeviceToHost), "download");
    if (st == ASTRO_OK && C) st = scheck(s, hipMemcpy(C, s->C, b, hipMemcpyDeviceToHost), "download");
    if (st == ASTRO_OK && src) st = scheck(s, hipMemcpy(src, s->S, b, hipMemcpyDeviceToHost), "download");
    if (st == ASTRO_OK && occ) st = scheck(s, hipMemcpy(occ, s->occ, b, hipMemcpyDeviceToHost), "download");
    return st;
}

int32_t astro_surrogate_advance(astro_surrogate* s, int32_t mode, int32_t* used_surrogate) {
    if (!s || mode < ASTRO_SURROGATE_PLAIN || mode > ASTRO_SURROGATE_GATED) return ASTRO_INVALID_ARGUMENT;
    ++s->calls;
    bool use_net = mode != ASTRO_SURROGATE_PLAIN;
    if (mode == ASTRO_SURROGATE_GATED && s->cooldown > 0) {
        --s->cooldown;
        use_net = false;
    }
    if (!use_net) {
        plain_advance(s, s->T, s->C, s->Ts, s->Cs);
        if (mode == ASTRO_SURROGATE_GATED) ++s->fallbacks;
        if (used_surrogate) *used_surrogate = 0;
        return scheck(s, hipGetLastError(), "advance");
    }
    const bool judge = mode == ASTRO_SURROGATE_GATED && s->d.judge_every > 0 && (s->calls % s->d.judge_every) == 1 % s->d.judge_every;
    if (mode == ASTRO_SURROGATE_GATED) {
        hipMemcpyAsync(s->T0, s->T, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
        hipMemcpyAsync(s->C0, s->C, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
    }
    k_pack<<<grid1(s->n), 256, 0, s->stream>>>(pack_args_padded(s, s->T, s->C), s->nm);
    forward_padded(s, true);
    if (used_surrogate) *used_surrogate = 1;
    if (mode != ASTRO_SURROGATE_GATED) return scheck(s, hipGetLastError(), "advance");
    // Heat budget: with Neumann walls the plain path changes the total heat
    // by exactly dt_call * sum(S); the net's change must agree.
    const double want = static_cast<double>(s->nm.dt_call) * field_sum(s, s->S);
    const double got = field_sum(s, s->T) - field_sum(s, s->T0);
    const double scale = std::max(std::fabs(want), 1e-6 * static_cast<double>(s->n) * s->nm.dt_out);
    s->last_budget_res = std::fabs(got - want) / scale;
    bool reject = !(s->last_budget_res <= s->d.budget_tol);   // NaN rejects
    if (judge && !reject) {
        // the plain path from the same start judges the call
        hipMemcpyAsync(s->Tf, s->T0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
        hipMemcpyAsync(s->Cf, s->C0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
        float *tf = s->Tf, *cf = s->Cf, *ts = s->Ts, *cs = s->Cs;
        plain_advance(s, tf, cf, ts, cs);
        ++s->judged;
        s->last_judge_err = rel_increment_err(s, s->T, tf, s->T0);
        if (!(s->last_judge_err <= s->d.judge_tol)) {   // NaN rejects
            reject = true;
            // keep the plain result already computed
            hipMemcpyAsync(s->T, tf, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
            hipMemcpyAsync(s->C, cf, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
            s->cooldown = s->d.cooldown;
            ++s->fallbacks;
            if (used_surrogate) *used_surrogate = 0;
            return scheck(s, hipGetLastError(), "advance");
        }
    }
    if (reject) {
        hipMemcpyAsync(s->T, s->T0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
        hipMemcpyAsync(s->C, s->C0, sizeof(float) * s->n, hipMemcpyDeviceToDevice, s->stream);
        plain_advance(s, s->T, s->C, s->Ts, s->Cs);
        s->cooldown = s->d.cooldown;
        ++s->fallbacks;
        if (used_surrogate) *used_surrogate = 0;
    }
    return scheck(s, hipGetLastError(), "advance");
}

int32_t astro_surrogate_counters_get(const astro_surrogate* s, astro_surrogate_counters* c) {
    if (!s || !c) return ASTRO_INVALID_ARGUMENT;
    c->calls = s->calls;
    c->fallbacks = s->fallbacks;
    c->judged = s->judged;
    c->last_judge_err = s->last_judge_err;
    c->last_budget_residual = s->last_budget_res;
    c->out_scale_temp = s->nm.dt_out;
    c->out_scale_nutrient = s->nm.dc_out;
    return ASTRO_OK;
}

int32_t astro_surrogate_weights(astro_surrogate* s, int32_t layer, float* w, size_t n, int32_t write) {
    if (!s || layer < 0 || layer >= NLAYER || !w) return ASTRO_INVALID_ARGUMENT;
    const astro_surrogate::Layer& Ly = s->L[layer];
    if (n != static_cast<size_t>(Ly.params())) return ASTRO_INVALID_ARGUMENT;
    if (write) return upload_layer(s, layer, w);
    int32_t st = scheck(s, hipStreamSynchronize(s->stream), "synchronize");
    if (st == ASTRO_OK) st = scheck(s, hipMemcpy(w, Ly.master, sizeof(float) * n, hipMemcpyDeviceToHost), "weights");
    return st;
}

// The network on given normalized inputs (BF16-rounded by the caller or
// not): x [V][8] in, pred [V][2] out; with `target`, also the loss gradient
// of every layer's master weights (grads: layer-major, params each).
int32_t astro_surrogate_eval(astro_surrogate* s, const float* x, size_t n, const float* target, float* pred,
                             float* grads, size_t ngrads) {
    if (!s || !x || n != static_cast<size_t>(s->n)) return ASTRO_INVALID_ARGUMENT;
    size_t need = 0;
    for (int l = 0; l < NLAYER; ++l) need += s->L[l].params();
    if (grads && ngrads != need) return ASTRO_INVALID_ARGUMENT;
    std::vector<bf16> xb(n * CIN0);
    for (size_t k = 0; k < n * CIN0; ++k) xb[k] = static_cast<bf16>(x[k]);
    int32_t st = scheck(s, hipMemcpy(s->x0, xb.data(), sizeof(bf16) * xb.size(), hipMemcpyHostToDevice), "upload");
    if (st != ASTRO_OK) return st;
    // predictions alone come from the inference kernels; with a target, from
    // the training kernels whose activations the gradients need
    if (target) {
        forward(s);
    } else {
        k_pad_input<<<grid1(s->n), 256, 0, s->stream>>>(s->x0, s->x0p, s->d.nx, s->d.ny, s->d.nz, s->pad);
        forward_padded(s, false);
    }
    if (target) {
        if ((st = scheck(s, hipMemcpyAsync(s->tgt, target, sizeof(float) * n * NOUT, hipMemcpyHostToDevice, s->stream),
                         "upload")) != ASTRO_OK)
            return st;
        std::vector<std::vector<float>> g(NLAYER);
        if ((st = backward(s, false, 0.0f, grads ? &g : nullptr)) != ASTRO_OK) return st;
        if (grads) {
            size_t off = 0;
            for (int l = 0; l < NLAYER; ++l) {
                std::copy(g[l].begin(), g[l].end(), grads + off);
                off += g[l].size();
            }
        }
    }
    st = scheck(s, hipStreamSynchronize(s->stream), "eval");
    if (st == ASTRO_OK && pred)
        st = scheck(s, hipMemcpy(pred, s->pred, sizeof(float) * n * NOUT, hipMemcpyDeviceToHost), "download");
    return st;
}

void* astro_surrogate_stream(astro_surrogate* s) { return s ? static_cast<void*>(s->stream) : nullptr; }

}  // extern "C"
