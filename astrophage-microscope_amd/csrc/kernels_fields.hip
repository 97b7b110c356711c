// kernels_fields.hip -- FTCS diffusion, the light sweep and occlusion stamp,
// and the field half of the telemetry reduction.
//
// FTCS (src/fields/diffuse.cu:19-44): one explicit ping-pong substep. The
// square depth-averaged grid keeps the reference's 5-point expression so it
// rounds identically; a grid with nz > 1 uses the 7-point form
//   u' = u + cx((l+r)-2u) + cy((d+u)-2u) + cz((b+f)-2u).
// Both are pure HBM streams (4 B read + 4 B written per point and substep):
// blocks are 64 x 4 threads so each wave reads one contiguous 256 B row, the
// 3D kernel marches along z keeping the column in registers, and tiles are
// remapped so each XCD streams one contiguous slab (neighbour rows hit L2).
#include "kernels.h"

namespace astro_dev {

namespace {

__device__ __forceinline__ float ghost_of(float c, int bc, float amb, float robin) {
    if (bc == ASTRO_BC_DIRICHLET) return amb;
    if (bc == ASTRO_BC_ROBIN) return c - robin * (c - amb);
    return c;
}

struct KFtcs2d {
    struct Args {
        const float* __restrict__ src;
        float* __restrict__ dst;
        int n;
        float coeff;
        int bc;
        float amb;
        float robin;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [src, dst, n, coeff, bc, amb, robin] = a;

    const int tx = (n + 63) / 64, ty = (n + 3) / 4;
    const int t = xcd_tile(vg.block, vg.grid);
    if (t >= tx * ty) return;
    const int i = (t % tx) * 64 + (threadIdx.x & 63);
    const int j = (t / tx) * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const size_t id = static_cast<size_t>(j) * n + i;
    const float c = src[id];
    const float l = (i > 0) ? src[id - 1] : ghost_of(c, bc, amb, robin);
    const float r = (i < n - 1) ? src[id + 1] : ghost_of(c, bc, amb, robin);
    const float d = (j > 0) ? src[id - n] : ghost_of(c, bc, amb, robin);
    const float u = (j < n - 1) ? src[id + n] : ghost_of(c, bc, amb, robin);
    dst[id] = c + coeff * (l + r + d + u - 4.0f * c);
    }
};

struct KFtcs3d {
    struct Args {
        const float* __restrict__ src;
        float* __restrict__ dst;
        int nx;
        int ny;
        int nz;
        int zc;
        float cx;
        float cy;
        float cz;
        int bc;
        float amb;
        float rx;
        float ry;
        float rz;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [src, dst, nx, ny, nz, zc, cx, cy, cz, bc, amb, rx, ry, rz] = a;

    const int tx = (nx + 63) / 64, ty = (ny + 3) / 4, tz = (nz + zc - 1) / zc;
    const int t = xcd_tile(vg.block, vg.grid);
    if (t >= tx * ty * tz) return;
    const int i = (t % tx) * 64 + (threadIdx.x & 63);
    const int j = ((t / tx) % ty) * 4 + (threadIdx.x >> 6);
    const int k0 = (t / (tx * ty)) * zc;
    if (i >= nx || j >= ny) return;
    const int k1 = k0 + zc < nz ? k0 + zc : nz;
    const size_t pl = static_cast<size_t>(nx) * ny;
    const size_t col = static_cast<size_t>(j) * nx + i;
    float below = (k0 > 0) ? src[(k0 - 1) * pl + col] : 0.0f;
    float cur = src[k0 * pl + col];
    for (int k = k0; k < k1; ++k) {
        const size_t id = k * pl + col;
        const float above = (k < nz - 1) ? src[id + pl] : 0.0f;
        const float c = cur;
        const float l = (i > 0) ? src[id - 1] : ghost_of(c, bc, amb, rx);
        const float r = (i < nx - 1) ? src[id + 1] : ghost_of(c, bc, amb, rx);
        const float d = (j > 0) ? src[id - nx] : ghost_of(c, bc, amb, ry);
        const float u = (j < ny - 1) ? src[id + nx] : ghost_of(c, bc, amb, ry);
        float acc = cx * ((l + r) - 2.0f * c) + cy * ((d + u) - 2.0f * c);
        if (nz > 1) {
            const float b = (k > 0) ? below : ghost_of(c, bc, amb, rz);
            const float f = (k < nz - 1) ? above : ghost_of(c, bc, amb, rz);
            acc = acc + cz * ((b + f) - 2.0f * c);
        }
        dst[id] = c + acc;
        below = cur;
        cur = above;
    }
    }
};

// The same 7-point substep, vectorised: a lane owns 4 consecutive x points
// (16-byte loads and stores), a wave one 256-point row segment, a block 4
// rows, and each block ZC planes. All of a block's inputs -- its ZC + 2
// planes of its rows, the rows above and below, and the row-end neighbours --
// are requested before any is used, so a wave keeps ~14 independent loads in
// flight; in-row neighbours come from the adjacent lanes' registers. Needs
// nx % 4 == 0; per point the arithmetic is exactly k_ftcs3d's.
__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

constexpr int FTCS_ZC = 4;

// W lanes per x row (4 points each): a wave covers 64 / W rows, a block
// 4 x 64 / W rows, so narrow grids keep every lane busy.
struct Ftcs3Args {
    const float* __restrict__ src;
    float* __restrict__ dst;
    int nx;
    int ny;
    int nz;
    float cx;
    float cy;
    float cz;
    int bc;
    float amb;
    float rx;
    float ry;
    float rz;
};

template <int W>
struct KFtcs3dV4 {
    using Args = Ftcs3Args;
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [src, dst, nx, ny, nz, cx, cy, cz, bc, amb, rx, ry, rz] = a;

    constexpr int ZC = FTCS_ZC;
    constexpr int RB = 4 * (64 / W);   // rows per block
    const int tx = (nx + 4 * W - 1) / (4 * W), ty = (ny + RB - 1) / RB, tz = (nz + ZC - 1) / ZC;
    const int t = xcd_tile(vg.block, vg.grid);
    if (t >= tx * ty * tz) return;   // whole block: shuffles below stay uniform
    const int lane = threadIdx.x & 63;
    const int xl = lane % W;           // lane within its row
    // z chunks vary fastest, then x, then y: the chunks of one column are
    // consecutive tiles of one XCD, so each chunk's halo planes are the
    // planes its neighbours just read through the same L2.
    const int i0 = ((t / tz) % tx) * (4 * W) + xl * 4;
    const int j = (t / (tz * tx)) * RB + static_cast<int>(threadIdx.x) / W;
    const int k0 = (t % tz) * ZC;
    const bool on = i0 < nx && j < ny;
    const size_t pl = static_cast<size_t>(nx) * ny;
    const size_t col = on ? static_cast<size_t>(j) * nx + i0 : 0;
    const float4 zero4{0.0f, 0.0f, 0.0f, 0.0f};
    float4 planes[ZC + 2], dn4[ZC], up4[ZC];
    float lf[ZC], rg[ZC];
#pragma unroll
    for (int m = 0; m < ZC + 2; ++m) {
        const int k = k0 - 1 + m;
        planes[m] = (on && k >= 0 && k < nz) ? ld4(src + k * pl + col) : zero4;
    }
#pragma unroll
    for (int m = 0; m < ZC; ++m) {
        const int k = k0 + m;
        const bool ok = on && k < nz;
        const size_t id = k * pl + col;
        dn4[m] = (ok && j > 0) ? ld4(src + id - nx) : zero4;
        up4[m] = (ok && j < ny - 1) ? ld4(src + id + nx) : zero4;
        lf[m] = (ok && xl == 0 && i0 > 0) ? src[id - 1] : 0.0f;
        rg[m] = (ok && (xl == W - 1 || i0 + 4 >= nx) && i0 + 4 < nx) ? src[id + 4] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < ZC; ++m) {
        const int k = k0 + m;
        const float4 cur = planes[m + 1], below = planes[m], above = planes[m + 2];
        // x neighbours of the lane's first and last point
        float lft = __shfl_up(cur.w, 1, 64);
        float rgt = __shfl_down(cur.x, 1, 64);
        if (xl == 0) lft = lf[m];
        if (xl == W - 1 || i0 + 4 >= nx) rgt = rg[m];
        if (!on || k >= nz) continue;
        const float c[4] = {cur.x, cur.y, cur.z, cur.w};
        const float dd[4] = {dn4[m].x, dn4[m].y, dn4[m].z, dn4[m].w};
        const float uu[4] = {up4[m].x, up4[m].y, up4[m].z, up4[m].w};
        const float bb[4] = {below.x, below.y, below.z, below.w}, ff[4] = {above.x, above.y, above.z, above.w};
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float cc = c[e];
            const float l = e > 0 ? c[e - 1] : (i0 > 0 ? lft : ghost_of(cc, bc, amb, rx));
            const float r = e < 3 ? c[e + 1] : (i0 + 4 < nx ? rgt : ghost_of(cc, bc, amb, rx));
            const float dn = j > 0 ? dd[e] : ghost_of(cc, bc, amb, ry);
            const float up = j < ny - 1 ? uu[e] : ghost_of(cc, bc, amb, ry);
            float acc = cx * ((l + r) - 2.0f * cc) + cy * ((dn + up) - 2.0f * cc);
            if (nz > 1) {
                const float b = k > 0 ? bb[e] : ghost_of(cc, bc, amb, rz);
                const float f = k < nz - 1 ? ff[e] : ghost_of(cc, bc, amb, rz);
                acc = acc + cz * ((b + f) - 2.0f * cc);
            }
            o[e] = cc + acc;
        }
        *reinterpret_cast<float4*>(dst + k * pl + col) = float4{o[0], o[1], o[2], o[3]};
    }
    }
};

// occlusion_stamp_kernel (src/sim/emission.cu:17-27).
template <class P>
struct KStamp {
    static constexpr bool stride = true;   // any block count covers the work
    struct Args {
        Cells<P> c;
        const Counters* ctr;
        Grid irr;
    };
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [c, ctr, irr] = a;

    const int n = ctr->nphys;
    const auto body = [&](const int i) {
    const uint32_t f = c.flags[i];
    if (!(f & ASTRO_FLAG_OCCUPIED)) return;
    const double area = (f & ASTRO_FLAG_ALIVE) ? ASTRO_CELL_CROSS_SECTION
                                               : ASTRO_CELL_CROSS_SECTION * (1.0 - ASTRO_CELL_ALBEDO_DEAD);
    const uint32_t idx = grid_nearest(irr, static_cast<float>(c.x[i]), static_cast<float>(c.y[i]),
                                      static_cast<float>(c.z[i]));
    deposit_add(&irr.deposit[idx], area, irr.scale);
    };
    for (int i = vg.block * BLOCK + threadIdx.x; i < n; i += vg.grid * BLOCK) body(i);
    }
};

// sweep_kernel (src/sim/emission.cu:37-53): one sequential Beer-Lambert march
// per line, the product kept in the reference's order so transmittances (and
// an exact 0) match bit for bit. The occlusion accumulator is consumed here:
// nonzero entries are cleared as they are read, which replaces the reference's
// dense clear_kernel at the start of the next tick.
struct SweepArgs {
    unsigned long long* blocked;
    float* out;
    int nx;
    int ny;
    int nz;
    int axis;
    int sign;
    double scale;
    double face;
    float source;
    float ambient;
};

constexpr int SWEEP_PRELOAD = 16;   // deposits a column walker loads ahead

struct KSweep {
    using Args = SweepArgs;
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [blocked, out, nx, ny, nz, axis, sign, scale, face, source, ambient] = a;

    const int other = axis == 0 ? ny : nx;
    const int len = axis == 0 ? nx : ny;
    const int line = vg.block * BLOCK + threadIdx.x;
    if (line >= other * nz) return;
    const int k = line / other, l = line % other;
    const auto index = [&](int s) -> size_t {
        const int step = sign > 0 ? s : (len - 1 - s);
        return axis == 0 ? (static_cast<size_t>(k) * ny + l) * nx + step
                         : (static_cast<size_t>(k) * ny + step) * nx + l;
    };
    // Each chunk's deposits are loaded before any is cleared, so the loads
    // are in flight together instead of one per step.
    double t = 1.0;
    for (int s0 = 0; s0 < len; s0 += SWEEP_PRELOAD) {
        unsigned long long raw[SWEEP_PRELOAD];
#pragma unroll
        for (int q = 0; q < SWEEP_PRELOAD; ++q) raw[q] = s0 + q < len ? blocked[index(s0 + q)] : 0ull;
#pragma unroll
        for (int q = 0; q < SWEEP_PRELOAD; ++q) {
            if (s0 + q >= len) break;
            const size_t idx = index(s0 + q);
            out[idx] = static_cast<float>(static_cast<double>(source) * t) + ambient;
            if (raw[q] != 0ull) blocked[idx] = 0ull;
            t *= column_transmittance(from_fixed(raw[q], scale), face);
        }
    }
    }
};

// The march for lines along x (axis 0), spread over lanes. A voxel without
// a deposit has the factor column_transmittance(0) = 1 - 0 / face = 1.0
// exactly, and multiplying by 1.0 changes nothing, so the running product at
// each point is the reference's sequential product of the nonzero-deposit
// factors before it, in march order. Each line gets `seg` adjacent lanes
// (64 / seg lines per wave), each lane four consecutive steps of every chunk
// of 4 * seg steps:
//   * the lanes load their deposits coalesced, clear the nonzero ones and
//     form their four factors;
//   * within each line, the lanes holding a nonzero factor are visited in
//     march order: each applies its four factors to the running product in
//     order (t *= f, as the reference), and the lanes before it up to the
//     previous such lane take the product as it stood, untouched by factors;
//   * every lane then writes its four points from the product entering its
//     first step, multiplying through its own factors in order.
// Every product is formed by the same multiplications, in the same order, as
// the one-lane march (its extra multiplications are by 1.0), so transmittances
// and irradiances are the same bit for bit; the march costs one visit per
// lane that holds a deposit instead of one step per point.
struct SweepLanesArgs {
    SweepArgs s;
    int seg;   // lanes per line: a power of two, 4 * seg >= nx or seg == 64
};

struct KSweepLanes {
    using Args = SweepLanesArgs;
    static __device__ void run(const Args& a, VGrid vg) {
        const auto& [blocked, out, nx, ny, nz, axis, sign, scale, face, source, ambient] = a.s;
        const int seg = a.seg;
        const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
        const int lines = ny * nz;
        const int first = (vg.block * (BLOCK / 64) + wave) * (64 / seg);
        if (first >= lines) return;   // whole wave
        const int j = lane & (seg - 1), sbase = lane - j;
        const int line = first + lane / seg;
        const bool on = line < lines;
        const unsigned long long segmask = seg == 64 ? ~0ull : ((1ull << seg) - 1ull) << sbase;
        const double src = static_cast<double>(source);
        double tline = 1.0;   // the line's running product, the same in all its lanes
        for (int c0 = 0; c0 < nx; c0 += 4 * seg) {
            unsigned long long raw[4];
            size_t idx[4];
            bool ok[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int pos = c0 + 4 * j + q;
                ok[q] = on && pos < nx;
                idx[q] = static_cast<size_t>(line) * nx + static_cast<size_t>(sign > 0 ? pos : nx - 1 - pos);
                raw[q] = ok[q] ? blocked[idx[q]] : 0ull;
            }
            double f[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ok[q] && raw[q] != 0ull) blocked[idx[q]] = 0ull;
                f[q] = column_transmittance(from_fixed(raw[q], scale), face);
            }
            const bool holds = (raw[0] | raw[1] | raw[2] | raw[3]) != 0ull;
            unsigned long long pending = __ballot(holds) & segmask;
            int prev = sbase - 1;   // last lane of the line already visited
            double tin = tline;     // the product entering this lane's first step
            while (__ballot(pending != 0ull) != 0ull) {
                const int b = pending ? __ffsll(static_cast<long long>(pending)) - 1 : lane;
                const double chain = ((tline * f[0]) * f[1] * f[2]) * f[3];
                const double next = __shfl(chain, b, 64);
                if (pending) {
                    if (lane > prev && lane <= b) tin = tline;
                    tline = next;
                    prev = b;
                    pending &= pending - 1ull;
                }
            }
            if (lane > prev) tin = tline;
            double t = tin;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ok[q]) out[idx[q]] = static_cast<float>(src * t) + ambient;
                t *= f[q];
            }
        }
    }
};

__global__ __launch_bounds__(BLOCK) void k_fill(float* v, size_t n, float value) {
    const size_t i = static_cast<size_t>(blockIdx.x) * BLOCK + threadIdx.x;
    if (i < n) v[i] = value;
}

// grid_brush's kernel (src/fields/diffuse.cu:62-77): value += amount * (1 - t^2)^2
// with t = r / radius inside the disc around (cx, cy), r measured from the
// voxel centre in the plane; a 3D grid takes the same amount at every depth.
// Float arithmetic in the reference's order. sqrtf and the divide are
// correctly rounded (hipcc's default -fhip-fp32-correctly-rounded-divide-sqrt,
// as CUDA's defaults); __fsqrt_rn would not be: without OCML's rounded
// operations it is the native approximation.
__global__ __launch_bounds__(BLOCK) void k_brush(float* v, int nx, int ny, size_t n, float dx, float dy, float cx,
                                                 float cy, float radius, float amount) {
    const float halfx = 0.5f * static_cast<float>(nx) * dx;
    const float halfy = 0.5f * static_cast<float>(ny) * dy;
    const size_t stride = static_cast<size_t>(gridDim.x) * BLOCK;
    for (size_t idx = static_cast<size_t>(blockIdx.x) * BLOCK + threadIdx.x; idx < n; idx += stride) {
        const int i = static_cast<int>(idx % static_cast<size_t>(nx));
        const int j = static_cast<int>((idx / static_cast<size_t>(nx)) % static_cast<size_t>(ny));
        const float x = (static_cast<float>(i) + 0.5f) * dx - halfx;
        const float y = (static_cast<float>(j) + 0.5f) * dy - halfy;
        const float r = sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy));
        if (r > radius) continue;
        const float t = r / radius;
        const float w = (1.0f - t * t) * (1.0f - t * t);
        v[idx] = v[idx] + amount * w;
    }
}

__device__ __forceinline__ unsigned long long wsum(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ unsigned long long wmax(unsigned long long v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    return v;
}

// stats_field_kernel (src/sim/stats.cu:45-60): each value rounded to fixed
// point on its own, then summed as integers.
__global__ __launch_bounds__(BLOCK) void k_stats_fields(Grid T, Grid C, Grid N, int c_on, int n_on, size_t nt,
                                                        size_t nc, size_t nn, unsigned long long* acc) {
    const size_t most = nt > nc ? (nt > nn ? nt : nn) : (nc > nn ? nc : nn);
    unsigned long long st = 0, mx = 0, sc = 0, sn = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * BLOCK + threadIdx.x; i < most;
         i += static_cast<size_t>(gridDim.x) * BLOCK) {
        if (i < nt) {
            const double t = static_cast<double>(T.value[i]);
            const unsigned long long ft = static_cast<unsigned long long>(to_fixed(t, ASTRO_DEPOSIT_SCALE_STATS));
            st += ft;
            mx = ft > mx ? ft : mx;
        }
        if (c_on && i < nc) sc += static_cast<unsigned long long>(to_fixed(static_cast<double>(C.value[i]), ASTRO_STATS_SCALE_FIELD));
        if (n_on && i < nn) sn += static_cast<unsigned long long>(to_fixed(static_cast<double>(N.value[i]), ASTRO_STATS_SCALE_FIELD));
    }
    st = wsum(st); sc = wsum(sc); sn = wsum(sn); mx = wmax(mx);
    if ((threadIdx.x & 63) == 0) {
        if (st) atomicAdd(&acc[SW_TEMP_MEDIUM], st);
        if (sc) atomicAdd(&acc[SW_CO2], sc);
        if (sn) atomicAdd(&acc[SW_N2], sn);
        atomicMax(&acc[SW_MAX_TEMP], mx);
    }
}

}  // namespace

// The three FTCS variants: square 2D, general 3D with nx % 4 == 0 (float4
// lanes), and the scalar general 3D kernel. Chambers of one launch are grouped
// by variant, one batched launch per variant present.
void launch_ftcs(const Launch& L, std::span<const Grid> gs, std::span<const Ftcs> fs, const char* name) {
    Batch<KFtcs2d> sq;
    Batch<KFtcs3dV4<64>> v64;
    Batch<KFtcs3dV4<32>> v32;
    Batch<KFtcs3dV4<16>> v16;
    Batch<KFtcs3dV4<8>> v8;
    Batch<KFtcs3d> sc;
    for (size_t i = 0; i < gs.size(); ++i) {
        const Grid& g = gs[i];
        const Ftcs& f = fs[i];
        if (f.square2d) {
            const int tiles = ((g.nx + 63) / 64) * ((g.ny + 3) / 4);
            sq.add({g.value, g.scratch, g.nx, f.coeff, f.bc, f.amb, f.rx}, blocks_for_xcd(tiles));
        } else if (g.nx % 4 == 0) {
            // the narrowest row group that spans nx in one pass (64 lanes at
            // most), unless the recorder carries a tuned width
            int w = g.nx >= 256 ? 64 : (g.nx > 64 ? 32 : (g.nx > 32 ? 16 : 8));
            if (L.rec) {
                const auto it = L.rec->ftcs_lanes.find(name);
                if (it != L.rec->ftcs_lanes.end() && it->second > 0) w = it->second;
            }
            const int rb = 4 * (64 / w);
            const int tiles = ((g.nx + 4 * w - 1) / (4 * w)) * ((g.ny + rb - 1) / rb) * ((g.nz + FTCS_ZC - 1) / FTCS_ZC);
            const Ftcs3Args a{g.value, g.scratch, g.nx, g.ny, g.nz, f.cx, f.cy, f.cz, f.bc, f.amb, f.rx, f.ry, f.rz};
            const int nb = blocks_for_xcd(tiles);
            if (w == 64) v64.add(a, nb);
            else if (w == 32) v32.add(a, nb);
            else if (w == 16) v16.add(a, nb);
            else v8.add(a, nb);
        } else {
            const int txy = ((g.nx + 63) / 64) * ((g.ny + 3) / 4);
            int zc = g.nz;
            while (zc > 4 && txy * ((g.nz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
            const int tiles = txy * ((g.nz + zc - 1) / zc);
            sc.add({g.value, g.scratch, g.nx, g.ny, g.nz, zc, f.cx, f.cy, f.cz, f.bc, f.amb, f.rx, f.ry, f.rz},
                   blocks_for_xcd(tiles));
        }
    }
    if (!sq.args.empty()) launch_k<KFtcs2d>(L, name, sq.args, sq.blocks);
    if (!v64.args.empty()) launch_k<KFtcs3dV4<64>>(L, name, v64.args, v64.blocks);
    if (!v32.args.empty()) launch_k<KFtcs3dV4<32>>(L, name, v32.args, v32.blocks);
    if (!v16.args.empty()) launch_k<KFtcs3dV4<16>>(L, name, v16.args, v16.blocks);
    if (!v8.args.empty()) launch_k<KFtcs3dV4<8>>(L, name, v8.args, v8.blocks);
    if (!sc.args.empty()) launch_k<KFtcs3d>(L, name, sc.args, sc.blocks);
}

template <class P>
void launch_stamp(const Launch& L, Views<P> ws) {
    Batch<KStamp<P>> b;
    for (const WorldDev<P>& w : ws) b.add({w.c, w.ctr, w.g[ASTRO_FIELD_IRRADIANCE]}, stride_grid(w.pcap));
    launch_k<KStamp<P>>(L, "occlusion_stamp", b.args, b.blocks);
}

// Lines along y are coalesced across lanes as they are; lines along x go
// through the LDS-transposing kernel.
void launch_sweep(const Launch& L, std::span<const Grid> irr, Ticks ts) {
    Batch<KSweep> cols;
    Batch<KSweepLanes> rows;
    for (size_t i = 0; i < irr.size(); ++i) {
        const Grid& g = irr[i];
        const TickParams& t = ts[i];
        const SweepArgs a{g.deposit, g.value, g.nx, g.ny, g.nz, t.sweep_axis, t.sweep_sign, g.scale, t.face_area,
                          t.source, t.ambient_irr};
        const int lines = (t.sweep_axis == 0 ? g.ny : g.nx) * g.nz;
        if (t.sweep_axis == 0) {
            // lanes per line: the fewest whose four steps each span the line, at most a wave
            int seg = 1;
            while (seg < 64 && 4 * seg < g.nx) seg *= 2;
            const int per_block = (BLOCK / 64) * (64 / seg);
            rows.add(SweepLanesArgs{a, seg}, (lines + per_block - 1) / per_block);
        } else {
            cols.add(a, (lines + BLOCK - 1) / BLOCK);
        }
    }
    if (!rows.args.empty()) launch_k<KSweepLanes>(L, "light_sweep", rows.args, rows.blocks);
    if (!cols.args.empty()) launch_k<KSweep>(L, "light_sweep", cols.args, cols.blocks);
}

void launch_fill(const Launch& L, float* v, size_t n, float value) {
    if (n == 0) return;
    k_fill<<<static_cast<unsigned>((n + BLOCK - 1) / BLOCK), BLOCK, 0, L.stream>>>(v, n, value);
}

void launch_brush(const Launch& L, const Grid& g, float cx, float cy, float radius, float amount) {
    const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
    if (n == 0) return;
    const size_t want = (n + BLOCK - 1) / BLOCK;
    const unsigned blocks = static_cast<unsigned>(want < 4096 ? want : 4096);
    k_brush<<<blocks, BLOCK, 0, L.stream>>>(g.value, g.nx, g.ny, n, g.fdx, g.fdy, cx, cy, radius, amount);
}

void launch_stats_fields(const Launch& L, const Grid& T, const Grid& C, const Grid& N, bool c_on, bool n_on,
                         unsigned long long* stats) {
    const size_t nt = static_cast<size_t>(T.nx) * T.ny * T.nz;
    const size_t nc = c_on ? static_cast<size_t>(C.nx) * C.ny * C.nz : 0;
    const size_t nn = n_on ? static_cast<size_t>(N.nx) * N.ny * N.nz : 0;
    k_stats_fields<<<1024, BLOCK, 0, L.stream>>>(T, C, N, c_on ? 1 : 0, n_on ? 1 : 0, nt, nc, nn, stats);
}

template void launch_stamp<double>(const Launch&, Views<double>);
template void launch_stamp<float>(const Launch&, Views<float>);

}  // namespace astro_dev
