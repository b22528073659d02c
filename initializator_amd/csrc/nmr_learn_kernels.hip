// nmr_learn_kernels.hip -- device kernels of the learned corrector: a 3-layer
// 3-D convolution net (6 -> 16 -> 16 -> 2 channels, 3x3x3 kernels, "same"
// zero padding, ReLU between layers) with BF16 weights and activations and
// FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), its backward pass (input and weight
// gradients), Adam, and the glue between the solver's FP32 fields and the
// net's tensors.
//
// Tensors live on the grid padded by one zero voxel per side, channels last:
// element (x, y, z, c) of an nx*ny*nz grid with C channels is at
// (((z+1)*(ny+2) + y+1)*(nx+2) + x+1)*C + c.  The halo stays zero, which is
// the zero padding of the convolution.
//
// Convolution as an implicit GEMM: out[v][co] = sum_{t,ci} W[co][ci][t] *
// in[v + off(t)][ci].  One wavefront computes a row of 16 x-consecutive
// voxels for all 16 output channels: the MFMA's A operand is the weight
// matrix (16 output channels x 32 reduction indices), its B operand the
// input (32 reduction indices x 16 voxels), and the reduction index k of a
// 32-chunk covers 32/CIN taps of CIN channels.  The A fragments of all
// chunks stay in registers for the whole kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "nmr_kernels.h"

namespace nmr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float fl4 __attribute__((ext_vector_type(4)));

// Round to nearest even (NaN kept quiet), and back.
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool bf_pos(uint16_t h) { return h != 0 && (h & 0x8000u) == 0; }

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// tap t = kz*9 + ky*3 + kx, offsets -1..1 on each axis
__device__ __forceinline__ long tap_off(int t, long px, long plane)
{
    const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
    return (long)(kz - 1) * plane + (long)(ky - 1) * px + (long)(kx - 1);
}

// MODE 0: + bias, ReLU, BF16 out (16 ch); 1: + bias, FP32 out (2 ch);
// 2: gated by ref > 0, BF16 out (16 ch) -- the input gradient through ReLU.
template <int CIN, int MODE>
__global__ void __launch_bounds__(256) conv_kernel(ConvArgs a)
{
    constexpr int TPC = 32 / CIN;               // taps per reduction chunk
    constexpr int NCH = (27 + TPC - 1) / TPC;   // chunks
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    bf16x8 A[NCH];
    const uint4* wa4 = reinterpret_cast<const uint4*>(a.wa);
#pragma unroll
    for (int c = 0; c < NCH; c++) A[c] = as_bf8(wa4[c * 64 + lane]);
    const long px = a.nx + 2, plane = px * (long)(a.ny + 2);
    const int cin0 = (8 * g) % CIN;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    for (int r = wave; r < a.nrows; r += nwaves) {
        const int row = a.rows ? a.rows[r] : r;
        const int seg = row % a.nseg, yz = row / a.nseg;
        const int y = yz % a.ny, z = yz / a.ny;
        const int x = seg * 16 + n;
        const int xv = x < a.nx ? x : a.nx - 1;
        const long base = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * px + (xv + 1);
        uint4 B[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int t = c * TPC + (8 * g) / CIN;
            B[c] = make_uint4(0u, 0u, 0u, 0u);
            if (t < 27)
                B[c] = *reinterpret_cast<const uint4*>(a.in + (base + tap_off(t, px, plane)) * CIN + cin0);
        }
        fl4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < NCH; c++) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], as_bf8(B[c]), acc, 0, 0, 0);
        if (x >= a.nx) continue;
        // lane holds output channels 4g..4g+3 of voxel n
        if constexpr (MODE == 1) {
            if (g == 0) {
                a.outf[base * 2 + 0] = acc[0] + a.bias[0];
                a.outf[base * 2 + 1] = acc[1] + a.bias[1];
            }
        } else {
            uint16_t h[4];
            if constexpr (MODE == 0) {
#pragma unroll
                for (int j = 0; j < 4; j++) h[j] = f2bf(fmaxf(acc[j] + a.bias[4 * g + j], 0.f));
            } else {
                const uint2 rv = *reinterpret_cast<const uint2*>(a.ref + base * 16 + 4 * g);
                const uint16_t rr[4] = {(uint16_t)(rv.x & 0xffffu), (uint16_t)(rv.x >> 16),
                                        (uint16_t)(rv.y & 0xffffu), (uint16_t)(rv.y >> 16)};
#pragma unroll
                for (int j = 0; j < 4; j++) h[j] = bf_pos(rr[j]) ? f2bf(acc[j]) : (uint16_t)0;
            }
            *reinterpret_cast<uint2*>(a.out + base * 16 + 4 * g) =
                make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
        }
    }
}

// The same convolution over tiles of 16 x 8 x 4 output voxels: the tile's
// input region (18 x 10 x 6 voxels, all channels) is staged in LDS once and
// every B operand is read from there, instead of 27 gathers per voxel from
// L1/L2.  Each wavefront computes 8 of the tile's 32 output rows with the
// row kernel's MFMA sequence (same operands, same chunk order), so the two
// kernels produce the same bits.  Persistent: workgroup b takes tiles b,
// b + gridDim.x, ...; every row of the grid is computed (no row list).
constexpr int kTileY = 8, kTileZ = 4;
template <int CIN, int MODE>
__global__ void __launch_bounds__(256) conv_tile_kernel(ConvArgs a)
{
    constexpr int TPC = 32 / CIN;
    constexpr int NCH = (27 + TPC - 1) / TPC;
    constexpr int HX = 18, HY = kTileY + 2, HZ = kTileZ + 2;
    constexpr int NV = HX * HY * HZ;
    constexpr int V16 = CIN / 8; // 16-byte parts per voxel
    __shared__ uint4 sin[NV * V16];
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    const int wave = threadIdx.x >> 6;
    bf16x8 A[NCH];
    const uint4* wa4 = reinterpret_cast<const uint4*>(a.wa);
#pragma unroll
    for (int c = 0; c < NCH; c++) A[c] = as_bf8(wa4[c * 64 + lane]);
    const long px = a.nx + 2;
    const int nty = (a.ny + kTileY - 1) / kTileY, ntz = (a.nz + kTileZ - 1) / kTileZ;
    const int ntiles = a.nseg * nty * ntz;
    const int part0 = ((8 * g) % CIN) / 8;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int seg = tile % a.nseg, r2 = tile / a.nseg;
        const int x0 = seg * 16, y0 = (r2 % nty) * kTileY, z0 = (r2 / nty) * kTileZ;
        // staged voxel (xx, yy, zz) = padded voxel (x0 + xx, y0 + yy, z0 + zz),
        // clamped to the padded grid (clamped ones feed only outputs that are
        // not stored)
        for (int i = threadIdx.x; i < NV * V16; i += 256) {
            const int v = i / V16, pt = i % V16;
            const int xx = v % HX, yy = (v / HX) % HY, zz = v / (HX * HY);
            const long gx = min(x0 + xx, a.nx + 1), gy = min(y0 + yy, a.ny + 1), gz = min(z0 + zz, a.nz + 1);
            sin[i] = *reinterpret_cast<const uint4*>(a.in + ((gz * (a.ny + 2) + gy) * px + gx) * CIN + pt * 8);
        }
        __syncthreads();
        for (int rr = wave; rr < kTileY * kTileZ; rr += 4) {
            const int ly = rr % kTileY, lz = rr / kTileY;
            fl4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int t = c * TPC + (8 * g) / CIN;
                uint4 B = make_uint4(0u, 0u, 0u, 0u);
                if (t < 27) {
                    const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
                    B = sin[(((lz + kz) * HY + (ly + ky)) * HX + (n + kx)) * V16 + part0];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], as_bf8(B), acc, 0, 0, 0);
            }
            const int x = x0 + n, y = y0 + ly, z = z0 + lz;
            if (x >= a.nx || y >= a.ny || z >= a.nz) continue;
            const long base = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * px + (x + 1);
            // lane holds output channels 4g..4g+3 of voxel n
            if constexpr (MODE == 1) {
                if (g == 0) {
                    a.outf[base * 2 + 0] = acc[0] + a.bias[0];
                    a.outf[base * 2 + 1] = acc[1] + a.bias[1];
                }
            } else {
                uint16_t h[4];
                if constexpr (MODE == 0) {
#pragma unroll
                    for (int j = 0; j < 4; j++) h[j] = f2bf(fmaxf(acc[j] + a.bias[4 * g + j], 0.f));
                } else {
                    const uint2 rv = *reinterpret_cast<const uint2*>(a.ref + base * 16 + 4 * g);
                    const uint16_t rr4[4] = {(uint16_t)(rv.x & 0xffffu), (uint16_t)(rv.x >> 16),
                                             (uint16_t)(rv.y & 0xffffu), (uint16_t)(rv.y >> 16)};
#pragma unroll
                    for (int j = 0; j < 4; j++) h[j] = bf_pos(rr4[j]) ? f2bf(acc[j]) : (uint16_t)0;
                }
                *reinterpret_cast<uint2*>(a.out + base * 16 + 4 * g) =
                    make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
            }
        }
        __syncthreads();
    }
}

// The same convolution marching along z through columns of 16 x 16 output
// rows: the input planes of the column (18 x 18 voxels, all channels) pass
// through a ring of four LDS slots, so that each is loaded once per column
// instead of once per 16 x 8 x 4 tile (1.27 input reads per output voxel
// instead of 2.1), and the next plane's global loads are in flight while the
// current output plane is computed.  Output plane z reads padded input
// planes z, z + 1, z + 2 (slots z & 3, ...); plane z + 3 goes into slot
// (z + 3) & 3, which no wavefront reads during plane z, and the barrier at
// the end of each plane orders it before its first read.  Wavefront w
// computes rows w, w + 4, w + 8, w + 12 of a plane as four independent
// accumulators that share each weight fragment; per row the MFMA sequence is
// the row kernel's (same operands, same chunk order), so the bits equal
// those of conv_kernel and conv_tile_kernel.  Persistent: workgroup b takes
// columns b, b + gridDim.x, ...; a column is a.zc output planes (the last
// one what is left).  MODE 4 (8 input channels) is MODE 0 with the input's
// channels 0 and 1 staged from the FP32 state a.smx/smy times the pore
// channel, rounded as feature_kernel rounds them, so that the feature
// tensor's other channels (pore, velocity), which do not change during a
// run, are written once.
constexpr int kZmY = 16;
template <int N>
struct Plane {
    uint4 v[N];
};
template <int N>
struct PlaneS {
    uint4 v[N];
    float fx[N], fy[N];
};
template <int CIN, int MODE>
__global__ void __launch_bounds__(256, 3) conv_zm_kernel(ConvArgs a)
{
    constexpr int TPC = 32 / CIN;
    constexpr int NCH = (27 + TPC - 1) / TPC;
    constexpr int HX = 18, HY = kZmY + 2;
    constexpr int V16 = CIN / 8;
    constexpr int PV = HX * HY * V16;          // uint4 per staged plane
    constexpr int NLD = (PV + 255) / 256;
    constexpr int ROWB = HX * V16 * 16;         // bytes per staged row
    __shared__ uint4 sin[4 * PV];
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    const int wave = threadIdx.x >> 6;
    bf16x8 A[NCH];
    const uint4* wa4 = reinterpret_cast<const uint4*>(a.wa);
#pragma unroll
    for (int c = 0; c < NCH; c++) A[c] = as_bf8(wa4[c * 64 + lane]);
    // this lane's B element of chunk c within a staged plane, in bytes from
    // the wavefront's first row; hi bit c: the lane's tap lies in the later
    // of the two planes a chunk spans; `last`: the lane's tap of the last
    // chunk exists (t < 27)
    int inner[NCH];
    uint32_t hi = 0;
    const int tg = (8 * g) / CIN, part0 = ((8 * g) % CIN) / 8;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const int t = min(c * TPC + tg, 26);
        const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
        inner[c] = ((ky * HX + n + kx) * V16 + part0) * 16;
        if (kz != (c * TPC) / 9) hi |= 1u << c;
    }
    const bool last = (NCH - 1) * TPC + tg < 27;
    static_assert(MODE != 4 || CIN == 8, "state-fed input is the first layer's");
    float bias[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == 0 || MODE == 4) {
#pragma unroll
        for (int j = 0; j < 4; j++) bias[j] = a.bias[4 * g + j];
    } else if constexpr (MODE == 1) {
        bias[0] = a.bias[0];
        bias[1] = a.bias[1];
    }
    const long px = a.nx + 2;
    const long pstride = (long)(a.ny + 2) * px * CIN; // elements per padded plane
    const int zc = a.zc > 0 ? a.zc : 16;
    const int nyb = (a.ny + kZmY - 1) / kZmY, nzb = (a.nz + zc - 1) / zc;
    const int ncol = a.nseg * nyb * nzb;
    const char* lds = reinterpret_cast<const char*>(sin);
    for (int col = blockIdx.x; col < ncol; col += gridDim.x) {
        const int seg = col % a.nseg, r2 = col / a.nseg;
        const int x0 = seg * 16, y0 = (r2 % nyb) * kZmY, zs = (r2 / nyb) * zc;
        const int ze = min(zs + zc, a.nz);
        // staged element i = threadIdx.x + 256 j of a plane: voxel (xx, yy) =
        // padded (x0 + xx, y0 + yy), clamped to the padded grid (clamped ones
        // feed only outputs that are not stored); a thread's element past the
        // plane loads element 0 again and is not stored
        long goff[NLD];
        long sidx[NLD]; // MODE 4: the state index of an interior voxel in its plane, else -1
#pragma unroll
        for (int j = 0; j < NLD; j++) {
            const int i = threadIdx.x + 256 * j < PV ? threadIdx.x + 256 * j : 0;
            const int v = i / V16, pt = i % V16;
            const long gx = min(x0 + v % HX, a.nx + 1), gy = min(y0 + v / HX, a.ny + 1);
            goff[j] = (gy * px + gx) * CIN + pt * 8;
            sidx[j] = (gx >= 1 && gx <= a.nx && gy >= 1 && gy <= a.ny) ? (gy - 1) * a.spitch + gx - 1 : -1;
        }
        using PL = typename std::conditional<MODE == 4, PlaneS<NLD>, Plane<NLD>>::type;
        const long splane = (long)a.ny * a.spitch;
        auto load = [&](int pz) {
            PL r;
#pragma unroll
            for (int j = 0; j < NLD; j++) {
                r.v[j] = *reinterpret_cast<const uint4*>(a.in + (long)pz * pstride + goff[j]);
                if constexpr (MODE == 4) {
                    const bool in = sidx[j] >= 0 && pz >= 1 && pz <= a.nz;
                    const long c = in ? (long)(pz - 1) * splane + sidx[j] : 0;
                    r.fx[j] = in ? a.smx[c] : 0.f;
                    r.fy[j] = in ? a.smy[c] : 0.f;
                }
            }
            return r;
        };
        auto stash = [&](int pz, const PL& r) {
#pragma unroll
            for (int j = 0; j < NLD; j++)
                if (threadIdx.x + 256 * j < PV) {
                    uint4 w = r.v[j];
                    if constexpr (MODE == 4) {
                        if (sidx[j] >= 0 && pz >= 1 && pz <= a.nz) {
                            // feature_kernel's channels 0, 1: (Mx, My) * pore
                            const float pore = (w.y & 0xffffu) != 0u ? 1.f : 0.f;
                            w.x = (uint32_t)f2bf(r.fx[j] * pore) | ((uint32_t)f2bf(r.fy[j] * pore) << 16);
                        }
                    }
                    sin[(pz & 3) * PV + threadIdx.x + 256 * j] = w;
                }
        };
        {
            const PL r0 = load(zs), r1 = load(zs + 1), r2v = load(zs + 2);
            stash(zs, r0);
            stash(zs + 1, r1);
            stash(zs + 2, r2v);
        }
        __syncthreads();
        const int x = x0 + n;
        for (int z = zs; z < ze; z++) {
            // output voxel of row i of this wavefront
            long base[4];
            bool store[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int y = y0 + wave + 4 * i;
                store[i] = x < a.nx && y < a.ny;
                base[i] = ((long)(z + 1) * (a.ny + 2) + (min(y, a.ny - 1) + 1)) * px + (min(x, a.nx - 1) + 1);
            }
            uint2 gate[4];
            if constexpr (MODE == 2) {
#pragma unroll
                for (int i = 0; i < 4; i++) gate[i] = *reinterpret_cast<const uint2*>(a.ref + base[i] * 16 + 4 * g);
            }
            const bool more = z + 3 <= ze + 1;
            PL nx4{};
            if (more) nx4 = load(z + 3);
            int sb[3];
#pragma unroll
            for (int kz = 0; kz < 3; kz++) sb[kz] = ((z + kz) & 3) * PV * 16 + wave * ROWB;
            fl4 acc[4];
#pragma unroll
            for (int i = 0; i < 4; i++) acc[i] = fl4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int kzl = (c * TPC) / 9, kzh = min(c * TPC + TPC - 1, 26) / 9;
                const int ad = ((kzl == kzh || !((hi >> c) & 1u)) ? sb[kzl] : sb[kzh]) + inner[c];
                uint4 B[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    B[i] = *reinterpret_cast<const uint4*>(lds + ad + i * 4 * ROWB);
                    if (c == NCH - 1 && !last) B[i] = make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], as_bf8(B[i]), acc[i], 0, 0, 0);
            }
            // lane holds output channels 4g..4g+3 of voxel n
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!store[i]) continue;
                if constexpr (MODE == 1) {
                    if (g == 0) {
                        a.outf[base[i] * 2 + 0] = acc[i][0] + bias[0];
                        a.outf[base[i] * 2 + 1] = acc[i][1] + bias[1];
                    }
                } else {
                    uint16_t h[4];
                    if constexpr (MODE == 0 || MODE == 4) {
#pragma unroll
                        for (int j = 0; j < 4; j++) h[j] = f2bf(fmaxf(acc[i][j] + bias[j], 0.f));
                    } else {
                        const uint16_t rr4[4] = {(uint16_t)(gate[i].x & 0xffffu), (uint16_t)(gate[i].x >> 16),
                                                 (uint16_t)(gate[i].y & 0xffffu), (uint16_t)(gate[i].y >> 16)};
#pragma unroll
                        for (int j = 0; j < 4; j++) h[j] = bf_pos(rr4[j]) ? f2bf(acc[i][j]) : (uint16_t)0;
                    }
                    *reinterpret_cast<uint2*>(a.out + base[i] * 16 + 4 * g) =
                        make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
                }
            }
            if (more) stash(z + 3, nx4);
            __syncthreads();
        }
    }
}

// Weight gradient dW[t][ci][co] = sum_v in[v + off(t)][ci] * gout[v][co] and
// bias gradient sum_v gout[v][co], per wavefront over 32-voxel x segments,
// into part[wave][27*256 + 16].  MFMA: A = input (16 ci x 32 voxels), B =
// output gradient (32 voxels x 16 co), both gathered per lane.
template <int CIN>
__global__ void __launch_bounds__(256) wgrad_kernel(WgradArgs a)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, n = lane & 15;
    const long px = a.nx + 2, plane = px * (long)(a.ny + 2);
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    fl4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; t++) acc[t] = fl4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const int nseg32 = (a.nx + 31) / 32;
    const int nrows = nseg32 * a.ny * a.nz;
    const bool chan = n < CIN;
    for (int r = wave; r < nrows; r += nwaves) {
        const int seg = r % nseg32, yz = r / nseg32;
        const int y = yz % a.ny, z = yz / a.ny;
        const long rowb = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * px;
        // B: gout[x0 + 8g + j][co = n]
        uint16_t bv[8];
        long vb[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int x = seg * 32 + 8 * g + j;
            const bool ok = x < a.nx;
            vb[j] = rowb + (ok ? x : a.nx - 1) + 1;
            bv[j] = ok ? a.gout[vb[j] * 16 + n] : (uint16_t)0;
            bsum += bf2f(bv[j]);
        }
        const uint4 B = make_uint4((uint32_t)bv[0] | ((uint32_t)bv[1] << 16), (uint32_t)bv[2] | ((uint32_t)bv[3] << 16),
                                   (uint32_t)bv[4] | ((uint32_t)bv[5] << 16), (uint32_t)bv[6] | ((uint32_t)bv[7] << 16));
#pragma unroll
        for (int t = 0; t < 27; t++) {
            const long o = tap_off(t, px, plane);
            uint16_t av[8];
#pragma unroll
            for (int j = 0; j < 8; j++) av[j] = chan ? a.in[(vb[j] + o) * CIN + n] : (uint16_t)0;
            const uint4 Av = make_uint4((uint32_t)av[0] | ((uint32_t)av[1] << 16),
                                        (uint32_t)av[2] | ((uint32_t)av[3] << 16),
                                        (uint32_t)av[4] | ((uint32_t)av[5] << 16),
                                        (uint32_t)av[6] | ((uint32_t)av[7] << 16));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(Av), as_bf8(B), acc[t], 0, 0, 0);
        }
    }
    float* dst = a.part + (size_t)wave * kWgradPart;
    // D[i = ci = 4g + j][col = co = n]
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) dst[t * 256 + (4 * g + j) * 16 + n] = acc[t][j];
    // bias: lanes n, n+16, n+32, n+48 hold the same co
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (g == 0) dst[27 * 256 + n] = bsum;
}

// Sum of the wavefront partials in a fixed order, scattered to the
// canonical gradient layout W[co][ci][t] (cout_real x cin_real x 27) followed
// by the bias.
__global__ void __launch_bounds__(256) wreduce_kernel(const float* __restrict__ part, int nwaves, int cin_real,
                                                      int cout_real, float* __restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kWgradPart) return;
    double s = 0.0;
    for (int w = 0; w < nwaves; w++) s += (double)part[(size_t)w * kWgradPart + e];
    if (e < 27 * 256) {
        const int t = e / 256, ci = (e / 16) % 16, co = e % 16;
        if (ci < cin_real && co < cout_real) grad[((size_t)co * cin_real + ci) * 27 + t] = (float)s;
    } else {
        const int co = e - 27 * 256;
        if (co < cout_real) grad[(size_t)cout_real * cin_real * 27 + co] = (float)s;
    }
}

// Packs the canonical FP32 weights of one layer, W[co][ci][t], into the A
// fragments of conv_kernel: forward (rows = output channels co, reduction
// (t, ci) with cin_store channels per tap) or input gradient (rows = ci,
// reduction (t', co) over 16 stored channels, W'[t'][co][ci] = W[26-t'][ci][co]).
__global__ void __launch_bounds__(256) pack_kernel(const float* __restrict__ w, int cin_real, int cout_real,
                                                   int cin_store, int backward, uint16_t* __restrict__ frag)
{
    const int cin = backward ? 16 : cin_store;
    const int tpc = 32 / cin, nch = (27 + tpc - 1) / tpc;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nch * 64 * 8) return;
    const int c = e / 512, lane = (e / 8) % 64, j = e % 8;
    const int g = lane / 16, row = lane % 16;
    const int k = 8 * g + j;
    const int t = c * tpc + k / cin, ch = k % cin;
    float v = 0.f;
    if (t < 27) {
        if (!backward) {
            if (ch < cin_real && row < cout_real) v = w[((size_t)row * cin_real + ch) * 27 + t];
        } else {
            if (ch < cout_real && row < cin_real) v = w[((size_t)ch * cin_real + row) * 27 + (26 - t)];
        }
    }
    frag[e] = f2bf(v);
}

// Input tensor of the net: (Mx, My) * pore, pore, cell-centred velocity /
// vscale, two zero channels, BF16, 8 channels per voxel.
__global__ void __launch_bounds__(256) feature_kernel(FeatureArgs a)
{
    const long n = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % a.nx), y = (int)((v / a.nx) % a.ny), z = (int)(v / ((long)a.nx * a.ny));
    const size_t c = ((size_t)z * a.ny + y) * a.pitch + x;
    const float pore = a.mask[c] ? 1.f : 0.f;
    float f[8] = {a.mx[c] * pore, a.my ? a.my[c] * pore : 0.f, pore, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (a.vx) {
        const size_t cx = ((size_t)z * a.ny + y) * a.pitch1 + x;
        const size_t cy = ((size_t)z * (a.ny + 1) + y) * a.pitch + x;
        const size_t plane = (size_t)a.ny * a.pitch;
        f[3] = 0.5f * (a.vx[cx] + a.vx[cx + 1]) * a.vinv;
        f[4] = 0.5f * (a.vy[cy] + a.vy[cy + a.pitch]) * a.vinv;
        f[5] = 0.5f * (a.vz[c] + a.vz[c + plane]) * a.vinv;
    }
    const long pv = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * (a.nx + 2) + (x + 1);
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = (uint32_t)f2bf(f[2 * q]) | ((uint32_t)f2bf(f[2 * q + 1]) << 16);
    *reinterpret_cast<uint4*>(a.out + pv * 8) = make_uint4(w[0], w[1], w[2], w[3]);
}

// Same tensor from host-given FP32 features [nz][ny][nx][6] (test entry).
__global__ void __launch_bounds__(256) feature_copy_kernel(const float* __restrict__ f6, int nx, int ny, int nz,
                                                           uint16_t* __restrict__ out)
{
    const long n = (long)nx * ny * nz;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
    const long pv = ((long)(z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + (x + 1);
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float lo = 2 * q < 6 ? f6[v * 6 + 2 * q] : 0.f, hi = 2 * q + 1 < 6 ? f6[v * 6 + 2 * q + 1] : 0.f;
        w[q] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
    }
    *reinterpret_cast<uint4*>(out + pv * 8) = make_uint4(w[0], w[1], w[2], w[3]);
}

// Output gradient of the last layer as a BF16 16-channel tensor from
// host-given FP32 values [nz][ny][nx][2] (test entry).
__global__ void __launch_bounds__(256) gout_copy_kernel(const float* __restrict__ g2, int nx, int ny, int nz,
                                                        uint16_t* __restrict__ out)
{
    const long n = (long)nx * ny * nz;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
    const long pv = ((long)(z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + (x + 1);
    uint4* o = reinterpret_cast<uint4*>(out + pv * 16);
    o[0] = make_uint4((uint32_t)f2bf(g2[v * 2]) | ((uint32_t)f2bf(g2[v * 2 + 1]) << 16), 0u, 0u, 0u);
    o[1] = make_uint4(0u, 0u, 0u, 0u);
}

// Squared-error loss over pores of (pred - target/scale) and its gradient
// 2 (pred - t) / npore as the BF16 16-channel output gradient; per-block
// partial loss sums.
__global__ void __launch_bounds__(256) loss_kernel(LossArgs a)
{
    __shared__ double sh[256];
    const long n = (long)a.nx * a.ny * a.nz;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (v < n) {
        const int x = (int)(v % a.nx), y = (int)((v / a.nx) % a.ny), z = (int)(v / ((long)a.nx * a.ny));
        const size_t c = ((size_t)z * a.ny + y) * a.pitch + x;
        const long pv = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * (a.nx + 2) + (x + 1);
        float g0 = 0.f, g1 = 0.f;
        if (a.mask[c]) {
            const float d0 = a.pred[pv * 2] - a.tx[c] * a.tinv;
            const float d1 = a.pred[pv * 2 + 1] - a.ty[c] * a.tinv;
            l = (double)d0 * d0 + (double)d1 * d1;
            g0 = 2.f * d0 * a.winv;
            g1 = 2.f * d1 * a.winv;
        }
        uint4* o = reinterpret_cast<uint4*>(a.gout + pv * 16);
        o[0] = make_uint4((uint32_t)f2bf(g0) | ((uint32_t)f2bf(g1) << 16), 0u, 0u, 0u);
        o[1] = make_uint4(0u, 0u, 0u, 0u);
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.lpart[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters (bias-corrected moments).
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                   const float* __restrict__ g, int n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// target = after - coarse over both components (compact layout).
__global__ void __launch_bounds__(256) diff_kernel(const float* __restrict__ ax, const float* __restrict__ ay,
                                                   const float* __restrict__ bx, const float* __restrict__ by,
                                                   float* __restrict__ ox, float* __restrict__ oy, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ox[i] = ax[i] - bx[i];
    oy[i] = ay[i] - by[i];
}

// Adds scale * pred to the pores of (mx, my) and reduces, per block, the
// state's pore energy sum(Mx^2 + My^2), the largest |correction|, the
// largest |M| of the state before the correction, a non-finite count, and
// the signal sums over the trim window: part[block][6].
constexpr int kApplyBlocks = 2048; // workgroups (= partials) of the apply pass
__global__ void __launch_bounds__(256) apply_kernel(ApplyArgs a)
{
    __shared__ double sh[4][6];
    const long n = (long)a.nx * a.ny * a.nz;
    double st[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; // energy, max |d|, max |M|, non-finite, sum Mx, sum My
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const int x = (int)(v % a.nx), y = (int)((v / a.nx) % a.ny), z = (int)(v / ((long)a.nx * a.ny));
        const size_t c = ((size_t)z * a.ny + y) * a.pitch + x;
        if (!a.mask[c]) continue;
        float mx = a.mx[c], my = a.my[c];
        st[2] = fmax(st[2], (double)fmaxf(fabsf(mx), fabsf(my)));
        if (a.pred) {
            const long pv = ((long)(z + 1) * (a.ny + 2) + (y + 1)) * (a.nx + 2) + (x + 1);
            const float d0 = a.scale * a.pred[pv * 2], d1 = a.scale * a.pred[pv * 2 + 1];
            mx += d0;
            my += d1;
            a.mx[c] = mx;
            a.my[c] = my;
            st[1] = fmax(st[1], (double)fmaxf(fabsf(d0), fabsf(d1)));
            if (!(fabsf(d0) <= 3.0e38f) || !(fabsf(d1) <= 3.0e38f)) st[3] += 1.0;
        }
        if (!(fabsf(mx) <= 3.0e38f) || !(fabsf(my) <= 3.0e38f)) st[3] += 1.0;
        st[0] += (double)mx * mx + (double)my * my;
        if (x >= a.bx0 && x < a.bx1 && y >= a.by0 && y < a.by1 && z >= a.bz0 && z < a.bz1) {
            st[4] += mx;
            st[5] += my;
        }
    }
    // wavefront, then workgroup (fixed order)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const double o = __shfl_xor(st[k], off);
            st[k] = (k == 1 || k == 2) ? fmax(st[k], o) : st[k] + o;
        }
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; k++) sh[wave][k] = st[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double r = sh[0][k];
        for (int w = 1; w < 4; w++) r = (k == 1 || k == 2) ? fmax(r, sh[w][k]) : r + sh[w][k];
        a.part[(size_t)blockIdx.x * 6 + k] = r;
    }
}

__global__ void __launch_bounds__(256) stats_reduce_kernel(const double* __restrict__ part, int nblocks,
                                                           double* __restrict__ out)
{
    __shared__ double sh[6][256];
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        const double* q = part + (size_t)b * 6;
        v[0] += q[0];
        v[1] = fmax(v[1], q[1]);
        v[2] = fmax(v[2], q[2]);
        v[3] += q[3];
        v[4] += q[4];
        v[5] += q[5];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + h]);
            sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + h]);
            sh[3][threadIdx.x] += sh[3][threadIdx.x + h];
            sh[4][threadIdx.x] += sh[4][threadIdx.x + h];
            sh[5][threadIdx.x] += sh[5][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

// flag[row] = 1 when the 16-voxel x segment `row` holds a pore.
__global__ void __launch_bounds__(256) row_flag_kernel(const uint8_t* __restrict__ mask, int nx, int ny, int nz,
                                                       int pitch, int nseg, uint8_t* __restrict__ flag)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long)nseg * ny * nz) return;
    const int seg = (int)(r % nseg);
    const long yz = r / nseg;
    const uint8_t* m = mask + (size_t)yz * pitch;
    uint8_t f = 0;
    for (int x = seg * 16; x < min(seg * 16 + 16, nx); x++) f |= m[x];
    flag[r] = f ? 1 : 0;
}

} // namespace

// ---- launchers ----

hipError_t launch_conv(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s)
{
    const dim3 g(grid), b(256);
    if (cin == 8 && mode == 0) hipLaunchKernelGGL((conv_kernel<8, 0>), g, b, 0, s, a);
    else if (cin == 16 && mode == 0) hipLaunchKernelGGL((conv_kernel<16, 0>), g, b, 0, s, a);
    else if (cin == 16 && mode == 1) hipLaunchKernelGGL((conv_kernel<16, 1>), g, b, 0, s, a);
    else if (cin == 16 && mode == 2) hipLaunchKernelGGL((conv_kernel<16, 2>), g, b, 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

int conv_tiles(int nx, int ny, int nz)
{
    return ((nx + 15) / 16) * ((ny + kTileY - 1) / kTileY) * ((nz + kTileZ - 1) / kTileZ);
}

hipError_t launch_conv_tiled(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s)
{
    const dim3 g(grid), b(256);
    if (cin == 8 && mode == 0) hipLaunchKernelGGL((conv_tile_kernel<8, 0>), g, b, 0, s, a);
    else if (cin == 16 && mode == 0) hipLaunchKernelGGL((conv_tile_kernel<16, 0>), g, b, 0, s, a);
    else if (cin == 16 && mode == 1) hipLaunchKernelGGL((conv_tile_kernel<16, 1>), g, b, 0, s, a);
    else if (cin == 16 && mode == 2) hipLaunchKernelGGL((conv_tile_kernel<16, 2>), g, b, 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

int conv_columns(int nx, int ny, int nz, int zc)
{
    return ((nx + 15) / 16) * ((ny + kZmY - 1) / kZmY) * ((nz + zc - 1) / zc);
}

hipError_t launch_conv_zm(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s)
{
    const dim3 g(grid), b(256);
    if (cin == 8 && mode == 0) hipLaunchKernelGGL((conv_zm_kernel<8, 0>), g, b, 0, s, a);
    else if (cin == 8 && mode == 4) hipLaunchKernelGGL((conv_zm_kernel<8, 4>), g, b, 0, s, a);
    else if (cin == 16 && mode == 0) hipLaunchKernelGGL((conv_zm_kernel<16, 0>), g, b, 0, s, a);
    else if (cin == 16 && mode == 1) hipLaunchKernelGGL((conv_zm_kernel<16, 1>), g, b, 0, s, a);
    else if (cin == 16 && mode == 2) hipLaunchKernelGGL((conv_zm_kernel<16, 2>), g, b, 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_wgrad(int cin, const WgradArgs& a, int grid, hipStream_t s)
{
    if (cin == 8) hipLaunchKernelGGL((wgrad_kernel<8>), dim3(grid), dim3(256), 0, s, a);
    else if (cin == 16) hipLaunchKernelGGL((wgrad_kernel<16>), dim3(grid), dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_wreduce(const float* part, int nwaves, int cin_real, int cout_real, float* grad, hipStream_t s)
{
    hipLaunchKernelGGL(wreduce_kernel, dim3((kWgradPart + 255) / 256), dim3(256), 0, s, part, nwaves, cin_real,
                       cout_real, grad);
    return hipGetLastError();
}

hipError_t launch_pack(const float* w, int cin_real, int cout_real, int cin_store, int backward, uint16_t* frag,
                       hipStream_t s)
{
    const int cin = backward ? 16 : cin_store;
    const int tpc = 32 / cin, nch = (27 + tpc - 1) / tpc;
    const int n = nch * 64 * 8;
    hipLaunchKernelGGL(pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, cin_real, cout_real, cin_store,
                       backward, frag);
    return hipGetLastError();
}

static int blocks_for(long n) { return (int)((n + 255) / 256); }

hipError_t launch_features(const FeatureArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(feature_kernel, dim3(blocks_for((long)a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_feature_copy(const float* f6, int nx, int ny, int nz, uint16_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(feature_copy_kernel, dim3(blocks_for((long)nx * ny * nz)), dim3(256), 0, s, f6, nx, ny, nz,
                       out);
    return hipGetLastError();
}

hipError_t launch_gout_copy(const float* g2, int nx, int ny, int nz, uint16_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(gout_copy_kernel, dim3(blocks_for((long)nx * ny * nz)), dim3(256), 0, s, g2, nx, ny, nz,
                       out);
    return hipGetLastError();
}

int loss_blocks(int nx, int ny, int nz) { return blocks_for((long)nx * ny * nz); }

hipError_t launch_loss(const LossArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(loss_kernel, dim3(loss_blocks(a.nx, a.ny, a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, float* m, float* v, const float* g, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, s, p, m, v, g, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_diff(const float* ax, const float* ay, const float* bx, const float* by, float* ox, float* oy,
                       long n, hipStream_t s)
{
    hipLaunchKernelGGL(diff_kernel, dim3(blocks_for(n)), dim3(256), 0, s, ax, ay, bx, by, ox, oy, n);
    return hipGetLastError();
}

int apply_blocks(int nx, int ny, int nz) { return (int)std::max<long>(1, std::min<long>(kApplyBlocks, blocks_for((long)nx * ny * nz))); }

hipError_t launch_apply(const ApplyArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(apply_kernel, dim3(apply_blocks(a.nx, a.ny, a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace nmr

namespace nmr {

hipError_t launch_stats_reduce(const double* part, int nblocks, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(1), dim3(256), 0, s, part, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_row_flags(const uint8_t* mask, int nx, int ny, int nz, int pitch, uint8_t* flag, hipStream_t s)
{
    const int nseg = (nx + 15) / 16;
    const long n = (long)nseg * ny * nz;
    hipLaunchKernelGGL(row_flag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mask, nx, ny, nz, pitch,
                       nseg, flag);
    return hipGetLastError();
}

} // namespace nmr
