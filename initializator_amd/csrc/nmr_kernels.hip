// nmr_kernels.hip -- gfx950 kernels for the per-timestep Bloch-Torrey update.
//
// The update restated here is the reference's ExplicitSimple
// (src/NmrRelaxationSolve.cpp:3148-3282): forward Euler, relaxation toward
// pEq*n (T1 parallel, T2 perpendicular), linearised precession dt*g*(M x Bg),
// first-order upwind advection on staggered face velocities with wall
// indicators, harmonic-mean 7-point diffusion, Jacobi (read old M, write new).
//
// What differs from the reference's data path (the arithmetic is the same):
//  * State is FP32 in a compact SoA layout, x fastest, rows padded to a
//    multiple of 4 floats.  The ghost layer of the reference
//    (BoundaryConditionsForVectorField, cpp:1664-1754) is replaced by clamped
//    neighbour reads: an out-of-domain neighbour takes the centre's value
//    and the centre's mask, which is exactly what the zero-derivative ghost
//    copy gives (cpp:880-928 for the mask, cpp:1895-1942 for D).
//  * Per-voxel D, T1, T2 (cpp:1756-1893) and the gradient field
//    (cpp:1944-1989) are not read from memory.  For one fluid with uniform
//    relaxivity they follow from the mask: a pore-pore face conducts D, a
//    pore-solid face conducts 2*D*1e-50/(D+2e-50) (below FP64 resolution of
//    the update, taken as 0), 1/T = 1/T_bulk + rho/d per solid face, and
//    Bg = G*d*(index+1/2).
//  * The signal (OutputX/Y/Z, cpp:3301-3461) is reduced inside the sweep:
//    every updated FP32 voxel is added in FP64 per lane, then across waves and the workgroup,
//    one FP64 partial per workgroup per step, summed in a fixed order by
//    nmr_reduce_partials -- deterministic, no atomics.
//
// Kernel shape: 256 threads; a lane owns 4 x-consecutive voxels (one 16-byte
// load per array), LX lanes span a tile row, RPT rows per thread.  The
// workgroup owns a TX x TY tile and marches a z-chunk: planes k-1, k, k+1 of
// its own voxels stay in registers (every plane is loaded once per chunk),
// plane k plus a one-voxel halo is staged in LDS for the x/y neighbours
// (double buffered, one barrier per plane).  Tiles are remapped so that each
// XCD receives a contiguous run of y-adjacent tiles, whose shared halo rows
// then hit in that XCD's L2.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "nmr_kernels.h"


namespace nmr {

typedef float f4 __attribute__((ext_vector_type(4)));

// Global accesses take a scalar base pointer plus a 32-bit byte offset, so
// they map onto the scalar-base + 32-bit-VGPR-offset addressing form.
__device__ __forceinline__ f4 ld4_nt(const float* base, uint32_t boff)
{
    return __builtin_nontemporal_load(reinterpret_cast<const f4*>(reinterpret_cast<const char*>(base) + boff));
}
__device__ __forceinline__ f4 ld4(const float* base, uint32_t boff)
{
    return *reinterpret_cast<const f4*>(reinterpret_cast<const char*>(base) + boff);
}
__device__ __forceinline__ float ld1(const float* base, uint32_t boff)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + boff);
}
__device__ __forceinline__ void st4_nt(float* base, uint32_t boff, f4 v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<f4*>(reinterpret_cast<char*>(base) + boff));
}
__device__ __forceinline__ uint32_t ldw_nt(const uint8_t* base, uint32_t boff)
{
    return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(base + boff));
}
// Loads and stores of the state: with KEEP, without the non-temporal hint,
// so that the device's memory-side cache (256 MB) keeps what one pass
// writes for the next -- worth it while the state's two buffers fit about
// twice that (the host picks KEEP); past that the hint pays.
template <bool KEEP>
__device__ __forceinline__ f4 ld4_state(const float* base, uint32_t boff)
{
    if constexpr (KEEP) return ld4(base, boff);
    else return ld4_nt(base, boff);
}
template <bool KEEP>
__device__ __forceinline__ void st4_state(float* base, uint32_t boff, f4 v)
{
    if constexpr (KEEP) *reinterpret_cast<f4*>(reinterpret_cast<char*>(base) + boff) = v;
    else st4_nt(base, boff, v);
}

// Harmonic mean 2ab/(a + b) of two (scaled) diffusivities for the face
// between them, with the hardware reciprocal (v_rcp_f32, about 1 ulp) in
// place of an IEEE division.  Sums below 1e-30 -- solids, which the
// reference gives D = 1e-50 -- give 0: their reciprocal would overflow, and
// the exact mean is below 1e-30 there anyway.  Both general kernels use it,
// so they stay bit-identical.
__device__ __forceinline__ float hmean(float a, float b)
{
    const float s = a + b;
    return s > 1e-30f ? (2.f * a * b) * __builtin_amdgcn_rcpf(s) : 0.f;
}

// Float offset of voxel (x, y, z) of a field in bricks of 8 x-voxels by 4
// rows (brick (z, y/4, x/8); nby brick rows per plane, nb8 bricks per brick
// row), the layout brick_kernel converts to.
__device__ __forceinline__ size_t brick_off(int x, int y, int z, int nby, int nb8)
{
    return (((size_t)z * (size_t)nby + (size_t)(y >> 2)) * (size_t)nb8 + (size_t)(x >> 3)) * 32 +
           (size_t)((y & 3) * 8 + (x & 7));
}

// byte e (0/1) of a packed mask word as a float
__device__ __forceinline__ float ub(uint32_t w, int e) { return (float)((w >> (8 * e)) & 0xffu); }

// Block index -> tile index such that XCD x (blocks x, x+8, x+16, ...) gets
// the contiguous tile range [x*q + min(x,r), ...), q = nb/8, r = nb%8.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int xcd = b & 7, ib = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd * q + (xcd < r ? xcd : r) + ib;
}

// Replace the elements of a chunk that lie beyond the domain end (index >
// last) by element `last`: the zero-derivative ghost value.
__device__ __forceinline__ f4 clamp_tail(f4 v, int last)
{
    const float t = last == 0 ? v.x : (last == 1 ? v.y : v.z);
    if (last < 3) v.w = t;
    if (last < 2) v.z = t;
    if (last < 1) v.y = t;
    return v;
}
__device__ __forceinline__ uint32_t clamp_tail_w(uint32_t w, int last)
{
    if (last >= 3) return w;
    const uint32_t t = (w >> (8 * last)) & 0xffu;
    const uint32_t keep = (1u << (8 * (last + 1))) - 1u;
    return (w & keep) | ((t * 0x01010101u) & ~keep);
}

// Update of one 4-voxel x-chunk (ExplicitSimple, cpp:3174-3262), shared by the
// one-step and the two-step sweeps so that both give bit-identical voxels.
// Inputs: centre values mc, z/y neighbour chunks, x neighbours xl (x-1 of
// element 0) and xr (x+1 of element 3), packed 0/1 mask words of the chunk and
// of its six neighbour chunks (kxm/kxp already shifted by one voxel), and the
// upwind velocity sums of the chunk's voxels (adv_kernel: per axis the sum of
// the voxel's two face velocities, 0 where the upwind face carries none).
// Ghost cells hold zero-derivative copies, so no bounds tests are needed.
// Solid voxels keep their value.
template <int NC, bool VEL, bool EXACT>
__device__ __forceinline__ void update_chunk(const Coef& cf, const f4 (&mc)[NC], const f4 (&fzm)[NC],
                                             const f4 (&fzp)[NC], const f4 (&fym)[NC], const f4 (&fyp)[NC],
                                             const float (&xl)[NC], const float (&xr)[NC], uint32_t kc,
                                             uint32_t kxm, uint32_t kxp, uint32_t kym, uint32_t kyp,
                                             uint32_t kzm, uint32_t kzp, const f4& ax4, const f4& ay4,
                                             const f4& az4, int ix, float thy, float thz, float sn, float cm1,
                                             f4 (&out)[NC])
{
    const float hxh = 0.5f * cf.hx, hyh = 0.5f * cf.hy, hzh = 0.5f * cf.hz;
    // solid neighbours per axis, one byte per voxel
    const uint32_t solx = 0x02020202u - (kxm + kxp);
    const uint32_t soly = 0x02020202u - (kym + kyp);
    const uint32_t solz = 0x02020202u - (kzm + kzp);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        // face weights: diffusion between pores minus upwind advection.
        // Every fused multiply-add is written out (the file is built with
        // -ffp-contract=off) so that the arithmetic of a voxel does not depend
        // on the kernel or tile shape it was compiled into.
        float wxm = cf.cdx * ub(kxm, e), wxp = cf.cdx * ub(kxp, e);
        float wym = cf.cdy * ub(kym, e), wyp = cf.cdy * ub(kyp, e);
        float wzm = cf.cdz * ub(kzm, e), wzp = cf.cdz * ub(kzp, e);
        if constexpr (VEL) {
            // inflow across the minus face when the sum is positive, across
            // the plus face when negative; a zero sum adds h*0 (exact)
            wxm = fmaf(hxh, fmaxf(ax4[e], 0.f), wxm);
            wxp = fmaf(-hxh, fminf(ax4[e], 0.f), wxp);
            wym = fmaf(hyh, fmaxf(ay4[e], 0.f), wym);
            wyp = fmaf(-hyh, fminf(ay4[e], 0.f), wyp);
            wzm = fmaf(hzh, fmaxf(az4[e], 0.f), wzm);
            wzp = fmaf(-hzh, fminf(az4[e], 0.f), wzp);
        }
        const float nsx = ub(solx, e), nsy = ub(soly, e), nsz = ub(solz, e);
        const float rate2 = fmaf(cf.s2z, nsz, fmaf(cf.s2y, nsy, fmaf(cf.s2x, nsx, cf.r2)));
        float m0[NC], d[NC];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const float m = mc[q][e];
            m0[q] = m;
            const float fxm = (e > 0) ? mc[q][e - 1] : xl[q];
            const float fxp = (e < 3) ? mc[q][e + 1] : xr[q];
            float t = wxm * (fxm - m);
            t = fmaf(wxp, fxp - m, t);
            t = fmaf(wym, fym[q][e] - m, t);
            t = fmaf(wyp, fyp[q][e] - m, t);
            t = fmaf(wzm, fzm[q][e] - m, t);
            d[q] = fmaf(wzp, fzp[q][e] - m, t);
        }
        if constexpr (NC == 3) {
            const float rate1 = fmaf(cf.s1z, nsz, fmaf(cf.s1y, nsy, fmaf(cf.s1x, nsx, cf.r1)));
            const float par = fmaf(cf.n2, m0[2], fmaf(cf.n1, m0[1], cf.n0 * m0[0]));
            const float nn[3] = {cf.n0, cf.n1, cf.n2};
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const float mpar = nn[q] * par;
                d[q] = fmaf(rate1, fmaf(cf.peq, nn[q], -mpar), d[q]);
                d[q] = fmaf(rate2, mpar - m0[q], d[q]);
            }
            const float thx = cf.bgx * ((float)(ix + e) + 0.5f);
            // + dt*g*(m x Bg)
            d[0] = fmaf(thz, m0[1], fmaf(-thy, m0[2], d[0]));
            d[1] = fmaf(-thz, m0[0], fmaf(thx, m0[2], d[1]));
            d[2] = fmaf(thy, m0[0], fmaf(-thx, m0[1], d[2]));
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) d[q] = fmaf(-rate2, m0[q], d[q]);
            if constexpr (NC == 2) {
                if constexpr (EXACT) {
                    d[0] = fmaf(m0[1], sn, fmaf(m0[0], cm1, d[0]));
                    d[1] = fmaf(-m0[0], sn, fmaf(m0[1], cm1, d[1]));
                } else {
                    d[0] = fmaf(thz, m0[1], d[0]);
                    d[1] = fmaf(-thz, m0[0], d[1]);
                }
            }
        }
        const bool pore = ((kc >> (8 * e)) & 0xffu) != 0;
#pragma unroll
        for (int q = 0; q < NC; q++) out[q][e] = pore ? m0[q] + d[q] : m0[q];
    }
}

// One voxel of a step for the schemes and boundary conditions outside the
// z-sweep: the upwind step of ExplicitSimple (cpp:3148-3282, the same
// face-weight arithmetic as update_chunk, so that with zero-derivative ghosts
// the result is bit-identical to the sweep), or a stage of the MacCormack step
// (NMR_predictor_upgrade / NMR_corrector_upgrade, cu:637-967): advection with
// the cell-centred velocity u = (um + up)/2, backward differences in the
// predictor and forward differences in the corrector, a solid neighbour
// reading as the voxel's own value.  Neighbours outside the domain follow the
// axis's boundary condition (own value, opposite edge, or 0) and count as
// pores, as the reference's ghost mask copies the adjacent interior.
// Threads: 64 x-voxels by 4 rows per workgroup, one plane per blockIdx.z.
template <int NC, int ST>
__global__ void __launch_bounds__(256) general_kernel(GenArgs a)
{
    const int x = blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int z = blockIdx.z;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (x >= nx || y >= ny) return;
    const size_t P = (size_t)a.pitch;
    const size_t plane = (size_t)ny * P;
    const size_t c = (size_t)z * plane + (size_t)y * P + (size_t)x;
    if (a.mask[c] == 0) {
        if constexpr (ST == GEN_CORRECT) {
#pragma unroll
            for (int q = 0; q < NC; q++) a.out[q][c] = 0.f; // the reference's corrector holds 0 at solids
        }
        return; // upwind: solids keep their value (both buffers hold it)
    }
    const Coef& cf = a.c;
    // neighbour cells: index, kind (0 load with its mask, 1 own value, 2 zero,
    // 3 load as a pore), order x-, x+, y-, y+, z-, z+
    size_t ni[6];
    int kind[6];
    {
        const int lo[3] = {x, y, z}, n[3] = {nx, ny, nz};
        const size_t st[3] = {1, P, plane};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
#pragma unroll
            for (int sd = 0; sd < 2; sd++) {
                const int s = 2 * ax + sd;
                const bool out = sd == 0 ? lo[ax] == 0 : lo[ax] == n[ax] - 1;
                if (!out) {
                    ni[s] = sd == 0 ? c - st[ax] : c + st[ax];
                    kind[s] = 0;
                } else if (a.bc[ax] == BC_PERIODIC) {
                    ni[s] = sd == 0 ? c + (size_t)(n[ax] - 1) * st[ax] : c - (size_t)(n[ax] - 1) * st[ax];
                    kind[s] = 3;
                } else {
                    ni[s] = c;
                    kind[s] = a.bc[ax] == BC_DIRICHLET ? 2 : 1;
                }
            }
        }
    }
    float kn[6];
#pragma unroll
    for (int s = 0; s < 6; s++) kn[s] = kind[s] == 0 ? (a.mask[ni[s]] != 0 ? 1.f : 0.f) : 1.f;

    // face weights on (f_neighbour - f_centre)
    float w[6] = {cf.cdx * kn[0], cf.cdx * kn[1], cf.cdy * kn[2], cf.cdy * kn[3], cf.cdz * kn[4], cf.cdz * kn[5]};
    if (a.dn) { // per-voxel diffusivity: harmonic mean between pores; ghosts copy the centre
        const float dc = a.dn[c];
        const float cd[3] = {cf.cdx, cf.cdy, cf.cdz};
#pragma unroll
        for (int s = 0; s < 6; s++) {
            const float dnb = kind[s] == 0 || kind[s] == 3 ? (kind[s] == 3 ? dc : a.dn[ni[s]]) : dc;
            const float hm = hmean(dc, dnb);
            w[s] = cd[s / 2] * hm * kn[s];
        }
    }
    if (a.vel) {
        const float hxh = 0.5f * cf.hx, hyh = 0.5f * cf.hy, hzh = 0.5f * cf.hz;
        const size_t P1 = (size_t)a.pitch1;
        const size_t cx = (size_t)z * ny * P1 + (size_t)y * P1 + (size_t)x;
        const size_t cy = ((size_t)z * (ny + 1) + (size_t)y) * P + (size_t)x;
        const float um = a.vx[cx], up = a.vx[cx + 1];
        const float vm = a.vy[cy], vp = a.vy[cy + P];
        const float wm = a.vz[c], wp = a.vz[c + plane];
        const float su = um + up, sv = vm + vp, sw = wm + wp;
        if constexpr (ST == GEN_UPWIND) {
            // with amask a face to a solid carries no advection (NMR_2d, cu:190-210)
            bool open[6];
#pragma unroll
            for (int s = 0; s < 6; s++) open[s] = !a.amask || kn[s] != 0.f;
            w[0] = (um != 0.f && open[0]) ? fmaf(hxh, fmaxf(su, 0.f), w[0]) : w[0];
            w[1] = (up != 0.f && open[1]) ? fmaf(-hxh, fminf(su, 0.f), w[1]) : w[1];
            w[2] = (vm != 0.f && open[2]) ? fmaf(hyh, fmaxf(sv, 0.f), w[2]) : w[2];
            w[3] = (vp != 0.f && open[3]) ? fmaf(-hyh, fminf(sv, 0.f), w[3]) : w[3];
            w[4] = (wm != 0.f && open[4]) ? fmaf(hzh, fmaxf(sw, 0.f), w[4]) : w[4];
            w[5] = (wp != 0.f && open[5]) ? fmaf(-hzh, fminf(sw, 0.f), w[5]) : w[5];
        } else if constexpr (ST == GEN_PREDICT) { // + u hx (f_{i-1} - f_i); kn is 0 or 1
            w[0] = fmaf(kn[0] * hxh, su, w[0]);
            w[2] = fmaf(kn[2] * hyh, sv, w[2]);
            w[4] = fmaf(kn[4] * hzh, sw, w[4]);
        } else { // - u hx (f_{i+1} - f_i)
            w[1] = fmaf(-kn[1] * hxh, su, w[1]);
            w[3] = fmaf(-kn[3] * hyh, sv, w[3]);
            w[5] = fmaf(-kn[5] * hzh, sw, w[5]);
        }
    }
    const float nsx = 2.f - kn[0] - kn[1], nsy = 2.f - kn[2] - kn[3], nsz = 2.f - kn[4] - kn[5];
    const float rate2 = a.r2v ? a.r2v[c] : fmaf(cf.s2z, nsz, fmaf(cf.s2y, nsy, fmaf(cf.s2x, nsx, cf.r2)));
    const float* const* src = (ST == GEN_CORRECT) ? a.p : a.m;
    float m0[NC], d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const float m = src[q][c];
        m0[q] = m;
        float f[6];
#pragma unroll
        for (int s = 0; s < 6; s++) {
            if (kind[s] == 1) f[s] = m;
            else if (kind[s] == 2) f[s] = 0.f;
            else if (ST != GEN_UPWIND && kn[s] == 0.f) f[s] = m; // solid reads as own value
            else f[s] = src[q][ni[s]];
        }
        float t = w[0] * (f[0] - m);
        t = fmaf(w[1], f[1] - m, t);
        t = fmaf(w[2], f[2] - m, t);
        t = fmaf(w[3], f[3] - m, t);
        t = fmaf(w[4], f[4] - m, t);
        d[q] = fmaf(w[5], f[5] - m, t);
    }
    // precession angles: the gradient field's, or along axis bnax the
    // per-voxel field's (value selects, as in gupdate_chunk)
    const int bnax = a.bn ? a.bnax : -1;
    const float tb = a.bn ? cf.dtg * a.bn[c] : 0.f;
    const float thy = bnax == 1 ? tb : cf.bgy * ((float)y + 0.5f);
    const float thz = bnax == 2 ? tb : cf.bgz * ((float)z + 0.5f);
    if constexpr (NC == 3) {
        const float rate1 = a.r1v ? a.r1v[c] : fmaf(cf.s1z, nsz, fmaf(cf.s1y, nsy, fmaf(cf.s1x, nsx, cf.r1)));
        const float par = fmaf(cf.n2, m0[2], fmaf(cf.n1, m0[1], cf.n0 * m0[0]));
        const float nn[3] = {cf.n0, cf.n1, cf.n2};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float mpar = nn[q] * par;
            d[q] = fmaf(rate1, fmaf(cf.peq, nn[q], -mpar), d[q]);
            d[q] = fmaf(rate2, mpar - m0[q], d[q]);
        }
        const float thx = bnax == 0 ? tb : cf.bgx * ((float)x + 0.5f);
        d[0] = fmaf(thz, m0[1], fmaf(-thy, m0[2], d[0]));
        d[1] = fmaf(-thz, m0[0], fmaf(thx, m0[2], d[1]));
        d[2] = fmaf(thy, m0[0], fmaf(-thx, m0[1], d[2]));
    } else {
#pragma unroll
        for (int q = 0; q < NC; q++) d[q] = fmaf(-rate2, m0[q], d[q]);
        if constexpr (NC == 2) {
            d[0] = fmaf(thz, m0[1], d[0]);
            d[1] = fmaf(-thz, m0[0], d[1]);
        }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) {
        if constexpr (ST == GEN_CORRECT) a.out[q][c] = 0.5f * (a.m[q][c] + (m0[q] + d[q]));
        else a.out[q][c] = m0[q] + d[q];
    }
}

// Per-plane scalars of the exact precession rotation.
template <bool EXACT>
__device__ __forceinline__ void plane_rotation(float thz, float& sn, float& cm1)
{
    sn = 0.f;
    cm1 = 0.f;
    if constexpr (EXACT) {
        sn = sinf(thz);
        const float h = sinf(0.5f * thz);
        cm1 = -2.f * h * h; // cos(th) - 1 without cancellation
    }
}

// Signal of one updated chunk: pores inside the trim window.
template <int NC>
__device__ __forceinline__ void accumulate(double (&acc)[NC], const f4 (&out)[NC], uint32_t kc, const bool (&tx)[4])
{
#pragma unroll
    for (int e = 0; e < 4; e++)
        if (tx[e] && ((kc >> (8 * e)) & 0xffu) != 0) {
#pragma unroll
            for (int q = 0; q < NC; q++) acc[q] += (double)out[q][e];
        }
}

// Phase-weighted pore sums of the chunk: acc[q] += v*w1, acc[NC+q] += v*w2
// (FP32 operands, exact FP64 products, as field_sum_kernel forms them).
template <int NC>
__device__ __forceinline__ void accumulate_w(double (&acc)[2 * NC], const f4 (&out)[NC], uint32_t kc,
                                             const bool (&tx)[4], const f4& w1, const f4& w2)
{
#pragma unroll
    for (int e = 0; e < 4; e++)
        if (tx[e] && ((kc >> (8 * e)) & 0xffu) != 0) {
            const double a1 = (double)w1[e], a2 = (double)w2[e];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const double v = (double)out[q][e];
                acc[q] += v * a1;
                acc[NC + q] += v * a2;
            }
        }
}

// Workgroup sum of acc into part[blockIdx.x*NC + q] (fixed order).
template <int NC, int NT>
__device__ __forceinline__ void block_reduce(double (&acc)[NC], double (*sred)[NC], double* dst)
{
#pragma unroll
    for (int q = 0; q < NC; q++) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[q] = v;
    }
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NC; q++) sred[wave][q] = acc[q];
    }
    __syncthreads();
    if (tid < NC) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) s += sred[w][tid];
        dst[tid] = s;
    }
}

// One tile of the one-step sweep: tile index t (x fastest, then y, then the
// z-chunk).  Adds the signal of the tile's updated pore voxels into acc.
// BRK: the float arrays are stored as bricks of 128 bytes holding 8 x-voxels
// of 4 rows, so that a DRAM line holds a compact block of voxels (all-solid
// lines are then never fetched) and a wavefront of 16 lanes x 4 rows reads
// whole lines.
template <int NC, bool VEL, bool EXACT, int LX, int RPT, bool BRK = false, bool KEEP = false>
__device__ __forceinline__ void sweep_tile(const SweepArgs& a, const int t, double (&acc)[NC])
{
    constexpr int NT = 256;
    constexpr int RY = NT / LX;  // thread rows
    constexpr int TX = 4 * LX;   // tile width (voxels)
    constexpr int TY = RY * RPT; // tile height (rows)
    constexpr int LW = TX + 8;   // LDS row: col 3 = x0-1, 4..4+TX-1 tile, 4+TX = x0+TX
    constexpr int LH = TY + 2;   // row 0 = y0-1, 1..TY tile, TY+1 = y0+TY

    // Planes k-1 and k of the tile plus their one-voxel halo, in a ring of
    // three slots; ghost cells of the domain boundary hold the
    // zero-derivative copy, so the update reads neighbours without bounds
    // tests.  Three slots let the update read planes k-1 and k from LDS
    // (only the plane in flight, k+1, stays in registers) with one barrier
    // per plane: the slot written for plane k was last read two planes ago.
    __shared__ __attribute__((aligned(16))) float sm[3][NC][LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t sk[3][LH][LW];

    const int tid = threadIdx.x;
    const int lx = tid % LX;
    const int ry = tid / LX;

    const int tx = t % a.ntx;
    const int tyz = t / a.ntx;
    const int ty = tyz % a.nty;
    const int tz = tyz / a.nty;
    const int x0 = tx * TX, y0 = ty * TY;
    const int k0 = tz * a.lz;
    const int k1 = min(k0 + a.lz, a.nz);
    const int ix = x0 + 4 * lx;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const bool xin = ix < nx;
    const int xlast = nx - 1 - ix;             // last valid element of this chunk (>= 3: all valid)
    const bool xend = xin && xlast <= 3;       // chunk holds voxel nx-1
    const bool xpart = xin && xlast < 3;       // chunk reaches past nx: elements > xlast take the ghost value
    const bool xbeg = (ix == 0);               // chunk holds voxel 0
    const int c = 4 + 4 * lx;                  // LDS column of element 0

    const uint32_t Pb = (uint32_t)a.pitch * 4u;      // bytes per row (float arrays)
    const uint32_t plane_b = (uint32_t)ny * Pb;      // bytes per plane
    const uint32_t plane_m = (uint32_t)ny * (uint32_t)a.pitch;
    // byte offsets of the float arrays along x, y, z (they add up): rows of
    // pitch floats, or 128-byte bricks of 8 x-voxels by 4 rows
    const uint32_t nbx = ((uint32_t)a.pitch + 7u) >> 3; // bricks per row
    auto xo = [&](int x) -> uint32_t {
        return BRK ? ((uint32_t)(x >> 3) << 7) + ((uint32_t)(x & 7) << 2) : (uint32_t)x * 4u;
    };
    auto yo = [&](int y) -> uint32_t {
        return BRK ? (uint32_t)(y >> 2) * nbx * 128u + ((uint32_t)(y & 3) << 5) : (uint32_t)y * Pb;
    };
    auto zo = [&](int z) -> uint32_t {
        return BRK ? (uint32_t)z * (uint32_t)a.nby * nbx * 128u : (uint32_t)z * plane_b;
    };

    int iy[RPT];
    bool rin[RPT];
    uint32_t rowf[RPT], rowm[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        iy[r] = y0 + ry + r * RY;
        rin[r] = xin && (iy[r] < ny);
        rowf[r] = yo(iy[r]) + xo(ix);
        rowm[r] = (uint32_t)iy[r] * (uint32_t)a.pitch + (uint32_t)ix;
    }
    const bool top = (ry == 0);
    const bool bot = (ry == RY - 1);
    const int hy = top ? (y0 - 1) : (y0 + TY);
    const bool hrow = (top || bot) && xin && hy >= 0 && hy < ny;
    const uint32_t hrowf = yo(hrow ? hy : 0) + xo(ix);
    const uint32_t hrowm = (uint32_t)(hrow ? hy : 0) * (uint32_t)a.pitch + (uint32_t)ix;
    const bool hcl = (lx == 0) && x0 > 0;
    const bool hcr = (lx == LX - 1) && x0 + TX < nx;
    const int hcx = (lx == 0) ? x0 - 1 : x0 + TX;

    // All-solid 4-voxel chunks never change, so their stores and velocity
    // loads are skipped; when no pore-solid face carries velocity (a.inert),
    // their magnetization is never read by a pore either and is not loaded.
    // The mask is therefore loaded two planes ahead of the magnetization.
    const bool skip_m = a.inert != 0;
    f4 mp[RPT][NC], mc[RPT][NC], mn[RPT][NC];
    uint32_t kp[RPT], kc[RPT], kn[RPT], kq[RPT];
    f4 hc[NC], hn[NC];
    uint32_t hkc = 0, hkn = 0, hkq = 0;
    float xc[RPT][NC], xn[RPT][NC];
    uint32_t kxc[RPT], kxn[RPT];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        hc[q] = f4{0.f, 0.f, 0.f, 0.f};
        hn[q] = hc[q];
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        kxc[r] = kxn[r] = 0;
#pragma unroll
        for (int q = 0; q < NC; q++) xc[r][q] = xn[r][q] = 0.f;
    }
    auto clampz = [&](int kk) { return kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk); };

    auto load_mask = [&](int kk, uint32_t (&kw)[RPT]) {
        const uint32_t pm = (uint32_t)clampz(kk) * plane_m;
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            kw[r] = 0;
            if (rin[r]) kw[r] = ldw_nt(a.mask, pm + rowm[r]);
        }
    };
    auto load_m = [&](int kk, const uint32_t (&kw)[RPT], f4 (&m)[RPT][NC]) {
        const uint32_t pf = zo(clampz(kk));
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (rin[r] && (!skip_m || kw[r] != 0)) {
#pragma unroll
                for (int q = 0; q < NC; q++) m[r][q] = ld4_state<KEEP>(a.min[q], pf + rowf[r]);
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) m[r][q] = f4{0.f, 0.f, 0.f, 0.f};
            }
        }
    };
    auto load_hmask = [&](int kk) -> uint32_t {
        if (!hrow) return 0;
        return *reinterpret_cast<const uint32_t*>(a.mask + (uint32_t)clampz(kk) * plane_m + hrowm);
    };
    auto load_hm = [&](int kk, uint32_t hk, f4 (&h)[NC]) {
        if (hrow && (!skip_m || hk != 0)) {
            const uint32_t pf = zo(clampz(kk));
#pragma unroll
            for (int q = 0; q < NC; q++) h[q] = ld4(a.min[q], pf + hrowf);
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) h[q] = f4{0.f, 0.f, 0.f, 0.f};
        }
    };
    // Elements past nx take the ghost value (registers and LDS agree).  The
    // loads above return raw rows; the fix is applied where a plane is
    // rotated in, one iteration after its loads were issued, so that no
    // wait for memory follows a load directly.
    auto fix_mask = [&](uint32_t (&kw)[RPT]) {
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (xpart) kw[r] = clamp_tail_w(kw[r], xlast);
    };
    auto fix_m = [&](f4 (&m)[RPT][NC]) {
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (xpart) {
#pragma unroll
                for (int q = 0; q < NC; q++) m[r][q] = clamp_tail(m[r][q], xlast);
            }
    };
    auto fix_h = [&](uint32_t& hk, f4 (&h)[NC]) {
        if (xpart) {
            hk = clamp_tail_w(hk, xlast);
#pragma unroll
            for (int q = 0; q < NC; q++) h[q] = clamp_tail(h[q], xlast);
        }
    };
    auto load_xhalo = [&](int kk, float (&xv)[RPT][NC], uint32_t (&kx)[RPT]) {
        if (hcl || hcr) {
            const uint32_t pf = zo(clampz(kk)), pm = (uint32_t)clampz(kk) * plane_m;
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                if (iy[r] < ny) {
                    const uint32_t of = pf + yo(iy[r]) + xo(hcx);
#pragma unroll
                    for (int q = 0; q < NC; q++) xv[r][q] = ld1(a.min[q], of);
                    kx[r] = a.mask[pm + (uint32_t)iy[r] * (uint32_t)a.pitch + (uint32_t)hcx];
                }
            }
        }
    };

    // upwind velocity sums of a plane's chunks
    auto load_vel = [&](int kk, const uint32_t (&kw)[RPT], f4 (&ax)[RPT], f4 (&ay)[RPT], f4 (&az)[RPT]) {
        const uint32_t pz = zo(kk);
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (rin[r] && kw[r] != 0) {
                ax[r] = ld4_nt(a.adv[0], pz + rowf[r]);
                ay[r] = ld4_nt(a.adv[1], pz + rowf[r]);
                az[r] = ld4_nt(a.adv[2], pz + rowf[r]);
            } else {
                ax[r] = ay[r] = az[r] = f4{0.f, 0.f, 0.f, 0.f};
            }
        }
    };
    f4 ax4[RPT], ay4[RPT], az4[RPT];

    // prologue: masks of planes k0-1, k0, k0+1; magnetization of k0-1, k0
    load_mask(k0 - 1, kp);
    load_mask(k0, kc);
    load_mask(k0 + 1, kn);
    hkc = load_hmask(k0);
    hkn = load_hmask(k0 + 1);
    fix_mask(kp);
    fix_mask(kc);
    fix_mask(kn);
    if (xpart) {
        hkc = clamp_tail_w(hkc, xlast);
        hkn = clamp_tail_w(hkn, xlast);
    }
    load_m(k0 - 1, kp, mp);
    load_m(k0, kc, mc);
    load_hm(k0, hkc, hc);
    load_xhalo(k0, xc, kxc);
    fix_m(mp);
    fix_m(mc);
    if (xpart) {
#pragma unroll
        for (int q = 0; q < NC; q++) hc[q] = clamp_tail(hc[q], xlast);
    }
    // Retire the prologue's loads here (vmcnt(0)): otherwise the wait the
    // compiler places for them lands inside the loop, in front of the LDS
    // staging, where it would also wait for the loads just issued for the
    // next plane.
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // plane k0-1 into slot 0 (its chunks only: the update reads it as the z
    // neighbour at the same position)
#pragma unroll
    for (int r = 0; r < RPT; r++)
        if (rin[r]) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[0][q][1 + ry + r * RY][c]) = mp[r][q];
        }

    const Coef cf = a.c;
    bool trim_x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) trim_x[e] = (ix + e >= a.bx0) && (ix + e < a.bx1);

    int pbuf = 0, buf = 1; // slots of planes k-1 and k
    for (int k = k0; k < k1; k++) {
        // prefetch: masks of plane k+2, magnetization of plane k+1
        load_mask(k + 2, kq);
        hkq = load_hmask(k + 2);
        load_m(k + 1, kn, mn);
        load_hm(k + 1, hkn, hn);
        load_xhalo(k + 1, xn, kxn);
        // velocity faces of plane k (only for chunks holding a pore)
        if constexpr (VEL) load_vel(k, kc, ax4, ay4, az4);

        // stage plane k (with ghosts) into LDS
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (!rin[r]) continue;
            const int lr = 1 + ry + r * RY;
            f4 v[NC];
            const uint32_t w = kc[r];
#pragma unroll
            for (int q = 0; q < NC; q++) v[q] = mc[r][q];
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][lr][c]) = v[q];
            *reinterpret_cast<uint32_t*>(&sk[buf][lr][c]) = w;
            if (xend) { // ghost column nx (inside this chunk or the next column)
                const int xl = xlast < 3 ? xlast : 3;
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][c + 4] = v[q][xl];
                sk[buf][lr][c + 4] = (uint8_t)((w >> (8 * xl)) & 0xffu);
            }
            if (xbeg) { // ghost column -1
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][c - 1] = v[q].x;
                sk[buf][lr][c - 1] = (uint8_t)(w & 0xffu);
            }
            if (hcl || hcr) {
                const int col = hcl ? c - 1 : c + 4;
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][col] = xc[r][q];
                sk[buf][lr][col] = (uint8_t)kxc[r];
            }
            if (iy[r] == 0) { // ghost row -1
#pragma unroll
                for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][lr - 1][c]) = v[q];
                *reinterpret_cast<uint32_t*>(&sk[buf][lr - 1][c]) = w;
            }
            if (iy[r] == ny - 1) { // ghost row ny
#pragma unroll
                for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][lr + 1][c]) = v[q];
                *reinterpret_cast<uint32_t*>(&sk[buf][lr + 1][c]) = w;
            }
        }
        if (hrow) {
            const int lr = top ? 0 : TY + 1;
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][lr][c]) = hc[q];
            *reinterpret_cast<uint32_t*>(&sk[buf][lr][c]) = hkc;
        }
        __syncthreads();

        // per-plane scalars
        const bool trim_z = k >= a.bz0 && k < a.bz1;
        const float thz = cf.bgz * ((float)k + 0.5f);
        float sn, cm1;
        plane_rotation<EXACT>(thz, sn, cm1);

#pragma unroll
        for (int r = 0; r < RPT; r++) {
            // all-solid chunks keep their value: no update (wavefronts whose
            // chunks are all solid skip the arithmetic)
            if (!rin[r] || kc[r] == 0) continue;
            const int lr = 1 + ry + r * RY;
            const int y = iy[r];
            const bool trim_y = y >= a.by0 && y < a.by1;
            const float thy = cf.bgy * ((float)y + 0.5f);
            f4 ym[NC], yp[NC];
            float xl[NC], xr[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                ym[q] = *reinterpret_cast<const f4*>(&sm[buf][q][lr - 1][c]);
                yp[q] = *reinterpret_cast<const f4*>(&sm[buf][q][lr + 1][c]);
                xl[q] = sm[buf][q][lr][c - 1];
                xr[q] = sm[buf][q][lr][c + 4];
            }
            const uint32_t kym = *reinterpret_cast<const uint32_t*>(&sk[buf][lr - 1][c]);
            const uint32_t kyp = *reinterpret_cast<const uint32_t*>(&sk[buf][lr + 1][c]);
            const uint32_t kxm = (kc[r] << 8) | (uint32_t)sk[buf][lr][c - 1];
            const uint32_t kxp = (kc[r] >> 8) | ((uint32_t)sk[buf][lr][c + 4] << 24);
            f4 out[NC], mcl[NC], mpl[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                mcl[q] = *reinterpret_cast<const f4*>(&sm[buf][q][lr][c]);
                mpl[q] = *reinterpret_cast<const f4*>(&sm[pbuf][q][lr][c]);
            }
            update_chunk<NC, VEL, EXACT>(cf, mcl, mpl, mn[r], ym, yp, xl, xr, kc[r], kxm, kxp, kym, kyp, kp[r],
                                         kn[r], ax4[r], ay4[r], az4[r], ix, thy, thz, sn,
                                         cm1, out);
            if (trim_z && trim_y) accumulate<NC>(acc, out, kc[r], trim_x);
            const uint32_t of = zo(k) + rowf[r];
#pragma unroll
            for (int q = 0; q < NC; q++) st4_state<KEEP>(a.mout[q], of, out[q]);
        }

        // rotate the register planes and the LDS slots
#pragma unroll
        for (int r = 0; r < RPT; r++) {
#pragma unroll
            for (int q = 0; q < NC; q++) {
                mc[r][q] = mn[r][q];
                xc[r][q] = xn[r][q];
            }
            kp[r] = kc[r];
            kc[r] = kn[r];
            kn[r] = kq[r];
            kxc[r] = kxn[r];
        }
        fix_m(mc);
        fix_mask(kn);
        fix_h(hkq, hn);
#pragma unroll
        for (int q = 0; q < NC; q++) hc[q] = hn[q];
        hkc = hkn;
        hkn = hkq;
        pbuf = buf;
        buf = buf == 2 ? 0 : buf + 1;
    }
}

// Compacted brick sweep: tile 64 x 16 (one 4-voxel chunk of a row per
// thread), the arithmetic of update_chunk on exactly the values the tile
// sweep passes it, so the fields are the same bits.
//
// In the porous media here about 28 % of the 4-voxel chunks hold a pore,
// but 97 % of the wavefronts of the tile sweep hold at least one, so the
// update runs with about 28 % of its lanes busy.  Here every thread stages
// its own chunk of each plane into a ring of four LDS planes (k-1, k, k+1
// and the next), the chunks of a plane that hold a pore are listed (wave
// ballots, a prefix over the waves) and the list is handed out to as few
// wavefronts as it needs, 64 chunks each; a worker lane reads its chunk and
// the six neighbour chunks from LDS.  The list of plane k+2 is built while
// plane k is updated, so a worker knows its chunk of plane k+1 when the
// update of plane k ends and loads that chunk's velocity faces right after
// it, across the next plane's staging.  Which wavefront takes
// the first 64 chunks rotates with the workgroup's slot on its CU, spreading
// the update over the SIMDs.  One barrier per plane: the ring has a slot more than the
// three planes an update reads, and the lists and counts three slots.
template <int NC, bool VEL, bool EXACT>
__device__ __forceinline__ void csweep_tile(const SweepArgs& a, const int t, double (&acc)[NC])
{
    constexpr int LX = 16, TX = 64, TY = 16, NT = 256;
    constexpr int LW = TX + 8; // LDS row: col 3 = x0-1, 4..4+TX-1 tile, 4+TX = x0+TX
    constexpr int LH = TY + 2; // row 0 = y0-1, 1..TY tile, TY+1 = y0+TY
    __shared__ __attribute__((aligned(16))) float sm[4][NC][LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t sk[4][LH][LW];
    __shared__ uint8_t slist[3][NT];
    __shared__ int scnt[3][4];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lx = tid % LX, ry = tid / LX;
    const int tx = t % a.ntx;
    const int tyz = t / a.ntx;
    const int ty = tyz % a.nty;
    const int tz = tyz / a.nty;
    const int x0 = tx * TX, y0 = ty * TY;
    const int k0 = tz * a.lz;
    const int k1 = min(k0 + a.lz, a.nz);
    const int nx = a.nx, ny = a.ny, nz = a.nz;

    const uint32_t plane_m = (uint32_t)ny * (uint32_t)a.pitch;
    const uint32_t nbx = ((uint32_t)a.pitch + 7u) >> 3;
    auto xo = [&](int x) -> uint32_t { return ((uint32_t)(x >> 3) << 7) + ((uint32_t)(x & 7) << 2); };
    auto yo = [&](int y) -> uint32_t { return (uint32_t)(y >> 2) * nbx * 128u + ((uint32_t)(y & 3) << 5); };
    auto zo = [&](int z) -> uint32_t { return (uint32_t)z * (uint32_t)a.nby * nbx * 128u; };
    auto clampz = [&](int kk) { return kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk); };
    auto slot = [](int p) { return (p + 4) & 3; }; // p >= -1
    auto lst = [](int p) { return p % 3; };        // p >= 0

    // ---- the owner's chunk (staging) ----
    const int ix = x0 + 4 * lx, iy = y0 + ry;
    const bool rin = ix < nx && iy < ny;
    const int xlast = nx - 1 - ix;
    const bool xend = rin && xlast <= 3;
    const bool xpart = rin && xlast < 3;
    const bool xbeg = ix == 0;
    const int c = 4 + 4 * lx, lr = 1 + ry;
    const uint32_t rowf = yo(iy < ny ? iy : 0) + xo(ix);
    const uint32_t rowm = (uint32_t)(iy < ny ? iy : 0) * (uint32_t)a.pitch + (uint32_t)ix;
    const bool top = ry == 0, bot = ry == TY - 1;
    const int hy = top ? (y0 - 1) : (y0 + TY);
    const bool hrow = (top || bot) && ix < nx && hy >= 0 && hy < ny;
    const uint32_t hrowf = yo(hrow ? hy : 0) + xo(ix);
    const uint32_t hrowm = (uint32_t)(hrow ? hy : 0) * (uint32_t)a.pitch + (uint32_t)ix;
    const bool hcl = lx == 0 && x0 > 0;
    const bool hcr = lx == LX - 1 && x0 + TX < nx;
    const int hcx = lx == 0 ? x0 - 1 : x0 + TX;
    const bool skip_m = a.inert != 0;

    auto load_mask = [&](int kk) -> uint32_t {
        return rin ? ldw_nt(a.mask, (uint32_t)clampz(kk) * plane_m + rowm) : 0u;
    };
    auto load_hmask = [&](int kk) -> uint32_t {
        return hrow ? *reinterpret_cast<const uint32_t*>(a.mask + (uint32_t)clampz(kk) * plane_m + hrowm) : 0u;
    };
    auto load_m = [&](int kk, uint32_t kw, f4 (&m)[NC]) {
        const uint32_t pf = zo(clampz(kk));
#pragma unroll
        for (int q = 0; q < NC; q++) m[q] = (rin && (!skip_m || kw != 0)) ? ld4_nt(a.min[q], pf + rowf) : f4{0.f, 0.f, 0.f, 0.f};
    };
    auto load_hm = [&](int kk, uint32_t hk, f4 (&h)[NC]) {
        const uint32_t pf = zo(clampz(kk));
#pragma unroll
        for (int q = 0; q < NC; q++) h[q] = (hrow && (!skip_m || hk != 0)) ? ld4(a.min[q], pf + hrowf) : f4{0.f, 0.f, 0.f, 0.f};
    };
    auto load_xhalo = [&](int kk, float (&xv)[NC], uint32_t& kx) {
        if ((hcl || hcr) && iy < ny) {
            const int zz = clampz(kk);
            const uint32_t of = zo(zz) + yo(iy) + xo(hcx);
#pragma unroll
            for (int q = 0; q < NC; q++) xv[q] = ld1(a.min[q], of);
            kx = a.mask[(uint32_t)zz * plane_m + (uint32_t)iy * (uint32_t)a.pitch + (uint32_t)hcx];
        }
    };
    // Elements past nx take the ghost value (the last voxel's).
    auto fix_w = [&](uint32_t w) { return xpart ? clamp_tail_w(w, xlast) : w; };
    auto fix_v = [&](f4 (&m)[NC]) {
#pragma unroll
        for (int q = 0; q < NC; q++)
            if (xpart) m[q] = clamp_tail(m[q], xlast);
    };
    // Plane p (values v, mask w, halo row h / hk, halo column xv / kx) into
    // its ring slot with the ghost cells of the domain boundary.
    auto stage = [&](int p, const f4 (&v)[NC], uint32_t w, const f4 (&h)[NC], uint32_t hk, const float (&xv)[NC],
                     uint32_t kx) {
        const int s = slot(p);
        if (rin) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[s][q][lr][c]) = v[q];
            *reinterpret_cast<uint32_t*>(&sk[s][lr][c]) = w;
            if (xend) { // ghost column nx (inside this chunk or the next column)
                const int xl = xlast < 3 ? xlast : 3;
#pragma unroll
                for (int q = 0; q < NC; q++) sm[s][q][lr][c + 4] = v[q][xl];
                sk[s][lr][c + 4] = (uint8_t)((w >> (8 * xl)) & 0xffu);
            }
            if (xbeg) { // ghost column -1
#pragma unroll
                for (int q = 0; q < NC; q++) sm[s][q][lr][c - 1] = v[q].x;
                sk[s][lr][c - 1] = (uint8_t)(w & 0xffu);
            }
            if (hcl || hcr) {
                const int col = hcl ? c - 1 : c + 4;
#pragma unroll
                for (int q = 0; q < NC; q++) sm[s][q][lr][col] = xv[q];
                sk[s][lr][col] = (uint8_t)kx;
            }
            if (iy == 0) { // ghost row -1
#pragma unroll
                for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[s][q][lr - 1][c]) = v[q];
                *reinterpret_cast<uint32_t*>(&sk[s][lr - 1][c]) = w;
            }
            if (iy == ny - 1) { // ghost row ny
#pragma unroll
                for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[s][q][lr + 1][c]) = v[q];
                *reinterpret_cast<uint32_t*>(&sk[s][lr + 1][c]) = w;
            }
        }
        if (hrow) {
            const int hr = top ? 0 : TY + 1;
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[s][q][hr][c]) = h[q];
            *reinterpret_cast<uint32_t*>(&sk[s][hr][c]) = hk;
        }
    };
    // Count of plane p's chunks that hold a pore (p inside the z-chunk), per wave.
    auto ballot_count = [&](int p, uint32_t w) -> unsigned long long {
        const bool act = p < k1 && rin && w != 0u;
        const unsigned long long b = __ballot(act);
        if (lane == 0) scnt[lst(p)][wave] = (int)__popcll(b);
        return b;
    };
    // After the barrier: this owner's place in plane p's list.
    auto list_write = [&](int p, unsigned long long b) {
        if ((b >> lane) & 1ull) {
            int base = 0;
            for (int w2 = 0; w2 < wave; w2++) base += scnt[lst(p)][w2];
            const int pos = base + (int)__popcll(b & ((1ull << lane) - 1ull));
            slist[lst(p)][pos] = (uint8_t)tid;
        }
    };

    // ---- the worker's chunk ----
    // The workgroups resident on one CU share its four SIMDs; the one that
    // takes a workgroup's first 64 chunks rotates with the workgroup's slot
    // on the CU (TG_ID, bits 19:16 of the HW_ID register).
    const int rot = (int)(__builtin_amdgcn_s_getreg(4 | (16 << 6) | (3 << 11)) & 3u);
    const int wslot = ((wave - rot) & 3) * 64 + lane;
    auto assign = [&](int p) -> int { // chunk of plane p this lane updates, or -1
        if (p >= k1) return -1;
        const int n = scnt[lst(p)][0] + scnt[lst(p)][1] + scnt[lst(p)][2] + scnt[lst(p)][3];
        return wslot < n ? (int)slist[lst(p)][wslot] : -1;
    };
    struct Faces {
        f4 ax, ay, az; // upwind velocity sums of the worker's chunk
    };
    auto load_faces = [&](int p, int cid, Faces& f) {
        f.ax = f.ay = f.az = f4{0.f, 0.f, 0.f, 0.f};
        if (VEL && cid >= 0) {
            const int wx = x0 + 4 * (cid % LX), wy = y0 + cid / LX;
            const uint32_t of = zo(p) + yo(wy) + xo(wx);
            f.ax = ld4_nt(a.adv[0], of);
            f.ay = ld4_nt(a.adv[1], of);
            f.az = ld4_nt(a.adv[2], of);
        }
    };

    // ---- prologue: planes k0-1 and k0 staged, plane k0+1 in registers,
    // the lists of planes k0 and k0+1, the faces of plane k0 ----
    f4 m1[NC], m2[NC], h1[NC], h2[NC], mt[NC], ht[NC];
    float x1[NC], x2[NC], xt[NC];
    uint32_t kx1 = 0, kx2 = 0, kxt = 0;
#pragma unroll
    for (int q = 0; q < NC; q++) x1[q] = x2[q] = xt[q] = 0.f;
    uint32_t wa = load_mask(k0 - 1), wb = load_mask(k0), k1w = load_mask(k0 + 1), k2w = load_mask(k0 + 2);
    uint32_t hka = load_hmask(k0 - 1), hkb = load_hmask(k0), hk1 = load_hmask(k0 + 1), hk2 = load_hmask(k0 + 2);
    load_m(k0 - 1, wa, mt);
    load_hm(k0 - 1, hka, ht);
    load_xhalo(k0 - 1, xt, kxt);
    load_m(k0, wb, m2);
    load_hm(k0, hkb, h2);
    load_xhalo(k0, x2, kx2);
    load_m(k0 + 1, k1w, m1);
    load_hm(k0 + 1, hk1, h1);
    load_xhalo(k0 + 1, x1, kx1);
    wa = fix_w(wa);
    wb = fix_w(wb);
    k1w = fix_w(k1w);
    hka = fix_w(hka);
    hkb = fix_w(hkb);
    hk1 = fix_w(hk1);
    fix_v(mt);
    fix_v(ht);
    fix_v(m2);
    fix_v(h2);
    fix_v(m1);
    fix_v(h1);
    stage(k0 - 1, mt, wa, ht, hka, xt, kxt);
    stage(k0, m2, wb, h2, hkb, x2, kx2);
    const unsigned long long b0 = ballot_count(k0, wb);
    const unsigned long long b1 = ballot_count(k0 + 1, k1w);
    __syncthreads();
    list_write(k0, b0);
    list_write(k0 + 1, b1);
    __syncthreads();
    int cid = assign(k0);
    Faces f0;
    load_faces(k0, cid, f0);

    const Coef cf = a.c;
    for (int k = k0; k < k1; k++) {
        // 1. stage plane k+1; 2. count plane k+2's pore chunks
        stage(k + 1, m1, k1w, h1, hk1, x1, kx1);
        const unsigned long long b2 = ballot_count(k + 2, k2w);
        // 3. loads: mask of plane k+3, plane k+2
        const uint32_t k3w = load_mask(k + 3), hk3 = load_hmask(k + 3);
        load_m(k + 2, k2w, m2);
        load_hm(k + 2, hk2, h2);
        load_xhalo(k + 2, x2, kx2);
        __syncthreads();
        // 5. plane k+2's list; 6. this lane's chunk of plane k+1 (its faces
        // are loaded after the update, so they do not hold registers during it)
        list_write(k + 2, b2);
        const int cid1 = assign(k + 1);

        // 7. update of plane k
        if (cid >= 0) {
            const int wlx = cid % LX, wry = cid / LX;
            const int wx = x0 + 4 * wlx, wy = y0 + wry;
            const int wc_ = 4 + 4 * wlx, wlr = 1 + wry;
            const int s0 = slot(k), sp = slot(k - 1), sn_ = slot(k + 1);
            f4 mc[NC], mp[NC], mn[NC], ym[NC], yp[NC];
            float xl[NC], xr[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                mc[q] = *reinterpret_cast<const f4*>(&sm[s0][q][wlr][wc_]);
                mp[q] = *reinterpret_cast<const f4*>(&sm[sp][q][wlr][wc_]);
                mn[q] = *reinterpret_cast<const f4*>(&sm[sn_][q][wlr][wc_]);
                ym[q] = *reinterpret_cast<const f4*>(&sm[s0][q][wlr - 1][wc_]);
                yp[q] = *reinterpret_cast<const f4*>(&sm[s0][q][wlr + 1][wc_]);
                xl[q] = sm[s0][q][wlr][wc_ - 1];
                xr[q] = sm[s0][q][wlr][wc_ + 4];
            }
            const uint32_t kc = *reinterpret_cast<const uint32_t*>(&sk[s0][wlr][wc_]);
            const uint32_t kp = *reinterpret_cast<const uint32_t*>(&sk[sp][wlr][wc_]);
            const uint32_t kn = *reinterpret_cast<const uint32_t*>(&sk[sn_][wlr][wc_]);
            const uint32_t kym = *reinterpret_cast<const uint32_t*>(&sk[s0][wlr - 1][wc_]);
            const uint32_t kyp = *reinterpret_cast<const uint32_t*>(&sk[s0][wlr + 1][wc_]);
            const uint32_t kxm = (kc << 8) | (uint32_t)sk[s0][wlr][wc_ - 1];
            const uint32_t kxp = (kc >> 8) | ((uint32_t)sk[s0][wlr][wc_ + 4] << 24);
            const float thz = cf.bgz * ((float)k + 0.5f);
            float sn, cm1;
            plane_rotation<EXACT>(thz, sn, cm1);
            const float thy = cf.bgy * ((float)wy + 0.5f);
            f4 out[NC];
            update_chunk<NC, VEL, EXACT>(cf, mc, mp, mn, ym, yp, xl, xr, kc, kxm, kxp, kym, kyp, kp, kn, f0.ax,
                                         f0.ay, f0.az, wx, thy, thz, sn, cm1, out);
            if (k >= a.bz0 && k < a.bz1 && wy >= a.by0 && wy < a.by1) {
                bool trim_x[4];
#pragma unroll
                for (int e = 0; e < 4; e++) trim_x[e] = (wx + e >= a.bx0) && (wx + e < a.bx1);
                accumulate<NC>(acc, out, kc, trim_x);
            }
            const uint32_t of = zo(k) + yo(wy) + xo(wx);
#pragma unroll
            for (int q = 0; q < NC; q++) st4_nt(a.mout[q], of, out[q]);
        }

        // 8. the next plane's faces; rotate
        load_faces(k + 1, cid1, f0);
        cid = cid1;
#pragma unroll
        for (int q = 0; q < NC; q++) {
            m1[q] = m2[q];
            h1[q] = h2[q];
            x1[q] = x2[q];
        }
        fix_v(m1);
        fix_v(h1);
        kx1 = kx2;
        k1w = fix_w(k2w);
        hk1 = fix_w(hk2);
        k2w = k3w;
        hk2 = hk3;
    }
}

template <int NC, bool VEL, bool EXACT>
__global__ void __launch_bounds__(256) csweep_kernel(SweepArgs a)
{
    __shared__ double sred[4][NC];
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc[q] = 0.0;
    csweep_tile<NC, VEL, EXACT>(a, xcd_tile(blockIdx.x, a.nblocks), acc);
    block_reduce<NC, 256>(acc, sred, a.partials + (size_t)blockIdx.x * NC);
}

// KEEP (one or two components, linear phase): capped at 128 VGPRs, so that
// 4 workgroups fit per CU.  With the state served by the memory-side cache
// the extra wavefront per SIMD pays (256^3: 58.7 against 61.0 us per step,
// the z-chunk filling the 1 024 slots; phi 0.35: 67.1 against 68.2 us);
// with the hint it did not.
template <int NC, bool VEL, bool EXACT, int LX, int RPT, bool BRK = false, bool KEEP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((KEEP && NC <= 2 && !EXACT) ? 4 : 1)))
sweep_kernel(SweepArgs a)
{
    __shared__ double sred[4][NC];
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc[q] = 0.0;
    sweep_tile<NC, VEL, EXACT, LX, RPT, BRK, KEEP>(a, xcd_tile(blockIdx.x, a.nblocks), acc);
    block_reduce<NC, 256>(acc, sred, a.partials + (size_t)blockIdx.x * NC);
}

// Batched sweep: one launch advances every instance of a batch by one step.
// The grid holds one workgroup per tile, or is sized to the resident
// workgroups (persistent); workgroup w takes tiles w, w + gridDim.x, ... of
// a list in descending cost (permuted in runs so that each XCD takes
// neighbouring tiles), each tile naming its instance.  Instance descriptors and the step counter live in device memory
// so that a chunk of launches can be captured in a graph and replayed.
template <int NC, bool VEL, bool EXACT, int LX, int RPT, bool BRK = false>
__global__ void __launch_bounds__(256) batch_kernel(BatchArgs b)
{
    __shared__ double sred[4][NC];
    const int step = *b.step_base + b.local_step;
    for (int t = blockIdx.x; t < b.ntiles; t += gridDim.x) {
        const int2 tl = b.tiles[t];
        const BatchInst* I = b.inst + tl.x;
        SweepArgs a;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            a.min[q] = I->m[b.parity][q];
            a.mout[q] = I->m[b.parity ^ 1][q];
        }
        a.mask = I->mask;
#pragma unroll
        for (int q = 0; q < 3; q++) a.adv[q] = I->adv[q];
        a.partials = nullptr;
        a.nx = I->nx;
        a.ny = I->ny;
        a.nz = I->nz;
        a.pitch = I->pitch;
        a.pitch1 = I->pitch1;
        a.ntx = I->ntx;
        a.nty = I->nty;
        a.ntz = I->ntz;
        a.lz = I->lz;
        a.nblocks = 0;
        a.bx0 = I->bx0; a.bx1 = I->bx1; a.by0 = I->by0; a.by1 = I->by1; a.bz0 = I->bz0; a.bz1 = I->bz1;
        a.inert = I->inert;
        a.nby = I->nby;
        a.c = I->c;
        // pulsed gradient along z: +G on [0, nd), -G on [nD, nD + nd) (the
        // second lobe wins where they overlap)
        a.c.bgz = (step >= I->nD && step < I->nD + I->nd) ? -I->bgz : (step < I->nd ? I->bgz : 0.f);
        double acc[NC];
#pragma unroll
        for (int q = 0; q < NC; q++) acc[q] = 0.0;
        sweep_tile<NC, VEL, EXACT, LX, RPT, BRK>(a, tl.y, acc);
        block_reduce<NC, 256>(acc, sred, b.partials + (size_t)t * NC);
        __syncthreads();
    }
}

// Per-instance signal rows of a chunk of batched steps: block (s, i) sums the
// partials of instance i's tiles (fixed order) into signal[i][row][q],
// row = *step_base - row0 + s + 1 (row0: the step counter at the run's start).
__global__ void __launch_bounds__(64) batch_reduce_kernel(const double* __restrict__ part, int ntiles, int nc,
                                                          const int* __restrict__ off, const int* __restrict__ ids,
                                                          const int* __restrict__ step_base, int row0, int64_t rows,
                                                          double* __restrict__ signal)
{
    __shared__ double sh[64];
    const int s = blockIdx.x, i = blockIdx.y;
    const double* p = part + (size_t)s * ntiles * nc;
    const int64_t row = (int64_t)(*step_base - row0) + s + 1;
    for (int q = 0; q < nc; q++) {
        double v = 0.0;
        for (int k = off[i] + (int)threadIdx.x; k < off[i + 1]; k += 64) v += p[(size_t)ids[k] * nc + q];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int h = 32; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) signal[((size_t)i * rows + row) * nc + q] = sh[0];
        __syncthreads();
    }
}

// Advances the device step counter after a chunk (one lane, vector store).
__global__ void step_advance_kernel(int* step_base, int n)
{
    if (threadIdx.x == 0) step_base[0] += n;
}

// Two time steps per pass (temporal blocking).  The workgroup owns TY full-
// width rows (nx <= 4*LX, one x tile) and a z-chunk [k0, k1).  Each thread
// owns one "level-1" row y = y0-1+ry of the halo-extended tile and one
// 4-voxel x-chunk.  Per plane p the pass
//   1. stages M^n(p) of rows y0-2 .. y0+TY+1 (with ghosts) in LDS A,
//   2. computes M^{n+1}(p) on rows y0-1 .. y0+TY (level 1, kept in registers
//      for the z-neighbours and published in LDS B after the level-2 step),
//   3. computes M^{n+2}(p-1) on rows y0 .. y0+TY-1 (level 2) from LDS B and
//      the level-1 registers of planes p-2, p-1, p, and writes it to HBM.
// M^{n+1} never leaves the chip, so one pass moves the 29 B/voxel of a single
// step for two steps.  Both levels call update_chunk, so the result is bit-
// identical to two one-step sweeps.  Signals of both steps are reduced.
template <int NC, bool VEL, bool EXACT, int LX>
__global__ void __launch_bounds__(512) sweep2_kernel(Sweep2Args a)
{
    constexpr int NT = 512;
    constexpr int RY = NT / LX;  // level-1 rows
    constexpr int TY = RY - 2;   // output rows
    constexpr int LW = 4 * LX + 8;
    constexpr int HA = RY + 2;   // LDS A rows: y0-2 .. y0+TY+1 (index y - y0 + 2)
    constexpr int HB = RY;       // LDS B rows: y0-1 .. y0+TY   (index y - y0 + 1)

    __shared__ __attribute__((aligned(16))) float sA[NC][HA][LW];
    __shared__ __attribute__((aligned(16))) float sB[NC][HB][LW];
    __shared__ __attribute__((aligned(16))) uint8_t kA[HA][LW];
    __shared__ __attribute__((aligned(16))) uint8_t kB[HB][LW];
    __shared__ double sred[NT / 64][NC];

    const int tid = threadIdx.x;
    const int lx = tid % LX;
    const int ry = tid / LX;
    const int t = xcd_tile(blockIdx.x, a.nblocks);
    const int ty = t % a.nty;
    const int tz = t / a.nty;
    const int y0 = ty * TY;
    const int k0 = tz * a.lz;
    const int k1 = min(k0 + a.lz, a.nz);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int ix = 4 * lx;
    const bool xin = ix < nx;
    const int xlast = nx - 1 - ix;
    const bool xend = xin && xlast <= 3;
    const bool xbeg = (ix == 0);
    const int c = 4 + 4 * lx;

    const int y = y0 - 1 + ry;                     // level-1 row
    const bool act = xin && y >= 0 && y < ny;      // computes level 1
    const bool own = act && ry >= 1 && ry <= RY - 2; // level-2 output row
    const int aR = ry + 1, bR = ry;
    // extra LDS A rows y0-2 (thread row 0) and y0+TY+1 (thread row RY-1)
    const int ye = (ry == 0) ? y0 - 2 : y0 + TY + 1;
    const bool eact = (ry == 0 || ry == RY - 1) && xin && ye >= 0 && ye < ny;
    const int eR = (ry == 0) ? 0 : HA - 1;

    const uint32_t Pb = (uint32_t)a.pitch * 4u;
    const uint32_t plane_b = (uint32_t)ny * Pb;
    const uint32_t plane_m = (uint32_t)ny * (uint32_t)a.pitch;
    const int yc = act ? y : 0, yec = eact ? ye : 0;
    const uint32_t rowf = (uint32_t)yc * Pb + (uint32_t)ix * 4u;
    const uint32_t rowm = (uint32_t)yc * (uint32_t)a.pitch + (uint32_t)ix;
    const uint32_t erowf = (uint32_t)yec * Pb + (uint32_t)ix * 4u;
    const uint32_t erowm = (uint32_t)yec * (uint32_t)a.pitch + (uint32_t)ix;

    auto clampz = [&](int p) { return p < 0 ? 0 : (p >= nz ? nz - 1 : p); };
    auto load_row = [&](bool on, int p, uint32_t rf, uint32_t rm, f4 (&m)[NC], uint32_t& kw) {
        if (on) {
            const uint32_t pf = (uint32_t)clampz(p) * plane_b, pm = (uint32_t)clampz(p) * plane_m;
#pragma unroll
            for (int q = 0; q < NC; q++) m[q] = ld4_nt(a.min[q], pf + rf);
            kw = ldw_nt(a.mask, pm + rm);
            if (xend) {
#pragma unroll
                for (int q = 0; q < NC; q++) m[q] = clamp_tail(m[q], xlast);
                kw = clamp_tail_w(kw, xlast);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) m[q] = f4{0.f, 0.f, 0.f, 0.f};
            kw = 0;
        }
    };
    // write a chunk (and its x ghosts, and the y ghost row when on the
    // domain edge) into an LDS plane
    auto stage = [&](float (*S)[LW], uint8_t (*K)[LW], int hrows, int row, const f4* v, uint32_t w, int yy,
                     int sq) {
        (void)sq;
#pragma unroll
        for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&S[q * hrows + row][c]) = v[q];
        *reinterpret_cast<uint32_t*>(&K[row][c]) = w;
        if (xend) {
            const int xl = xlast < 3 ? xlast : 3;
#pragma unroll
            for (int q = 0; q < NC; q++) S[q * hrows + row][c + 4] = v[q][xl];
            K[row][c + 4] = (uint8_t)((w >> (8 * xl)) & 0xffu);
        }
        if (xbeg) {
#pragma unroll
            for (int q = 0; q < NC; q++) S[q * hrows + row][c - 1] = v[q].x;
            K[row][c - 1] = (uint8_t)(w & 0xffu);
        }
        if (yy == 0 && row - 1 >= 0) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&S[q * hrows + row - 1][c]) = v[q];
            *reinterpret_cast<uint32_t*>(&K[row - 1][c]) = w;
        }
        if (yy == ny - 1 && row + 1 < hrows) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&S[q * hrows + row + 1][c]) = v[q];
            *reinterpret_cast<uint32_t*>(&K[row + 1][c]) = w;
        }
    };
    // neighbours of this thread's chunk in an LDS plane
    auto gather = [&](float (*S)[LW], uint8_t (*K)[LW], int hrows, int row, uint32_t kc, f4 (&ym)[NC],
                      f4 (&yp)[NC], float (&xl)[NC], float (&xr)[NC], uint32_t& kym, uint32_t& kyp, uint32_t& kxm,
                      uint32_t& kxp) {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            ym[q] = *reinterpret_cast<const f4*>(&S[q * hrows + row - 1][c]);
            yp[q] = *reinterpret_cast<const f4*>(&S[q * hrows + row + 1][c]);
            xl[q] = S[q * hrows + row][c - 1];
            xr[q] = S[q * hrows + row][c + 4];
        }
        kym = *reinterpret_cast<const uint32_t*>(&K[row - 1][c]);
        kyp = *reinterpret_cast<const uint32_t*>(&K[row + 1][c]);
        kxm = (kc << 8) | (uint32_t)K[row][c - 1];
        kxp = (kc >> 8) | ((uint32_t)K[row][c + 4] << 24);
    };

    float (*SA)[LW] = &sA[0][0];
    float (*SB)[LW] = &sB[0][0];

    // register pipeline: M^n planes p-1, p, p+1 (n0, n1, n2) and prefetch p+2
    f4 n0[NC], n1[NC], n2[NC], n3[NC], e1[NC], e2[NC];
    uint32_t kn0, kn1, kn2, kn3, ke1, ke2, kq = 0;
    // level-1 results of planes p-2, p-1 and p (r0, r1, rn)
    f4 r0[NC], r1[NC], rn[NC];
    // upwind velocity sums of the current level-1 plane and of the previous one
    f4 ax4 = f4{0.f, 0.f, 0.f, 0.f}, ay4 = ax4, az4 = ax4;
    f4 pax4 = ax4, pay4 = ax4, paz4 = ax4;
#pragma unroll
    for (int q = 0; q < NC; q++) r0[q] = r1[q] = rn[q] = f4{0.f, 0.f, 0.f, 0.f};

    const int p0 = k0 - 1;
    load_row(act, p0 - 1, rowf, rowm, n0, kn0);
    load_row(act, p0, rowf, rowm, n1, kn1);
    load_row(act, p0 + 1, rowf, rowm, n2, kn2);
    load_row(eact, p0, erowf, erowm, e1, ke1);

    double acc1[NC], acc2[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc1[q] = acc2[q] = 0.0;
    bool trim_x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) trim_x[e] = (ix + e >= a.bx0) && (ix + e < a.bx1);
    const bool trim_y = y >= a.by0 && y < a.by1;
    const Coef c1 = a.c1, c2 = a.c2;
    const float thy1 = c1.bgy * ((float)y + 0.5f), thy2 = c2.bgy * ((float)y + 0.5f);

    for (int p = p0; p <= k1; p++) {
        const bool l1 = p >= 0 && p < nz;
        const bool l2 = p - 1 >= k0;
        // prefetch M^n plane p+2 (clamped) and the extra row of plane p+1
        load_row(act, p + 2, rowf, rowm, n3, kn3);
        load_row(eact, p + 1, erowf, erowm, e2, ke2);
        if constexpr (VEL) {
            if (act && l1) {
                const uint32_t of = (uint32_t)p * plane_b + rowf;
                ax4 = ld4(a.adv[0], of);
                ay4 = ld4(a.adv[1], of);
                az4 = ld4(a.adv[2], of);
            }
        }
        // 1. stage M^n(p)
        if (l1) {
            if (act) stage(SA, kA, HA, aR, n1, kn1, y, 0);
            if (eact) stage(SA, kA, HA, eR, e1, ke1, ye, 0);
        }
        __syncthreads();
        // 2. level 1 at plane p (step n -> n+1)
        if (l1 && act) {
            f4 ym[NC], yp[NC];
            float xl[NC], xr[NC];
            uint32_t kym, kyp, kxm, kxp;
            gather(SA, kA, HA, aR, kn1, ym, yp, xl, xr, kym, kyp, kxm, kxp);
            const float thz = c1.bgz * ((float)p + 0.5f);
            float sn, cm1;
            plane_rotation<EXACT>(thz, sn, cm1);
            update_chunk<NC, VEL, EXACT>(c1, n1, n0, n2, ym, yp, xl, xr, kn1, kxm, kxp, kym, kyp, kn0, kn2, ax4, ay4,
                                         az4, ix, thy1, thz, sn, cm1, rn);
            if (xend) {
#pragma unroll
                for (int q = 0; q < NC; q++) rn[q] = clamp_tail(rn[q], xlast);
            }
            if (own && trim_y && p >= k0 && p < k1 && p >= a.bz0 && p < a.bz1) accumulate<NC>(acc1, rn, kn1, trim_x);
        }
        // 3. level 2 at plane p-1 (step n+1 -> n+2)
        if (l2 && own) {
            const int pm = p - 1;
            f4 ym[NC], yp[NC], zm[NC], zp[NC];
            float xl[NC], xr[NC];
            uint32_t kym, kyp, kxm, kxp;
            gather(SB, kB, HB, bR, kn0, ym, yp, xl, xr, kym, kyp, kxm, kxp);
#pragma unroll
            for (int q = 0; q < NC; q++) {
                zm[q] = (pm == 0) ? r1[q] : r0[q];
                zp[q] = (p >= nz) ? r1[q] : rn[q];
            }
            const float thz = c2.bgz * ((float)pm + 0.5f);
            float sn, cm1;
            plane_rotation<EXACT>(thz, sn, cm1);
            f4 out[NC];
            update_chunk<NC, VEL, EXACT>(c2, r1, zm, zp, ym, yp, xl, xr, kn0, kxm, kxp, kym, kyp, kq, kn1, pax4,
                                         pay4, paz4, ix, thy2, thz, sn, cm1, out);
            const uint32_t of = (uint32_t)pm * plane_b + rowf;
#pragma unroll
            for (int q = 0; q < NC; q++) st4_nt(a.mout[q], of, out[q]);
            if (trim_y && pm >= a.bz0 && pm < a.bz1) accumulate<NC>(acc2, out, kn0, trim_x);
        }
        __syncthreads();
        // publish M^{n+1}(p) for the level-2 step of the next plane
        if (l1 && act) stage(SB, kB, HB, bR, rn, kn1, y, 1);
        // rotate
#pragma unroll
        for (int q = 0; q < NC; q++) {
            n0[q] = n1[q];
            n1[q] = n2[q];
            n2[q] = n3[q];
            e1[q] = e2[q];
            r0[q] = r1[q];
            r1[q] = rn[q];
        }
        kq = kn0;
        kn0 = kn1;
        kn1 = kn2;
        kn2 = kn3;
        ke1 = ke2;
        if constexpr (VEL) {
            pax4 = ax4;
            pay4 = ay4;
            paz4 = az4;
        }
    }
    block_reduce<NC, NT>(acc1, sred, a.partials1 + (size_t)blockIdx.x * NC);
    __syncthreads();
    block_reduce<NC, NT>(acc2, sred, a.partials2 + (size_t)blockIdx.x * NC);
}

// Sets *flag when some face between two in-domain cells carries velocity and
// touches a solid cell (then solid magnetization can reach a pore through the
// upwind terms and the sweeps must load it).  One thread per cell checks the
// faces on its minus side; racing writers all store 1.
__global__ void __launch_bounds__(256) inert_check_kernel(InertArgs a)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const size_t P = a.pitch, P1 = a.pitch1;
    const size_t c = ((size_t)z * a.ny + y) * P + x;
    const bool pc = a.mask[c] != 0;
    bool bad = false;
    if (x > 0) {
        const float v = a.vx[((size_t)z * a.ny + y) * P1 + x];
        bad |= v != 0.f && !(pc && a.mask[c - 1] != 0);
    }
    if (y > 0) {
        const float v = a.vy[((size_t)z * (a.ny + 1) + y) * P + x];
        bad |= v != 0.f && !(pc && a.mask[c - P] != 0);
    }
    if (z > 0) {
        const float v = a.vz[((size_t)z * a.ny + y) * P + x];
        bad |= v != 0.f && !(pc && a.mask[c - (size_t)a.ny * P] != 0);
    }
    if (bad) *a.flag = 1;
}

hipError_t launch_inert_check(const InertArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(inert_check_kernel, dim3((a.nx + 255) / 256, a.ny, a.nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Face-velocity sums of one voxel per thread.  Upwind (the plain sweep): per
// axis the sum of the voxel's two face velocities, kept where the face on
// the inflow side carries velocity (the minus face for a positive sum, the
// plus face for a negative one) and 0 elsewhere -- ExplicitSimple
// (cpp:3236-3247) advects across a face only when its velocity is nonzero.
// The sweep then adds h*max(s, 0) to the minus face's weight and
// -h*min(s, 0) to the plus face's, the same values as the face-by-face
// test.  Centred (the MacCormack stages, whose advection uses the
// cell-centred velocity (um + up)/2): the plain sums.  Written in the
// state's row layout [nz][ny][pitch], 0 past nx.
__device__ __forceinline__ float upwind_sum(float m, float p)
{
    const float s = m + p;
    const bool keep = s > 0.f ? m != 0.f : (s < 0.f && p != 0.f);
    return keep ? s : 0.f;
}

__global__ void __launch_bounds__(256) adv_kernel(AdvArgs a)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.pitch) return;
    const size_t row = (size_t)z * a.ny + y;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    if (x < a.nx) {
        const size_t rx = row * a.pitch1 + x;
        const size_t ry = ((size_t)z * (a.ny + 1) + y) * a.pitch + x;
        const size_t rz = row * a.pitch + x;
        const float um = a.vx[rx], up = a.vx[rx + 1];
        const float vm = a.vy[ry], vp = a.vy[ry + a.pitch];
        const float wm = a.vz[rz], wp = a.vz[rz + (size_t)a.ny * a.pitch];
        if (a.centred) {
            sx = um + up;
            sy = vm + vp;
            sz = wm + wp;
        } else {
            sx = upwind_sum(um, up);
            sy = upwind_sum(vm, vp);
            sz = upwind_sum(wm, wp);
        }
    }
    const size_t o = row * a.pitch + x;
    a.out[0][o] = sx;
    a.out[1][o] = sy;
    a.out[2][o] = sz;
}

hipError_t launch_adv(const AdvArgs& a, hipStream_t s)
{
    if (a.nx <= 0 || a.ny <= 0 || a.nz <= 0 || a.pitch < a.nx || a.pitch1 <= a.nx) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.pitch + 255) / 256), (unsigned)a.ny, (unsigned)a.nz);
    hipLaunchKernelGGL(adv_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// Instantaneous RF / gradient pulse: every pore voxel's M is rotated by an
// angle about an axis, as AddImpulse + OrtTransform do it
// (src/NmrRelaxationSolve.cpp:2536-2608).  Angles and the rotation are
// evaluated in FP64 in the reference's order of operations; the state stays
// FP32.  Reference index i maps to x + 1 here (no ghost layer: ghost cells
// are refilled from the interior after every pulse in the reference anyway).
__global__ void __launch_bounds__(256) pulse_kernel(PulseArgs a)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= a.nx) return;
    const size_t o = ((size_t)z * a.ny + y) * a.pitch + x;
    if (a.mask[o] == 0) return;
    const size_t fo = a.nb8 > 0 ? brick_off(x, y, z, a.nby, a.nb8) : o; // state offset
    double len = sqrt(a.ax * a.ax + a.ay * a.ay + a.az * a.az);
    if (len == 0.0) len = 1.0;
    const double n0x = a.ax / len, n0y = a.ay / len, n0z = a.az / len;
    const double i = (double)(x + 1), j = (double)(y + 1), k = (double)(z + 1);
    double angle, rx, ry, rz;
    if (a.flag == 1) { // narrow gradient pulse, axis = unit gradient direction
        angle = a.param * a.gamma * ((i - 1) * a.dx * a.ax + (j - 1) * a.dy * a.ay + (k - 1) * a.dz * a.az);
        rx = n0x; ry = n0y; rz = n0z;
    } else if (a.flag == 2) { // RF pulse of angle param about the (unnormalised) axis
        angle = a.param;
        rx = a.ax; ry = a.ay; rz = a.az;
    } else if (a.flag == 3) { // multi-gradient pulse about a far-away origin
        const double distance = 10e+12;
        const double eff = sqrt(a.ax * a.ax + a.ay * a.ay + a.az * a.az);
        const double t = (i - 1.0 - distance) * a.dx * n0x + (j - 1.0 - distance) * a.dy * n0y +
                         (k - 1.0 - distance) * a.dz * n0z;
        angle = a.gamma * eff * sqrt(t * t);
        rx = n0x; ry = n0y; rz = n0z;
    } else { // flag 4: radial phase about the domain centre
        const double tx = (i - 0.5) * a.dx * a.ax - 0.5 * a.nx * a.dx * a.ax;
        const double ty = (j - 0.5) * a.dy * a.ay - 0.5 * a.ny * a.dy * a.ay;
        const double tz = (k - 0.5) * a.dz * a.az - 0.5 * a.nz * a.dz * a.az;
        angle = 0.0 - a.gamma * sqrt(tx * tx + ty * ty + tz * tz);
        rx = a.ax; ry = a.ay; rz = a.az;
    }
    const double c = cos(angle), s = sin(angle);
    const double vx = a.m[0][fo], vy = a.m[1][fo], vz = a.nc == 3 ? (double)a.m[2][fo] : 0.0;
    const double scal = rx * vx + ry * vy + rz * vz;
    a.m[0][fo] = (float)(vx * c + rx * (1.0 - c) * scal + s * (ry * vz - rz * vy));
    a.m[1][fo] = (float)(vy * c + ry * (1.0 - c) * scal + s * (rz * vx - rx * vz));
    if (a.nc == 3) a.m[2][fo] = (float)(vz * c + rz * (1.0 - c) * scal + s * (rx * vy - ry * vx));
}

hipError_t launch_pulse(const PulseArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(pulse_kernel, dim3((a.nx + 255) / 256, a.ny, a.nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Sum of the per-workgroup partials of nsteps steps: out[s*NC + q], fixed order.
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double* __restrict__ part,
                                                              int nwg, int nc, double* __restrict__ out)
{
    __shared__ double sh[256];
    const int s = blockIdx.x;
    const double* p = part + (size_t)s * (size_t)nwg * (size_t)nc;
    for (int q = 0; q < nc; q++) {
        double v = 0.0;
        for (int w = threadIdx.x; w < nwg; w += 256) v += p[(size_t)w * nc + q];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)s * nc + q] = sh[0];
        __syncthreads();
    }
}

// Pore sum of a field over the signal window: each workgroup takes 16 rows
// (y, z) of the window, a wavefront one row at a time with 64 consecutive
// x per load, and writes part[blockIdx.x*NS + q]; a second pass
// (reduce_partials_kernel) sums the workgroups in a fixed order.
constexpr int kFieldSumRows = 16;
template <int NC, bool W>
__global__ void __launch_bounds__(256) field_sum_kernel(FieldSumArgs a)
{
    constexpr int NS = W ? 2 * NC : NC;
    __shared__ double sred[4][NS];
    const int wy = a.by1 - a.by0, wz = a.bz1 - a.bz0;
    const long nrows = (wy > 0 && wz > 0) ? (long)wy * wz : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) acc[q] = 0.0;
    const long r1 = min(nrows, (long)(blockIdx.x + 1) * kFieldSumRows);
    for (long r = (long)blockIdx.x * kFieldSumRows + wave; r < r1; r += 4) {
        const int y = a.by0 + (int)(r % wy), z = a.bz0 + (int)(r / wy);
        const size_t row = ((size_t)z * (size_t)a.ny + (size_t)y) * (size_t)a.pitch;
        for (int x = a.bx0 + lane; x < a.bx1; x += 64) {
            if (a.mask[row + x] == 0) continue;
            const size_t fo = a.nb8 > 0 ? brick_off(x, y, z, a.nby, a.nb8) : row + x;
            if constexpr (W) {
                const double w1 = a.w[0][row + x], w2 = a.w[1][row + x];
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const double v = a.f[q][fo];
                    acc[q] += v * w1;
                    acc[NC + q] += v * w2;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) acc[q] += (double)a.f[q][fo];
            }
        }
    }
    block_reduce<NS, 256>(acc, sred, a.part + (size_t)blockIdx.x * NS);
}

// ---------------------------------------------------------------------------
// General z-sweep (MacCormack stages, any ghost rule, per-voxel coefficients).

// Update of one 4-voxel chunk with the per-voxel kernel's arithmetic
// (general_kernel): stage ST of a step, face weights from the mask and, with
// COEF, harmonic means of per-voxel diffusivities; ghost neighbours already
// hold their rule's value and the pore mask.  dxl / dxr: diffusivity of the x
// neighbours of elements 0 and 3; mav: the state at the start of the step
// (corrector stage).  vx4, vy4, vz0: the voxels' face-velocity sums per
// axis (adv_kernel): for the MacCormack stages the plain sums um + up (the
// centred velocity, doubled), for the upwind stage the upwind sums (um + up
// when the inflow face carries velocity, else 0), which give the face
// tests' bits.
template <int NC, bool VEL, int ST, bool COEF>
__device__ __forceinline__ void gupdate_chunk(const Coef& cf, const f4 (&mc)[NC], const f4 (&fzm)[NC],
                                              const f4 (&fzp)[NC], const f4 (&fym)[NC], const f4 (&fyp)[NC],
                                              const float (&xl)[NC], const float (&xr)[NC], uint32_t kc,
                                              uint32_t kxm, uint32_t kxp, uint32_t kym, uint32_t kyp,
                                              uint32_t kzm, uint32_t kzp, const f4& vx4, const f4& vy4,
                                              const f4& vz0, const f4& dc,
                                              float dxl, float dxr, const f4& dym, const f4& dyp, const f4& dzm,
                                              const f4& dzp, bool r1p, const f4& r1c, const f4& r2c,
                                              const f4 (&mav)[NC], bool amask, int ix, float thy, float thz,
                                              const f4& bnc, int bnax, float dtg, f4 (&out)[NC])
{
    const float hxh = 0.5f * cf.hx, hyh = 0.5f * cf.hy, hzh = 0.5f * cf.hz;
    const float cd[3] = {cf.cdx, cf.cdy, cf.cdz};
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const float kn[6] = {ub(kxm, e), ub(kxp, e), ub(kym, e), ub(kyp, e), ub(kzm, e), ub(kzp, e)};
        float w[6];
        if constexpr (COEF) {
            const float d0 = dc[e];
            const float dnb[6] = {e > 0 ? dc[e - 1] : dxl, e < 3 ? dc[e + 1] : dxr, dym[e], dyp[e], dzm[e], dzp[e]};
#pragma unroll
            for (int s = 0; s < 6; s++) {
                const float hm = hmean(d0, dnb[s]);
                w[s] = cd[s / 2] * hm * kn[s];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 6; s++) w[s] = cd[s / 2] * kn[s];
        }
        if constexpr (VEL) {
            if constexpr (ST == GEN_UPWIND) {
                // with NMR_F_ADVECT_MASK only across faces to a pore
                bool open[6];
#pragma unroll
                for (int q = 0; q < 6; q++) open[q] = !amask || kn[q] != 0.f;
                w[0] = open[0] ? fmaf(hxh, fmaxf(vx4[e], 0.f), w[0]) : w[0];
                w[1] = open[1] ? fmaf(-hxh, fminf(vx4[e], 0.f), w[1]) : w[1];
                w[2] = open[2] ? fmaf(hyh, fmaxf(vy4[e], 0.f), w[2]) : w[2];
                w[3] = open[3] ? fmaf(-hyh, fminf(vy4[e], 0.f), w[3]) : w[3];
                w[4] = open[4] ? fmaf(hzh, fmaxf(vz0[e], 0.f), w[4]) : w[4];
                w[5] = open[5] ? fmaf(-hzh, fminf(vz0[e], 0.f), w[5]) : w[5];
            } else if constexpr (ST == GEN_PREDICT) {
                w[0] = fmaf(kn[0] * hxh, vx4[e], w[0]);
                w[2] = fmaf(kn[2] * hyh, vy4[e], w[2]);
                w[4] = fmaf(kn[4] * hzh, vz0[e], w[4]);
            } else {
                w[1] = fmaf(-kn[1] * hxh, vx4[e], w[1]);
                w[3] = fmaf(-kn[3] * hyh, vy4[e], w[3]);
                w[5] = fmaf(-kn[5] * hzh, vz0[e], w[5]);
            }
        }
        const float nsx = 2.f - kn[0] - kn[1], nsy = 2.f - kn[2] - kn[3], nsz = 2.f - kn[4] - kn[5];
        float rate2;
        if constexpr (COEF) rate2 = r2c[e];
        else rate2 = fmaf(cf.s2z, nsz, fmaf(cf.s2y, nsy, fmaf(cf.s2x, nsx, cf.r2)));
        // precession angles: the gradient field's, or along axis bnax the
        // per-voxel field's (value selects, as in general_kernel)
        const float tb = bnax >= 0 ? dtg * bnc[e] : 0.f;
        const float ty = bnax == 1 ? tb : thy, tz = bnax == 2 ? tb : thz;
        float m0[NC], d[NC];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const float m = mc[q][e];
            m0[q] = m;
            const float f[6] = {(e > 0) ? mc[q][e - 1] : xl[q], (e < 3) ? mc[q][e + 1] : xr[q], fym[q][e], fyp[q][e],
                                fzm[q][e], fzp[q][e]};
            float t = w[0] * (f[0] - m);
            t = fmaf(w[1], f[1] - m, t);
            t = fmaf(w[2], f[2] - m, t);
            t = fmaf(w[3], f[3] - m, t);
            t = fmaf(w[4], f[4] - m, t);
            d[q] = fmaf(w[5], f[5] - m, t);
        }
        if constexpr (NC == 3) {
            float rate1;
            if (COEF && r1p) rate1 = r1c[e];
            else rate1 = fmaf(cf.s1z, nsz, fmaf(cf.s1y, nsy, fmaf(cf.s1x, nsx, cf.r1)));
            const float par = fmaf(cf.n2, m0[2], fmaf(cf.n1, m0[1], cf.n0 * m0[0]));
            const float nn[3] = {cf.n0, cf.n1, cf.n2};
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const float mpar = nn[q] * par;
                d[q] = fmaf(rate1, fmaf(cf.peq, nn[q], -mpar), d[q]);
                d[q] = fmaf(rate2, mpar - m0[q], d[q]);
            }
            const float tx = bnax == 0 ? tb : cf.bgx * ((float)(ix + e) + 0.5f);
            d[0] = fmaf(tz, m0[1], fmaf(-ty, m0[2], d[0]));
            d[1] = fmaf(-tz, m0[0], fmaf(tx, m0[2], d[1]));
            d[2] = fmaf(ty, m0[0], fmaf(-tx, m0[1], d[2]));
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) d[q] = fmaf(-rate2, m0[q], d[q]);
            if constexpr (NC == 2) {
                d[0] = fmaf(tz, m0[1], d[0]);
                d[1] = fmaf(-tz, m0[0], d[1]);
            }
        }
        const bool pore = ((kc >> (8 * e)) & 0xffu) != 0;
#pragma unroll
        for (int q = 0; q < NC; q++) {
            if constexpr (ST == GEN_CORRECT) out[q][e] = pore ? 0.5f * (mav[q][e] + (m0[q] + d[q])) : 0.f;
            else if constexpr (ST == GEN_PREDICT) out[q][e] = pore ? m0[q] + d[q] : 0.f; // the predictor holds 0 at solids
            else out[q][e] = pore ? m0[q] + d[q] : m0[q];
        }
    }
}

// Elements past nx (index > last) take the ghost value g.
__device__ __forceinline__ f4 fill_tail(f4 v, int last, float g)
{
    if (last < 3) v.w = g;
    if (last < 2) v.z = g;
    if (last < 1) v.y = g;
    return v;
}

// One tile of the general sweep: 256 threads, LX lanes x RY rows, one row
// per thread, a z-chunk of lz planes.  Ghost rules per axis: zero
// derivative (the edge value), periodic (the opposite edge) or zero value;
// every ghost carries the mask and diffusivity of the voxel it borders (the
// reference's ghost copies of actCell and diffusion), i.e. it is a pore for
// every voxel that is updated.  Ghost rows and columns are staged in LDS by
// the threads that own the tile's halo; ghost planes are loaded in place of
// the out-of-range planes.
template <int NC, bool VEL, int ST, bool COEF, int LX, bool BRK = false, bool W = false, bool KEEP = false>
__device__ __forceinline__ void gsweep_tile(const GSweepArgs& a, const int t, double (&acc)[W ? 2 * NC : NC])
{
    constexpr int NT = 256;
    constexpr int RY = NT / LX;
    constexpr int TX = 4 * LX;
    constexpr int TY = RY;
    constexpr int LW = TX + 8;
    constexpr int LH = TY + 2;
    constexpr int DH = COEF ? LH : 1, DW = COEF ? LW : 4;

    __shared__ __attribute__((aligned(16))) float sm[2][NC][LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t sk[2][LH][LW];
    __shared__ __attribute__((aligned(16))) float sd[2][DH][DW];

    const int tid = threadIdx.x;
    const int lx = tid % LX;
    const int ry = tid / LX;
    const int tx = t % a.ntx;
    const int tyz = t / a.ntx;
    const int ty = tyz % a.nty;
    const int tz = tyz / a.nty;
    const int x0 = tx * TX, y0 = ty * TY;
    const int k0 = tz * a.lz;
    const int k1 = min(k0 + a.lz, a.nz);
    const int ix = x0 + 4 * lx;
    const int iy = y0 + ry;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int bcx = a.bc[0], bcy = a.bc[1], bcz = a.bc[2];
    const bool xin = ix < nx;
    const int xlast = nx - 1 - ix;
    const bool xend = xin && xlast <= 3;
    const bool xpart = xin && xlast < 3;
    const bool rin = xin && iy < ny;
    const int c = 4 + 4 * lx;
    const int lr = 1 + ry;

    const uint32_t Pb = (uint32_t)a.pitch * 4u;
    const uint32_t plane_b = (uint32_t)ny * Pb;
    const uint32_t plane_m = (uint32_t)ny * (uint32_t)a.pitch;
    // byte offsets along x, y, z of the float arrays (row layout, or with
    // BRK the 128-byte bricks of 8 x-voxels by 4 rows of the plain sweep);
    // the per-voxel arrays (coefficients, field, phase weights) are laid out
    // as the state
    const uint32_t nbx = ((uint32_t)a.pitch + 7u) >> 3;
    auto xo = [&](int x) -> uint32_t {
        return BRK ? ((uint32_t)(x >> 3) << 7) + ((uint32_t)(x & 7) << 2) : (uint32_t)x * 4u;
    };
    auto yo = [&](int y) -> uint32_t {
        return BRK ? (uint32_t)(y >> 2) * nbx * 128u + ((uint32_t)(y & 3) << 5) : (uint32_t)y * Pb;
    };
    auto zo = [&](int z) -> uint32_t {
        return BRK ? (uint32_t)z * (uint32_t)a.nby * nbx * 128u : (uint32_t)z * plane_b;
    };
    const int yc = iy < ny ? iy : ny - 1; // rows past ny never load
    const uint32_t rowf = yo(yc) + xo(xin ? ix : 0);
    const uint32_t rowm = (uint32_t)yc * (uint32_t)a.pitch + (uint32_t)(xin ? ix : 0);

    // halo / ghost row: the top thread stages row y0-1, the bottom thread
    // row y0+TY, or the ghost row ny at LDS row 1 + (ny - y0) when the tile
    // reaches the domain end.  Values come from row hsrc (-1: zeros), mask
    // and diffusivity from row hmrow.
    const bool top = ry == 0, bot = ry == RY - 1;
    int hsrc = -1, hmrow = 0, hlr = 0;
    if (top) {
        hlr = 0;
        if (y0 > 0) hsrc = hmrow = y0 - 1;
        else {
            hmrow = 0;
            hsrc = bcy == BC_NEUMANN ? 0 : (bcy == BC_PERIODIC ? ny - 1 : -1);
        }
    } else if (bot) {
        const int yb = y0 + TY;
        if (yb < ny) {
            hsrc = hmrow = yb;
            hlr = TY + 1;
        } else {
            hmrow = ny - 1;
            hsrc = bcy == BC_NEUMANN ? ny - 1 : (bcy == BC_PERIODIC ? 0 : -1);
            hlr = 1 + (ny - y0);
        }
    }
    const bool hrow = (top || bot) && xin;
    const uint32_t hrowf = yo(hsrc >= 0 ? hsrc : 0) + xo(xin ? ix : 0);
    const uint32_t hrowm = (uint32_t)hmrow * (uint32_t)a.pitch + (uint32_t)(xin ? ix : 0);

    // x halo column: lane 0 loads column x0-1 (or, for periodic x in the
    // first tile, the ghost value at nx-1), lane LX-1 column x0+TX when it
    // is inside the domain.  The lane holding nx-1 loads the periodic ghost
    // value of column 0.
    const bool xgl = lx == 0 && x0 == 0;                 // left column is the domain ghost
    const bool hcl = lx == 0 && (x0 > 0 || bcx == BC_PERIODIC);
    const bool hcr = lx == LX - 1 && x0 + TX < nx;
    const int hcx = lx == 0 ? (x0 > 0 ? x0 - 1 : nx - 1) : x0 + TX;
    const bool gxr = xend && bcx == BC_PERIODIC;         // right ghost value from column 0

    const bool skip_m = (ST != GEN_UPWIND) || a.skip_m != 0 || a.amask != 0;
    auto clampz = [&](int kk) { return kk < 0 ? 0 : (kk >= nz ? nz - 1 : kk); };
    auto zsrc = [&](int kk) -> int {
        if (kk >= 0 && kk < nz) return kk;
        if (bcz == BC_NEUMANN) return kk < 0 ? 0 : nz - 1;
        if (bcz == BC_PERIODIC) return kk < 0 ? nz - 1 : 0;
        return -1;
    };

    f4 mp[NC], mc[NC], mn[NC], hc[NC], hn[NC];
    uint32_t kp = 0, kc = 0, kn = 0, kq = 0, hkc = 0, hkn = 0, hkq = 0;
    float xc[NC], xn[NC], gc[NC], gn[NC];
    uint32_t kxc = 0, kxn = 0;
    f4 dp = f4{0.f, 0.f, 0.f, 0.f}, dcv = dp, dnv = dp, hdc = dp, hdn = dp;
    float xdc = 0.f, xdn = 0.f;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        mp[q] = mc[q] = mn[q] = hc[q] = hn[q] = f4{0.f, 0.f, 0.f, 0.f};
        xc[q] = xn[q] = gc[q] = gn[q] = 0.f;
    }

    auto load_mask = [&](int kk) -> uint32_t {
        return xin ? ldw_nt(a.mask, (uint32_t)clampz(kk) * plane_m + rowm) : 0u;
    };
    auto load_m = [&](int kk, uint32_t kw, f4 (&m)[NC]) {
        const int zs = zsrc(kk);
        if (xin && zs >= 0 && (!skip_m || kw != 0)) {
            const uint32_t of = zo(zs) + rowf;
#pragma unroll
            for (int q = 0; q < NC; q++) m[q] = ld4_state<KEEP>(a.f[q], of);
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) m[q] = f4{0.f, 0.f, 0.f, 0.f};
        }
    };
    auto load_d = [&](int kk) -> f4 {
        if constexpr (COEF) {
            if (xin) return ld4_nt(a.dn, zo(clampz(kk)) + rowf);
        }
        return f4{0.f, 0.f, 0.f, 0.f};
    };
    auto load_hmask = [&](int kk) -> uint32_t {
        return hrow ? *reinterpret_cast<const uint32_t*>(a.mask + (uint32_t)clampz(kk) * plane_m + hrowm) : 0u;
    };
    auto load_hm = [&](int kk, uint32_t hk, f4 (&h)[NC]) {
        const int zs = zsrc(kk);
        if (hrow && hsrc >= 0 && zs >= 0 && (!skip_m || hk != 0)) {
            const uint32_t of = zo(zs) + hrowf;
#pragma unroll
            for (int q = 0; q < NC; q++) h[q] = ld4(a.f[q], of);
        } else {
#pragma unroll
            for (int q = 0; q < NC; q++) h[q] = f4{0.f, 0.f, 0.f, 0.f};
        }
    };
    auto load_hd = [&](int kk) -> f4 {
        if constexpr (COEF) {
            if (hrow) return ld4(a.dn, zo(clampz(kk)) + yo(hmrow) + xo(ix));
        }
        return f4{0.f, 0.f, 0.f, 0.f};
    };
    // x halo (value, mask, diffusivity) of plane kk, and the periodic right ghost
    auto load_xhalo = [&](int kk, float (&xv)[NC], uint32_t& kx, float& xd, float (&gv)[NC]) {
        const int zs = zsrc(kk);
        if ((hcl || hcr) && iy < ny) {
            if (zs >= 0) {
                const uint32_t of = zo(zs) + yo(iy) + xo(hcx);
#pragma unroll
                for (int q = 0; q < NC; q++) xv[q] = ld1(a.f[q], of);
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) xv[q] = 0.f;
            }
            if (!xgl) {
                const uint32_t cz = (uint32_t)clampz(kk);
                kx = a.mask[cz * plane_m + (uint32_t)iy * (uint32_t)a.pitch + (uint32_t)hcx];
                if constexpr (COEF) xd = ld1(a.dn, zo((int)cz) + yo(iy) + xo(hcx));
            }
        }
        if (gxr && iy < ny) {
            if (zs >= 0) {
                const uint32_t of = zo(zs) + yo(iy) + xo(0);
#pragma unroll
                for (int q = 0; q < NC; q++) gv[q] = ld1(a.f[q], of);
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) gv[q] = 0.f;
            }
        }
    };
    // ghost value past nx for the tail of a chunk and LDS column c+4: the
    // edge element xlast (zero derivative), the periodic value or 0.  Written
    // as value selects only: the same choice as nested conditionals with
    // branches, which hipcc once compiled into a selection that stored
    // element 1 in place of element 3 when nx is a multiple of 4.
    auto right_ghost = [&](const f4 (&m)[NC], const float (&gv)[NC], int q) -> float {
        const int xl = xlast < 3 ? xlast : 3;
        const float e01 = xl == 0 ? m[q].x : m[q].y;
        const float e23 = xl == 2 ? m[q].z : m[q].w;
        const float edge = xl < 2 ? e01 : e23;
        const float other = bcx == BC_PERIODIC ? gv[q] : 0.f;
        return bcx == BC_NEUMANN ? edge : other;
    };
    // Ghost fix of the centre plane's registers: the element past nx takes
    // the x ghost value, its mask and diffusivity those of element xlast.
    // Applied one iteration after the plane's loads were issued, so that no
    // wait for memory follows a load directly.  Only the centre plane's tail
    // is ever read by a voxel inside the domain.
    auto fix_centre = [&](f4 (&m)[NC], uint32_t& kw, f4& dv, const float (&gv)[NC]) {
        if (xpart) {
#pragma unroll
            for (int q = 0; q < NC; q++) m[q] = fill_tail(m[q], xlast, right_ghost(m, gv, q));
            kw = clamp_tail_w(kw, xlast);
            if constexpr (COEF) dv = clamp_tail(dv, xlast);
        }
    };

    f4 vx4 = f4{0.f, 0.f, 0.f, 0.f}, vy4 = vx4, wc = vx4, r1c = vx4, r2c = vx4, bnc = vx4;
    f4 pw1 = vx4, pw2 = vx4; // W: phase weights of the own chunk
    const int bnax = a.bn ? a.bnax : -1;
    f4 mav[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) mav[q] = f4{0.f, 0.f, 0.f, 0.f};
    // per-voxel inputs of plane kk (own chunk; only chunks holding a pore)
    auto load_own = [&](int kk, uint32_t kw) {
        if (rin && kw != 0) {
            const uint32_t pz = zo(kk);
            if constexpr (VEL) { // the voxels' face-velocity sums
                vx4 = ld4_nt(a.cs[0], pz + rowf);
                vy4 = ld4_nt(a.cs[1], pz + rowf);
                wc = ld4_nt(a.cs[2], pz + rowf);
            }
            if constexpr (COEF) {
                const uint32_t pc = zo(kk) + rowf;
                r2c = ld4_nt(a.r2v, pc);
                if constexpr (NC == 3) {
                    if (a.r1v) r1c = ld4_nt(a.r1v, pc);
                }
            }
            if (bnax >= 0) bnc = ld4_nt(a.bn, zo(kk) + rowf);
            if constexpr (W && ST != GEN_PREDICT) {
                pw1 = ld4_nt(a.w[0], zo(kk) + rowf);
                pw2 = ld4_nt(a.w[1], zo(kk) + rowf);
            }
            if constexpr (ST == GEN_CORRECT) {
#pragma unroll
                for (int q = 0; q < NC; q++) mav[q] = ld4_state<KEEP>(a.mavg[q], pz + rowf);
            }
        }
    };

    // prologue: masks of planes k0-1, k0, k0+1; values of k0-1, k0
    kp = load_mask(k0 - 1);
    kc = load_mask(k0);
    kn = load_mask(k0 + 1);
    hkc = load_hmask(k0);
    hkn = load_hmask(k0 + 1);
    load_m(k0 - 1, kp, mp);
    load_m(k0, kc, mc);
    load_hm(k0, hkc, hc);
    if constexpr (COEF) {
        dp = load_d(k0 - 1);
        dcv = load_d(k0);
        hdc = load_hd(k0);
    }
    load_xhalo(k0, xc, kxc, xdc, gc);
    fix_centre(mc, kc, dcv, gc);
    __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0): retire the prologue's loads before the loop

    const Coef cf = a.c;
    bool trim_x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) trim_x[e] = (ix + e >= a.bx0) && (ix + e < a.bx1);

    for (int k = k0; k < k1; k++) {
        const int buf = (k - k0) & 1;
        // prefetch: masks of plane k+2, values (and diffusivities) of plane k+1
        kq = load_mask(k + 2);
        hkq = load_hmask(k + 2);
        load_m(k + 1, kn, mn);
        load_hm(k + 1, hkn, hn);
        if constexpr (COEF) {
            dnv = load_d(k + 1);
            hdn = load_hd(k + 1);
        }
        load_xhalo(k + 1, xn, kxn, xdn, gn);
        load_own(k, kc);

        // stage plane k with its halo and ghosts
        if (rin) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][lr][c]) = mc[q];
            *reinterpret_cast<uint32_t*>(&sk[buf][lr][c]) = kc;
            if constexpr (COEF) *reinterpret_cast<f4*>(&sd[buf][lr][c]) = dcv;
            if (xend) { // ghost column nx (or padding, when nx falls inside the chunk)
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][c + 4] = right_ghost(mc, gc, q);
                const int xl = xlast < 3 ? xlast : 3;
                sk[buf][lr][c + 4] = (uint8_t)((kc >> (8 * xl)) & 0xffu);
                if constexpr (COEF) sd[buf][lr][c + 4] = dcv[xl];
            }
            if (xgl) { // ghost column -1
#pragma unroll
                for (int q = 0; q < NC; q++)
                    sm[buf][q][lr][c - 1] = bcx == BC_PERIODIC ? xc[q] : (bcx == BC_DIRICHLET ? 0.f : mc[q].x);
                sk[buf][lr][c - 1] = (uint8_t)(kc & 0xffu);
                if constexpr (COEF) sd[buf][lr][c - 1] = dcv.x;
            } else if (hcl || hcr) {
                const int col = hcl ? c - 1 : c + 4;
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][col] = xc[q];
                sk[buf][lr][col] = (uint8_t)kxc;
                if constexpr (COEF) sd[buf][lr][col] = xdc;
            }
        }
        if (hrow) {
#pragma unroll
            for (int q = 0; q < NC; q++) *reinterpret_cast<f4*>(&sm[buf][q][hlr][c]) = hc[q];
            *reinterpret_cast<uint32_t*>(&sk[buf][hlr][c]) = hkc;
            if constexpr (COEF) *reinterpret_cast<f4*>(&sd[buf][hlr][c]) = hdc;
        }
        __syncthreads();

        if (rin && kc != 0) {
            const bool trim = k >= a.bz0 && k < a.bz1 && iy >= a.by0 && iy < a.by1;
            const float thy = cf.bgy * ((float)iy + 0.5f), thz = cf.bgz * ((float)k + 0.5f);
            f4 ym[NC], yp[NC];
            float xl[NC], xr[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                ym[q] = *reinterpret_cast<const f4*>(&sm[buf][q][lr - 1][c]);
                yp[q] = *reinterpret_cast<const f4*>(&sm[buf][q][lr + 1][c]);
                xl[q] = sm[buf][q][lr][c - 1];
                xr[q] = sm[buf][q][lr][c + 4];
            }
            const uint32_t kym = *reinterpret_cast<const uint32_t*>(&sk[buf][lr - 1][c]);
            const uint32_t kyp = *reinterpret_cast<const uint32_t*>(&sk[buf][lr + 1][c]);
            const uint32_t kxm = (kc << 8) | (uint32_t)sk[buf][lr][c - 1];
            const uint32_t kxp = (kc >> 8) | ((uint32_t)sk[buf][lr][c + 4] << 24);
            f4 dym = f4{0.f, 0.f, 0.f, 0.f}, dyp = dym;
            float dxl = 0.f, dxr = 0.f;
            if constexpr (COEF) {
                dym = *reinterpret_cast<const f4*>(&sd[buf][lr - 1][c]);
                dyp = *reinterpret_cast<const f4*>(&sd[buf][lr + 1][c]);
                dxl = sd[buf][lr][c - 1];
                dxr = sd[buf][lr][c + 4];
            }
            f4 out[NC];
            gupdate_chunk<NC, VEL, ST, COEF>(cf, mc, mp, mn, ym, yp, xl, xr, kc, kxm, kxp, kym, kyp, kp, kn, vx4, vy4, wc,
                                             dcv, dxl, dxr, dym, dyp, dp, dnv, a.r1v != nullptr, r1c,
                                             r2c, mav, a.amask != 0,
                                             ix, thy, thz, bnc, bnax, cf.dtg, out);
            if constexpr (W) {
                if (ST != GEN_PREDICT && trim) accumulate_w<NC>(acc, out, kc, trim_x, pw1, pw2);
            } else {
                if (ST != GEN_PREDICT && trim) accumulate<NC>(acc, out, kc, trim_x);
            }
            const uint32_t of = zo(k) + rowf;
#pragma unroll
            for (int q = 0; q < NC; q++) st4_state<KEEP>(a.out[q], of, out[q]);
        }

        // rotate the register planes; ghost fixes of the planes loaded above
#pragma unroll
        for (int q = 0; q < NC; q++) {
            mp[q] = mc[q];
            mc[q] = mn[q];
            xc[q] = xn[q];
            gc[q] = gn[q];
            hc[q] = hn[q];
        }
        kp = kc;
        kc = kn;
        kn = kq;
        kxc = kxn;
        hkc = hkn;
        hkn = hkq;
        if constexpr (COEF) {
            dp = dcv;
            dcv = dnv;
            hdc = hdn;
            xdc = xdn;
        }
        fix_centre(mc, kc, dcv, gc);
    }
}

// W: phase-weighted sums, 2*NC per workgroup (the NC sums weighted by w[0],
// then the NC weighted by w[1], as field_sum_kernel lays them out)
// The coefficient variant of the corrector (and of the upwind stage) holds
// just over 168 VGPRs, i.e. 2 workgroups per CU; for one or two components
// the corrector is asked to fit 3 (168 VGPRs, at most one register spilled:
// 188.5 against 194.1 us per MacCormack step at 256^3).  Three components
// would spill dozens and keep 2.
template <int NC, bool VEL, int ST, bool COEF, int LX, bool BRK = false, bool W = false, bool KEEP = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((COEF && !W && ST == GEN_CORRECT && NC <= 2) ? 3 : 1))) gsweep_kernel(GSweepArgs a)
{
    constexpr int NS = W ? 2 * NC : NC;
    __shared__ double sred[4][NS];
    double acc[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) acc[q] = 0.0;
    gsweep_tile<NC, VEL, ST, COEF, LX, BRK, W, KEEP>(a, xcd_tile(blockIdx.x, a.nblocks), acc);
    if (ST != GEN_PREDICT && a.partials) block_reduce<NS, 256>(acc, sred, a.partials + (size_t)blockIdx.x * NS);
}

__global__ void __launch_bounds__(256) zero_solids_kernel(float* m0, float* m1, float* m2, int nc,
                                                          const uint8_t* __restrict__ mask, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        if (mask[i] != 0) continue;
        m0[i] = 0.f;
        if (nc > 1) m1[i] = 0.f;
        if (nc > 2) m2[i] = 0.f;
    }
}

// zero_solids_kernel for fields in bricks: voxel i of the row order (mask
// index) is cleared at its brick offset.
__global__ void __launch_bounds__(256) zero_solids_brick_kernel(float* m0, float* m1, float* m2, int nc,
                                                                const uint8_t* __restrict__ mask, int ny, int pitch,
                                                                size_t n)
{
    const int nby = (ny + 3) / 4, nb8 = (pitch / 4 + 1) / 2;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        if (mask[i] != 0) continue;
        const int x = (int)(i % (size_t)pitch);
        const size_t r = i / (size_t)pitch;
        const size_t o = brick_off(x, (int)(r % (size_t)ny), (int)(r / (size_t)ny), nby, nb8);
        m0[o] = 0.f;
        if (nc > 1) m1[o] = 0.f;
        if (nc > 2) m2[o] = 0.f;
    }
}

// Copies of n 16-byte chunks (dwordx4 loads and stores): the streaming
// ceiling the sweeps are compared with is the fastest of these shapes.
//   KIND 0: grid-stride loop, four chunks a grid-stride apart in flight per
//           thread, non-temporal;
//   KIND 1: one chunk per thread, one pass;
//   KIND 2: one chunk per thread, non-temporal;
//   KIND 3: four consecutive chunks per thread (64 B; a wavefront moves
//           4 KiB of consecutive bytes), non-temporal.
template <int KIND>
__global__ void __launch_bounds__(256) copy_kernel(const f4* __restrict__ src, f4* __restrict__ dst, size_t n)
{
    if constexpr (KIND == 0) {
        const size_t stride = (size_t)gridDim.x * 256;
        size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
        for (; i + 3 * stride < n; i += 4 * stride) {
            const f4 a = __builtin_nontemporal_load(src + i);
            const f4 b = __builtin_nontemporal_load(src + i + stride);
            const f4 c = __builtin_nontemporal_load(src + i + 2 * stride);
            const f4 d = __builtin_nontemporal_load(src + i + 3 * stride);
            __builtin_nontemporal_store(a, dst + i);
            __builtin_nontemporal_store(b, dst + i + stride);
            __builtin_nontemporal_store(c, dst + i + 2 * stride);
            __builtin_nontemporal_store(d, dst + i + 3 * stride);
        }
        for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    } else if constexpr (KIND == 1 || KIND == 2) {
        const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
        if (i >= n) return;
        if constexpr (KIND == 1) dst[i] = src[i];
        else __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    } else {
        // wavefront w of the grid covers chunks [256 w, 256 w + 256); load j
        // of lane l takes chunk 256 w + 64 j + l, so each load instruction of
        // the wavefront reads 1 KiB of consecutive bytes
        const size_t base = ((size_t)blockIdx.x * 256 + (threadIdx.x & ~63u)) * 4 + (threadIdx.x & 63u);
        f4 v[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (base + 64 * j < n) v[j] = __builtin_nontemporal_load(src + base + 64 * j);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (base + 64 * j < n) __builtin_nontemporal_store(v[j], dst + base + 64 * j);
    }
}

int copy_kinds() { return 4; }

hipError_t launch_copy(int kind, const float* src, float* dst, size_t nchunks, int cus, hipStream_t s)
{
    const f4* a = reinterpret_cast<const f4*>(src);
    f4* b = reinterpret_cast<f4*>(dst);
    const size_t one = (nchunks + 255) / 256, four = (nchunks + 1023) / 1024;
    if (one > 0x7fffffffULL) return hipErrorInvalidValue;
    switch (kind) {
    case 0:
        hipLaunchKernelGGL(copy_kernel<0>, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((size_t)cus * 16, four))),
                           dim3(256), 0, s, a, b, nchunks);
        break;
    case 1: hipLaunchKernelGGL(copy_kernel<1>, dim3((unsigned)one), dim3(256), 0, s, a, b, nchunks); break;
    case 2: hipLaunchKernelGGL(copy_kernel<2>, dim3((unsigned)one), dim3(256), 0, s, a, b, nchunks); break;
    case 3: hipLaunchKernelGGL(copy_kernel<3>, dim3((unsigned)four), dim3(256), 0, s, a, b, nchunks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launch helpers

struct VariantShape {
    int lx, rpt;
};
static const VariantShape kVariants[] = {
    {64, 1}, // TX 256, TY 4
    {64, 2}, // TX 256, TY 8
    {32, 1}, // TX 128, TY 8
    {32, 2}, // TX 128, TY 16
    {16, 1}, // TX 64,  TY 16
    {16, 2}, // TX 64,  TY 32
    {8, 1},  // TX 32,  TY 32
};
static const int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
// two-step sweeps: lanes per row (tile width 4*lx covers the whole x extent)
static const int kVariants2[] = {64, 32, 16};
static const int kNumVariants2 = (int)(sizeof(kVariants2) / sizeof(kVariants2[0]));

// After the two-step variants: the compacted brick sweep (tile 64 x 16; in
// the row layout it runs as variant 4, the same tile).  Opt-in: at 256^3 it
// issues 34 % fewer VALU instructions than the tile sweep (lanes busy 73 %
// against 42 %) but its waves wait longer, and it measured 73.8 against
// 70.6 us per step.
static const int kCompactVariant = kNumVariants + kNumVariants2;
int num_variants() { return kCompactVariant + 1; }
int compact_variant() { return kCompactVariant; }

bool variant_shape(int v, int* lx, int* rpt, int* tx, int* ty, int* steps)
{
    if (v < 0 || v >= num_variants()) return false;
    if (v == kCompactVariant) v = 4;
    if (v < kNumVariants) {
        *lx = kVariants[v].lx;
        *rpt = kVariants[v].rpt;
        *tx = 4 * kVariants[v].lx;
        *ty = (256 / kVariants[v].lx) * kVariants[v].rpt;
        *steps = 1;
    } else {
        *lx = kVariants2[v - kNumVariants];
        *rpt = 1;
        *tx = 4 * *lx;
        *ty = 512 / *lx - 2;
        *steps = 2;
    }
    return true;
}

// Row layout [Z][Y][4*nc4] floats (nc4 4-voxel chunks per row) <-> bricks
// of 8 x-voxels by 4 rows: brick (z, y/4, c/2) holds chunk c of rows
// y&~3 .. +3 at float (y&3)*8 + (c&1)*4.  One thread per 16-byte chunk.
__global__ void __launch_bounds__(256) brick_kernel(const float* __restrict__ src, float* __restrict__ dst, int nc4,
                                                    int Y, int Z, int to_brick)
{
    const long nch = (long)nc4 * Y * Z;
    const long nby = (Y + 3) / 4, nb8 = (nc4 + 1) / 2;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nch; i += (long)gridDim.x * 256) {
        const int cx = (int)(i % nc4);
        const long r = i / nc4;
        const int y = (int)(r % Y), z = (int)(r / Y);
        const size_t ro = (size_t)i * 4;
        const size_t bo = (((size_t)z * nby + (y >> 2)) * nb8 + (cx >> 1)) * 32 + (size_t)((y & 3) * 8 + (cx & 1) * 4);
        if (to_brick)
            *reinterpret_cast<f4*>(dst + bo) = *reinterpret_cast<const f4*>(src + ro);
        else
            *reinterpret_cast<f4*>(dst + ro) = *reinterpret_cast<const f4*>(src + bo);
    }
}

size_t brick_floats(int nc4, int Y, int Z) { return (size_t)((nc4 + 1) / 2) * ((Y + 3) / 4) * Z * 32; }

hipError_t launch_brick(const float* src, float* dst, int nbx, int Y, int Z, bool to_brick, hipStream_t s)
{
    const long nch = (long)nbx * Y * Z;
    const unsigned grid = (unsigned)std::max<long>(1, std::min<long>((nch + 255) / 256, 16384));
    hipLaunchKernelGGL(brick_kernel, dim3(grid), dim3(256), 0, s, src, dst, nbx, Y, Z, to_brick ? 1 : 0);
    return hipGetLastError();
}

bool brick_variant(int v) { return (v >= 0 && v <= 6) || v == kCompactVariant; }
// (the compacted sweep is opt-in: it measured slower than the 64 x 16 tile)
int brick_auto_variant(int nx) { return nx > 32 ? 4 : 6; }

template <int NC, bool VEL, bool EXACT>
static hipError_t launch_nc_brick(int v, const SweepArgs& a, hipStream_t s, bool keep)
{
    const dim3 grid(a.nblocks), block(256);
    if (keep && v == 4) { // the brick tile the round model picks; other shapes keep the hint
        hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 16, 1, true, true>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    if (v == kCompactVariant) {
        hipLaunchKernelGGL((csweep_kernel<NC, VEL, EXACT>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    switch (v) {
    case 0: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 64, 1, true>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 64, 2, true>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 32, 1, true>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 32, 2, true>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 16, 1, true>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 16, 2, true>), grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 8, 1, true>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NC, bool VEL, bool EXACT>
static hipError_t launch_nc(int v, const SweepArgs& a, hipStream_t s)
{
    const dim3 grid(a.nblocks), block(256);
    switch (v) {
    case 0: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 64, 1>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 64, 2>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 32, 1>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 32, 2>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 16, 1>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 16, 2>), grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL((sweep_kernel<NC, VEL, EXACT, 8, 1>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sweep(int nc, bool vel, bool exact, int variant, const SweepArgs& a, hipStream_t s, bool brick,
                        bool keep)
{
    if (!brick && variant == kCompactVariant) variant = 4; // the same tile in rows
    if (brick) {
        if (nc == 1)
            return vel ? launch_nc_brick<1, true, false>(variant, a, s, keep)
                       : launch_nc_brick<1, false, false>(variant, a, s, keep);
        if (nc == 2) {
            if (exact)
                return vel ? launch_nc_brick<2, true, true>(variant, a, s, keep)
                           : launch_nc_brick<2, false, true>(variant, a, s, keep);
            return vel ? launch_nc_brick<2, true, false>(variant, a, s, keep)
                       : launch_nc_brick<2, false, false>(variant, a, s, keep);
        }
        if (nc == 3)
            return vel ? launch_nc_brick<3, true, false>(variant, a, s, keep)
                       : launch_nc_brick<3, false, false>(variant, a, s, keep);
        return hipErrorInvalidValue;
    }
    if (nc == 1) return vel ? launch_nc<1, true, false>(variant, a, s) : launch_nc<1, false, false>(variant, a, s);
    if (nc == 2) {
        if (exact) return vel ? launch_nc<2, true, true>(variant, a, s) : launch_nc<2, false, true>(variant, a, s);
        return vel ? launch_nc<2, true, false>(variant, a, s) : launch_nc<2, false, false>(variant, a, s);
    }
    if (nc == 3) return vel ? launch_nc<3, true, false>(variant, a, s) : launch_nc<3, false, false>(variant, a, s);
    return hipErrorInvalidValue;
}

template <int NC, bool VEL, bool EXACT>
static hipError_t launch2_nc(int lx, const Sweep2Args& a, hipStream_t s)
{
    const dim3 grid(a.nblocks), block(512);
    switch (lx) {
    case 64: hipLaunchKernelGGL((sweep2_kernel<NC, VEL, EXACT, 64>), grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL((sweep2_kernel<NC, VEL, EXACT, 32>), grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL((sweep2_kernel<NC, VEL, EXACT, 16>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sweep2(int nc, bool vel, bool exact, int lx, const Sweep2Args& a, hipStream_t s)
{
    if (nc == 1) return vel ? launch2_nc<1, true, false>(lx, a, s) : launch2_nc<1, false, false>(lx, a, s);
    if (nc == 2) {
        if (exact) return vel ? launch2_nc<2, true, true>(lx, a, s) : launch2_nc<2, false, true>(lx, a, s);
        return vel ? launch2_nc<2, true, false>(lx, a, s) : launch2_nc<2, false, false>(lx, a, s);
    }
    if (nc == 3) return vel ? launch2_nc<3, true, false>(lx, a, s) : launch2_nc<3, false, false>(lx, a, s);
    return hipErrorInvalidValue;
}

template <int NC, bool VEL, bool EXACT>
static hipError_t launchb_nc(int v, int grid, const BatchArgs& b, hipStream_t s, bool brick)
{
    const dim3 g(grid), block(256);
#define NMR_BV(LX_, RPT_)                                                                                    \
    if (brick) hipLaunchKernelGGL((batch_kernel<NC, VEL, EXACT, LX_, RPT_, true>), g, block, 0, s, b);      \
    else hipLaunchKernelGGL((batch_kernel<NC, VEL, EXACT, LX_, RPT_>), g, block, 0, s, b);
    switch (v) {
    case 0: NMR_BV(64, 1) break;
    case 1: NMR_BV(64, 2) break;
    case 2: NMR_BV(32, 1) break;
    case 3: NMR_BV(32, 2) break;
    case 4: NMR_BV(16, 1) break;
    case 5: NMR_BV(16, 2) break;
    case 6: NMR_BV(8, 1) break;
    default: return hipErrorInvalidValue;
    }
#undef NMR_BV
    return hipGetLastError();
}

hipError_t launch_batch(int nc, bool vel, bool exact, int variant, int grid, const BatchArgs& b, hipStream_t s,
                        bool brick)
{
    if (nc == 1)
        return vel ? launchb_nc<1, true, false>(variant, grid, b, s, brick)
                   : launchb_nc<1, false, false>(variant, grid, b, s, brick);
    if (nc == 2) {
        if (exact)
            return vel ? launchb_nc<2, true, true>(variant, grid, b, s, brick)
                       : launchb_nc<2, false, true>(variant, grid, b, s, brick);
        return vel ? launchb_nc<2, true, false>(variant, grid, b, s, brick)
                   : launchb_nc<2, false, false>(variant, grid, b, s, brick);
    }
    if (nc == 3)
        return vel ? launchb_nc<3, true, false>(variant, grid, b, s, brick)
                   : launchb_nc<3, false, false>(variant, grid, b, s, brick);
    return hipErrorInvalidValue;
}

// Workgroups of a kernel resident on the whole device at once.
static int resident_blocks(const void* f)
{
    int per_cu = 0;
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 256, 0) != hipSuccess) return -1;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    return per_cu * ncu;
}

int batch_resident_blocks(int nc, bool vel, bool exact, int variant, bool brick)
{
    const void* f = nullptr;
#define NMR_PICK(NC_, V_, E_, B_)                                                               \
    switch (variant) {                                                                          \
    case 0: f = (const void*)batch_kernel<NC_, V_, E_, 64, 1, B_>; break;                       \
    case 1: f = (const void*)batch_kernel<NC_, V_, E_, 64, 2, B_>; break;                       \
    case 2: f = (const void*)batch_kernel<NC_, V_, E_, 32, 1, B_>; break;                       \
    case 3: f = (const void*)batch_kernel<NC_, V_, E_, 32, 2, B_>; break;                       \
    case 4: f = (const void*)batch_kernel<NC_, V_, E_, 16, 1, B_>; break;                       \
    case 5: f = (const void*)batch_kernel<NC_, V_, E_, 16, 2, B_>; break;                       \
    case 6: f = (const void*)batch_kernel<NC_, V_, E_, 8, 1, B_>; break;                        \
    default: return -1;                                                                         \
    }
#define NMR_PICKB(NC_, V_, E_) if (brick) { NMR_PICK(NC_, V_, E_, true) } else { NMR_PICK(NC_, V_, E_, false) }
    if (nc == 1) { if (vel) { NMR_PICKB(1, true, false) } else { NMR_PICKB(1, false, false) } }
    else if (nc == 2 && exact) { if (vel) { NMR_PICKB(2, true, true) } else { NMR_PICKB(2, false, true) } }
    else if (nc == 2) { if (vel) { NMR_PICKB(2, true, false) } else { NMR_PICKB(2, false, false) } }
    else if (nc == 3) { if (vel) { NMR_PICKB(3, true, false) } else { NMR_PICKB(3, false, false) } }
    else return -1;
#undef NMR_PICKB
#undef NMR_PICK
    return resident_blocks(f);
}

int sweep_resident_blocks(int nc, bool vel, bool exact, int variant, bool brick, bool keep)
{
    const void* f = nullptr;
    if (keep && brick && variant == 4) { // the kernel launch_sweep runs with keep
        if (nc == 1) f = vel ? (const void*)sweep_kernel<1, true, false, 16, 1, true, true>
                             : (const void*)sweep_kernel<1, false, false, 16, 1, true, true>;
        else if (nc == 2 && exact) f = vel ? (const void*)sweep_kernel<2, true, true, 16, 1, true, true>
                                           : (const void*)sweep_kernel<2, false, true, 16, 1, true, true>;
        else if (nc == 2) f = vel ? (const void*)sweep_kernel<2, true, false, 16, 1, true, true>
                                  : (const void*)sweep_kernel<2, false, false, 16, 1, true, true>;
        else if (nc == 3) f = vel ? (const void*)sweep_kernel<3, true, false, 16, 1, true, true>
                                  : (const void*)sweep_kernel<3, false, false, 16, 1, true, true>;
        return resident_blocks(f);
    }
    if (variant == kCompactVariant && brick) {
        if (nc == 1) f = vel ? (const void*)csweep_kernel<1, true, false> : (const void*)csweep_kernel<1, false, false>;
        else if (nc == 2 && exact)
            f = vel ? (const void*)csweep_kernel<2, true, true> : (const void*)csweep_kernel<2, false, true>;
        else if (nc == 2) f = vel ? (const void*)csweep_kernel<2, true, false> : (const void*)csweep_kernel<2, false, false>;
        else if (nc == 3) f = vel ? (const void*)csweep_kernel<3, true, false> : (const void*)csweep_kernel<3, false, false>;
        return resident_blocks(f);
    }
    if (variant == kCompactVariant) variant = 4;
#define NMR_PICK(NC_, V_, E_, B_)                                                               \
    switch (variant) {                                                                          \
    case 0: f = (const void*)sweep_kernel<NC_, V_, E_, 64, 1, B_>; break;                       \
    case 1: f = (const void*)sweep_kernel<NC_, V_, E_, 64, 2, B_>; break;                       \
    case 2: f = (const void*)sweep_kernel<NC_, V_, E_, 32, 1, B_>; break;                       \
    case 3: f = (const void*)sweep_kernel<NC_, V_, E_, 32, 2, B_>; break;                       \
    case 4: f = (const void*)sweep_kernel<NC_, V_, E_, 16, 1, B_>; break;                       \
    case 5: f = (const void*)sweep_kernel<NC_, V_, E_, 16, 2, B_>; break;                       \
    case 6: f = (const void*)sweep_kernel<NC_, V_, E_, 8, 1, B_>; break;                        \
    default: return -1;                                                                         \
    }
#define NMR_PICKB(NC_, V_, E_) if (brick) { NMR_PICK(NC_, V_, E_, true) } else { NMR_PICK(NC_, V_, E_, false) }
    if (nc == 1) { if (vel) { NMR_PICKB(1, true, false) } else { NMR_PICKB(1, false, false) } }
    else if (nc == 2 && exact) { if (vel) { NMR_PICKB(2, true, true) } else { NMR_PICKB(2, false, true) } }
    else if (nc == 2) { if (vel) { NMR_PICKB(2, true, false) } else { NMR_PICKB(2, false, false) } }
    else if (nc == 3) { if (vel) { NMR_PICKB(3, true, false) } else { NMR_PICKB(3, false, false) } }
    else return -1;
#undef NMR_PICKB
#undef NMR_PICK
    return resident_blocks(f);
}

int gsweep_resident_blocks(int nc, bool vel, bool coef, int lanes, bool brick, bool weighted, bool upwind)
{
    // the kernel a step launches with the most registers: the upwind stage,
    // or the MacCormack corrector (the predictor holds fewer); on bricks with
    // fused phase-weighted sums its instantiation with 2*NC accumulators
    const void* f = nullptr;
#define NMR_PICK_ST(NC_, V_, ST_, C_)                                                                \
    if (brick && weighted) f = (const void*)gsweep_kernel<NC_, V_, ST_, C_, 16, true, true>;          \
    else if (brick) f = (const void*)gsweep_kernel<NC_, V_, ST_, C_, 16, true>;                       \
    else if (lanes == 64) f = (const void*)gsweep_kernel<NC_, V_, ST_, C_, 64>;                       \
    else if (lanes == 32) f = (const void*)gsweep_kernel<NC_, V_, ST_, C_, 32>;                       \
    else return -1;
#define NMR_PICK(NC_, V_, C_)                                                                         \
    if (upwind) { NMR_PICK_ST(NC_, V_, GEN_UPWIND, C_) } else { NMR_PICK_ST(NC_, V_, GEN_CORRECT, C_) }
#define NMR_PICKC(NC_)                                                                                \
    if (coef) { if (vel) { NMR_PICK(NC_, true, true) } else { NMR_PICK(NC_, false, true) } }          \
    else { if (vel) { NMR_PICK(NC_, true, false) } else { NMR_PICK(NC_, false, false) } }
    if (nc == 1) { NMR_PICKC(1) }
    else if (nc == 2) { NMR_PICKC(2) }
    else if (nc == 3) { NMR_PICKC(3) }
    else return -1;
#undef NMR_PICKC
#undef NMR_PICK
#undef NMR_PICK_ST
    return resident_blocks(f);
}

hipError_t launch_batch_reduce(const double* part, int ntiles, int nc, const int* off, const int* ids,
                               const int* step_base, int row0, int64_t rows, double* signal, int nsteps, int ninst,
                               hipStream_t s)
{
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(nsteps, ninst), dim3(64), 0, s, part, ntiles, nc, off, ids,
                       step_base, row0, rows, signal);
    return hipGetLastError();
}

hipError_t launch_step_advance(int* step_base, int n, hipStream_t s)
{
    hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, s, step_base, n);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* part, int nsteps, int nwg, int nc, double* out, hipStream_t s)
{
    if (nsteps <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(nsteps), dim3(256), 0, s, part, nwg, nc, out);
    return hipGetLastError();
}

hipError_t launch_general(int nc, int stage, const GenArgs& a, hipStream_t s)
{
    const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz), block(256);
#define NMR_GEN(NC_)                                                                                 \
    switch (stage) {                                                                                 \
    case GEN_UPWIND: hipLaunchKernelGGL((general_kernel<NC_, GEN_UPWIND>), grid, block, 0, s, a); break; \
    case GEN_PREDICT: hipLaunchKernelGGL((general_kernel<NC_, GEN_PREDICT>), grid, block, 0, s, a); break; \
    case GEN_CORRECT: hipLaunchKernelGGL((general_kernel<NC_, GEN_CORRECT>), grid, block, 0, s, a); break; \
    default: return hipErrorInvalidValue;                                                            \
    }
    if (nc == 1) { NMR_GEN(1) }
    else if (nc == 2) { NMR_GEN(2) }
    else if (nc == 3) { NMR_GEN(3) }
    else return hipErrorInvalidValue;
#undef NMR_GEN
    return hipGetLastError();
}

int field_sum_blocks(const FieldSumArgs& a)
{
    const long wy = a.by1 - a.by0, wz = a.bz1 - a.bz0;
    const long nrows = (wy > 0 && wz > 0) ? wy * wz : 0;
    return (int)std::max<long>(1, (nrows + kFieldSumRows - 1) / kFieldSumRows);
}

hipError_t launch_field_sum(int nc, const FieldSumArgs& a, hipStream_t s)
{
    const bool w = a.w[0] != nullptr;
    const dim3 grid(field_sum_blocks(a)), block(256);
#define NMR_FS(NC_)                                                                              \
    if (w) hipLaunchKernelGGL((field_sum_kernel<NC_, true>), grid, block, 0, s, a);              \
    else hipLaunchKernelGGL((field_sum_kernel<NC_, false>), grid, block, 0, s, a);
    switch (nc) {
    case 1: NMR_FS(1) break;
    case 2: NMR_FS(2) break;
    case 3: NMR_FS(3) break;
    default: return hipErrorInvalidValue;
    }
#undef NMR_FS
    return hipGetLastError();
}

int gsweep_lanes(int nx) { return nx > 128 ? 64 : 32; }

template <int NC, bool VEL, bool COEF>
static hipError_t launch_gs(int stage, int lanes, const GSweepArgs& a, hipStream_t s)
{
    if (a.w[0]) return hipErrorInvalidValue; // weighted sums are fused on bricks only
    const dim3 grid(a.nblocks), block(256);
#define NMR_GS(ST_)                                                                                     \
    if (lanes == 64) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, ST_, COEF, 64>), grid, block, 0, s, a); \
    else if (lanes == 32) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, ST_, COEF, 32>), grid, block, 0, s, a); \
    else return hipErrorInvalidValue;
    switch (stage) {
    case GEN_UPWIND: NMR_GS(GEN_UPWIND) break;
    case GEN_PREDICT: NMR_GS(GEN_PREDICT) break;
    case GEN_CORRECT: NMR_GS(GEN_CORRECT) break;
    default: return hipErrorInvalidValue;
    }
#undef NMR_GS
    return hipGetLastError();
}

template <int NC, bool VEL, bool COEF>
static hipError_t launch_gs_brick(int stage, const GSweepArgs& a, hipStream_t s, bool keep)
{
    const dim3 grid(a.nblocks), block(256);
    const bool w = a.w[0] != nullptr; // fused phase-weighted sums
    if (w && !a.w[1]) return hipErrorInvalidValue;
    if (keep) { // the state without the non-temporal hint
        switch (stage) {
        case GEN_UPWIND:
            if (w) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_UPWIND, COEF, 16, true, true, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_UPWIND, COEF, 16, true, false, true>), grid, block, 0, s, a);
            break;
        case GEN_PREDICT:
            hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_PREDICT, COEF, 16, true, false, true>), grid, block, 0, s, a);
            break;
        case GEN_CORRECT:
            if (w) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_CORRECT, COEF, 16, true, true, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_CORRECT, COEF, 16, true, false, true>), grid, block, 0, s, a);
            break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (stage) {
    case GEN_UPWIND:
        if (w) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_UPWIND, COEF, 16, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_UPWIND, COEF, 16, true>), grid, block, 0, s, a);
        break;
    case GEN_PREDICT: hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_PREDICT, COEF, 16, true>), grid, block, 0, s, a); break;
    case GEN_CORRECT:
        if (w) hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_CORRECT, COEF, 16, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((gsweep_kernel<NC, VEL, GEN_CORRECT, COEF, 16, true>), grid, block, 0, s, a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gsweep(int nc, bool vel, int stage, bool coef, int lanes, const GSweepArgs& a, hipStream_t s,
                         bool brick, bool keep)
{
    if (brick) { // lanes 16; the per-voxel arrays are read from their brick copies
        if (lanes != 16) return hipErrorInvalidValue;
#define NMR_GSB(NC_)                                                                         \
    if (coef) return vel ? launch_gs_brick<NC_, true, true>(stage, a, s, keep)                \
                         : launch_gs_brick<NC_, false, true>(stage, a, s, keep);              \
    return vel ? launch_gs_brick<NC_, true, false>(stage, a, s, keep)                         \
               : launch_gs_brick<NC_, false, false>(stage, a, s, keep);
        switch (nc) {
        case 1: NMR_GSB(1)
        case 2: NMR_GSB(2)
        case 3: NMR_GSB(3)
        default: return hipErrorInvalidValue;
        }
#undef NMR_GSB
    }
#define NMR_GSC(NC_)                                                                \
    if (coef) return vel ? launch_gs<NC_, true, true>(stage, lanes, a, s)           \
                         : launch_gs<NC_, false, true>(stage, lanes, a, s);         \
    return vel ? launch_gs<NC_, true, false>(stage, lanes, a, s) : launch_gs<NC_, false, false>(stage, lanes, a, s);
    switch (nc) {
    case 1: NMR_GSC(1)
    case 2: NMR_GSC(2)
    case 3: NMR_GSC(3)
    default: return hipErrorInvalidValue;
    }
#undef NMR_GSC
}

hipError_t launch_zero_solids_brick(float* const* m, int nc, const uint8_t* mask, int ny, int nz, int pitch,
                                    hipStream_t s)
{
    const size_t n = (size_t)nz * ny * pitch;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(zero_solids_brick_kernel, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                       m[0], m[1], m[2], nc, mask, ny, pitch, n);
    return hipGetLastError();
}

hipError_t launch_zero_solids(float* const* m, int nc, const uint8_t* mask, size_t n, hipStream_t s)
{
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(zero_solids_kernel, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s, m[0],
                       m[1], m[2], nc, mask, n);
    return hipGetLastError();
}

} // namespace nmr
