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

#include "nmr_kernels.h"

namespace nmr {

__device__ __forceinline__ void load4(float (&d)[4], const float* p)
{
    const float4 v = *reinterpret_cast<const float4*>(p);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
}

__device__ __forceinline__ void store4(float* p, const float (&s)[4])
{
    float4 v;
    v.x = s[0];
    v.y = s[1];
    v.z = s[2];
    v.w = s[3];
    *reinterpret_cast<float4*>(p) = v;
}

__device__ __forceinline__ float pick(bool in, float nb, float self) { return in ? nb : self; }

__device__ __forceinline__ int mbyte(uint32_t w, int e) { return (int)((w >> (8 * e)) & 0xffu); }

// Block index -> tile index such that XCD x (blocks x, x+8, x+16, ...) gets
// the contiguous tile range [x*q + min(x,r), ...), q = nb/8, r = nb%8.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int xcd = b & 7, ib = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd * q + (xcd < r ? xcd : r) + ib;
}

template <int NC, bool VEL, bool EXACT, int LX, int RPT>
__global__ void __launch_bounds__(256) sweep_kernel(SweepArgs a)
{
    constexpr int NT = 256;
    constexpr int RY = NT / LX;  // thread rows
    constexpr int TX = 4 * LX;   // tile width (voxels)
    constexpr int TY = RY * RPT; // tile height (rows)
    constexpr int LW = TX + 8;   // LDS row: col 3 = left halo, 4..4+TX-1 own, 4+TX = right halo
    constexpr int LH = TY + 2;   // row 0 = y0-1 halo, 1..TY own, TY+1 = y0+TY halo

    __shared__ __attribute__((aligned(16))) float sm[2][NC][LH][LW];
    __shared__ __attribute__((aligned(16))) uint8_t sk[2][LH][LW];
    __shared__ double sred[NT / 64][NC];

    const int tid = threadIdx.x;
    const int lx = tid % LX;
    const int ry = tid / LX;

    const int t = xcd_tile(blockIdx.x, a.nblocks);
    const int tx = t % a.ntx;
    const int tyz = t / a.ntx;
    const int ty = tyz % a.nty;
    const int tz = tyz / a.nty;
    const int x0 = tx * TX, y0 = ty * TY;
    const int k0 = tz * a.lz;
    const int k1 = min(k0 + a.lz, a.nz);
    const int ix = x0 + 4 * lx;
    const bool xin = ix < a.nx;

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t P = (size_t)a.pitch;
    const size_t plane = (size_t)ny * P;

    int iy[RPT];
    bool rin[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        iy[r] = y0 + ry + r * RY;
        rin[r] = xin && (iy[r] < ny);
    }
    // halo roles
    const bool top = (ry == 0);
    const bool bot = (ry == RY - 1);
    const int hy = top ? (y0 - 1) : (y0 + TY);
    const bool hrow = (top || bot) && xin && hy >= 0 && hy < ny;
    const int hl = x0 - 1;  // left halo column
    const int hr = x0 + TX; // right halo column
    const bool hcl = (lx == 0) && hl >= 0;
    const bool hcr = (lx == LX - 1) && hr < nx;

    // register planes
    float mp[RPT][NC][4], mc[RPT][NC][4], mn[RPT][NC][4];
    uint32_t kp[RPT], kc[RPT], kn[RPT];
    float hc[NC][4] = {}, hn[NC][4] = {};
    uint32_t hkc = 0, hkn = 0;
    float xcl[RPT][NC] = {}, xcr[RPT][NC] = {}, xnl[RPT][NC] = {}, xnr[RPT][NC] = {};
    int kxcl[RPT] = {}, kxcr[RPT] = {}, kxnl[RPT] = {}, kxnr[RPT] = {};
    float wc[RPT][4];

    auto load_own = [&](int kk, float (&m)[RPT][NC][4], uint32_t (&kw)[RPT]) {
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (rin[r]) {
                const size_t o = (size_t)kk * plane + (size_t)iy[r] * P + ix;
#pragma unroll
                for (int q = 0; q < NC; q++) load4(m[r][q], a.min[q] + o);
                kw[r] = *reinterpret_cast<const uint32_t*>(a.mask + o);
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++)
#pragma unroll
                    for (int e = 0; e < 4; e++) m[r][q][e] = 0.f;
                kw[r] = 0;
            }
        }
    };
    auto load_halo = [&](int kk, float (&h)[NC][4], uint32_t& hk, float (&cl)[RPT][NC],
                         float (&cr)[RPT][NC], int (&kl)[RPT], int (&kr)[RPT]) {
        if (hrow) {
            const size_t o = (size_t)kk * plane + (size_t)hy * P + ix;
#pragma unroll
            for (int q = 0; q < NC; q++) load4(h[q], a.min[q] + o);
            hk = *reinterpret_cast<const uint32_t*>(a.mask + o);
        }
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const bool yok = iy[r] < ny;
            if (hcl && yok) {
                const size_t o = (size_t)kk * plane + (size_t)iy[r] * P + hl;
#pragma unroll
                for (int q = 0; q < NC; q++) cl[r][q] = a.min[q][o];
                kl[r] = a.mask[o];
            }
            if (hcr && yok) {
                const size_t o = (size_t)kk * plane + (size_t)iy[r] * P + hr;
#pragma unroll
                for (int q = 0; q < NC; q++) cr[r][q] = a.min[q][o];
                kr[r] = a.mask[o];
            }
        }
    };
    auto load_vz = [&](int kk, float (&w)[RPT][4]) {
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (rin[r]) {
                load4(w[r], a.vz + (size_t)kk * plane + (size_t)iy[r] * P + ix);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) w[r][e] = 0.f;
            }
        }
    };

    // prologue: plane k0-1 (clamped), plane k0 and its halo, Vz at face k0
    load_own(k0 > 0 ? k0 - 1 : k0, mp, kp);
    load_own(k0, mc, kc);
    load_halo(k0, hc, hkc, xcl, xcr, kxcl, kxcr);
    if constexpr (VEL) load_vz(k0, wc);

    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc[q] = 0.0;

    const Coef c = a.c;
    const bool trim_x[4] = {ix + 0 >= a.bx0 && ix + 0 < a.bx1, ix + 1 >= a.bx0 && ix + 1 < a.bx1,
                            ix + 2 >= a.bx0 && ix + 2 < a.bx1, ix + 3 >= a.bx0 && ix + 3 < a.bx1};

    for (int k = k0; k < k1; k++) {
        const int buf = (k - k0) & 1;
        const int kn_plane = (k + 1 < nz) ? k + 1 : k;
        // prefetch the next plane (own rows, halo) and the next Vz face
        load_own(kn_plane, mn, kn);
        load_halo(kn_plane, hn, hkn, xnl, xnr, kxnl, kxnr);
        float vxr[RPT][5], vyr[RPT][4], vyr1[RPT][4], wnx[RPT][4];
        if constexpr (VEL) {
            load_vz(k + 1, wnx);
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                if (rin[r]) {
                    const float* px = a.vx + (size_t)k * (size_t)ny * (size_t)a.pitch1 +
                                      (size_t)iy[r] * (size_t)a.pitch1 + ix;
                    float t4[4];
                    load4(t4, px);
#pragma unroll
                    for (int e = 0; e < 4; e++) vxr[r][e] = t4[e];
                    vxr[r][4] = (ix + 4 <= nx) ? px[4] : 0.f;
                    const float* py = a.vy + (size_t)k * (size_t)(ny + 1) * P + (size_t)iy[r] * P + ix;
                    load4(vyr[r], py);
                    load4(vyr1[r], py + P);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        vxr[r][e] = 0.f;
                        vyr[r][e] = 0.f;
                        vyr1[r][e] = 0.f;
                    }
                    vxr[r][4] = 0.f;
                }
            }
        }

        // stage plane k into LDS
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int lr = 1 + ry + r * RY;
#pragma unroll
            for (int q = 0; q < NC; q++) store4(&sm[buf][q][lr][4 + 4 * lx], mc[r][q]);
            *reinterpret_cast<uint32_t*>(&sk[buf][lr][4 + 4 * lx]) = kc[r];
            if (hcl) {
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][3] = xcl[r][q];
                sk[buf][lr][3] = (uint8_t)kxcl[r];
            }
            if (hcr) {
#pragma unroll
                for (int q = 0; q < NC; q++) sm[buf][q][lr][4 + TX] = xcr[r][q];
                sk[buf][lr][4 + TX] = (uint8_t)kxcr[r];
            }
        }
        if (hrow) {
            const int lr = top ? 0 : TY + 1;
#pragma unroll
            for (int q = 0; q < NC; q++) store4(&sm[buf][q][lr][4 + 4 * lx], hc[q]);
            *reinterpret_cast<uint32_t*>(&sk[buf][lr][4 + 4 * lx]) = hkc;
        }
        __syncthreads();

        const bool zm_in = k > 0, zp_in = k + 1 < nz;
        const bool trim_z = k >= a.bz0 && k < a.bz1;
        float thx = 0.f, thy = 0.f, thz = 0.f;
        if (NC == 3) {
            thz = c.bgz * ((float)k + 0.5f);
        } else if (NC == 2) {
            thz = c.bgz * ((float)k + 0.5f);
        }

#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (!rin[r]) continue;
            const int lr = 1 + ry + r * RY;
            const int y = iy[r];
            const bool ym_in = y > 0, yp_in = y + 1 < ny;
            const bool trim_y = y >= a.by0 && y < a.by1;
            if (NC == 3) thy = c.bgy * ((float)y + 0.5f);
            // neighbours from LDS
            float ym[NC][4], yp[NC][4], xl[NC], xr[NC];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const float4 a4 = *reinterpret_cast<const float4*>(&sm[buf][q][lr - 1][4 + 4 * lx]);
                const float4 b4 = *reinterpret_cast<const float4*>(&sm[buf][q][lr + 1][4 + 4 * lx]);
                ym[q][0] = a4.x; ym[q][1] = a4.y; ym[q][2] = a4.z; ym[q][3] = a4.w;
                yp[q][0] = b4.x; yp[q][1] = b4.y; yp[q][2] = b4.z; yp[q][3] = b4.w;
                xl[q] = sm[buf][q][lr][3 + 4 * lx];
                xr[q] = sm[buf][q][lr][8 + 4 * lx];
            }
            const uint32_t kym = *reinterpret_cast<const uint32_t*>(&sk[buf][lr - 1][4 + 4 * lx]);
            const uint32_t kyp = *reinterpret_cast<const uint32_t*>(&sk[buf][lr + 1][4 + 4 * lx]);
            const int kxl = sk[buf][lr][3 + 4 * lx];
            const int kxr = sk[buf][lr][8 + 4 * lx];

            float out[NC][4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int x = ix + e;
                const bool pore = mbyte(kc[r], e) != 0;
#pragma unroll
                for (int q = 0; q < NC; q++) out[q][e] = mc[r][q][e];
                if (!(x < nx) || !pore) continue;
                const bool xm_in = x > 0, xp_in = x + 1 < nx;
                // neighbour masks (out of domain -> the centre's, i.e. pore)
                const int mxm = xm_in ? (e > 0 ? mbyte(kc[r], e - 1) : kxl) : 1;
                const int mxp = xp_in ? (e < 3 ? mbyte(kc[r], e + 1) : kxr) : 1;
                const int mym = ym_in ? mbyte(kym, e) : 1;
                const int myp = yp_in ? mbyte(kyp, e) : 1;
                const int mzm = zm_in ? mbyte(kp[r], e) : 1;
                const int mzp = zp_in ? mbyte(kn[r], e) : 1;
                const float sx = (float)((mxm == 0) + (mxp == 0));
                const float sy = (float)((mym == 0) + (myp == 0));
                const float sz = (float)((mzm == 0) + (mzp == 0));
                const float rate2 = c.r2 + c.s2x * sx + c.s2y * sy + c.s2z * sz;
                const float rate1 = c.r1 + c.s1x * sx + c.s1y * sy + c.s1z * sz;
                const float pxm = (mxm != 0) ? 1.f : 0.f, pxp = (mxp != 0) ? 1.f : 0.f;
                const float pym = (mym != 0) ? 1.f : 0.f, pyp = (myp != 0) ? 1.f : 0.f;
                const float pzm = (mzm != 0) ? 1.f : 0.f, pzp = (mzp != 0) ? 1.f : 0.f;

                float um = 0.f, up = 0.f, vm = 0.f, vp = 0.f, wm = 0.f, wp = 0.f;
                if constexpr (VEL) {
                    um = vxr[r][e];
                    up = vxr[r][e + 1];
                    vm = vyr[r][e];
                    vp = vyr1[r][e];
                    wm = wc[r][e];
                    wp = wnx[r][e];
                }
                const float u = 0.5f * (um + up), v = 0.5f * (vm + vp), w = 0.5f * (wm + wp);
                const float un = fminf(u, 0.f), upos = fmaxf(u, 0.f);
                const float vn = fminf(v, 0.f), vpos = fmaxf(v, 0.f);
                const float wn_ = fminf(w, 0.f), wpos = fmaxf(w, 0.f);
                const float iup = (up != 0.f) ? 1.f : 0.f, ium = (um != 0.f) ? 1.f : 0.f;
                const float ivp = (vp != 0.f) ? 1.f : 0.f, ivm = (vm != 0.f) ? 1.f : 0.f;
                const float iwp = (wp != 0.f) ? 1.f : 0.f, iwm = (wm != 0.f) ? 1.f : 0.f;

                float m0[NC], d[NC];
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const float mv = mc[r][q][e];
                    m0[q] = mv;
                    const float fxm = xm_in ? (e > 0 ? mc[r][q][e - 1] : xl[q]) : mv;
                    const float fxp = xp_in ? (e < 3 ? mc[r][q][e + 1] : xr[q]) : mv;
                    const float fym = pick(ym_in, ym[q][e], mv);
                    const float fyp = pick(yp_in, yp[q][e], mv);
                    const float fzm = pick(zm_in, mp[r][q][e], mv);
                    const float fzp = pick(zp_in, mn[r][q][e], mv);
                    float dd = c.cdx * (pxp * (fxp - mv) + pxm * (fxm - mv)) +
                               c.cdy * (pyp * (fyp - mv) + pym * (fym - mv)) +
                               c.cdz * (pzp * (fzp - mv) + pzm * (fzm - mv));
                    if constexpr (VEL) {
                        const float adv = c.hx * (un * (fxp - mv) * iup + upos * (mv - fxm) * ium) +
                                          c.hy * (vn * (fyp - mv) * ivp + vpos * (mv - fym) * ivm) +
                                          c.hz * (wn_ * (fzp - mv) * iwp + wpos * (mv - fzm) * iwm);
                        dd -= adv;
                    }
                    d[q] = dd;
                }
                // relaxation
                if constexpr (NC == 3) {
                    const float par = c.n0 * m0[0] + c.n1 * m0[1] + c.n2 * m0[2];
                    const float nn[3] = {c.n0, c.n1, c.n2};
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const float mpar = nn[q] * par;
                        d[q] += rate1 * (c.peq * nn[q] - mpar) + rate2 * (mpar - m0[q]);
                    }
                    thx = c.bgx * ((float)x + 0.5f);
                    // + dt*g*(m x Bg)
                    d[0] += -thy * m0[2] + thz * m0[1];
                    d[1] += thx * m0[2] - thz * m0[0];
                    d[2] += -thx * m0[1] + thy * m0[0];
                } else {
#pragma unroll
                    for (int q = 0; q < NC; q++) d[q] -= rate2 * m0[q];
                    (void)rate1;
                    if constexpr (NC == 2) {
                        if constexpr (EXACT) {
                            const float sn = sinf(thz);
                            const float h = sinf(0.5f * thz);
                            const float cm1 = -2.f * h * h; // cos(th) - 1 without cancellation
                            d[0] += m0[0] * cm1 + m0[1] * sn;
                            d[1] += m0[1] * cm1 - m0[0] * sn;
                        } else {
                            d[0] += thz * m0[1];
                            d[1] -= thz * m0[0];
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < NC; q++) out[q][e] = m0[q] + d[q];
            }
            // store and accumulate the signal of the updated plane
            const size_t o = (size_t)k * plane + (size_t)y * P + ix;
#pragma unroll
            for (int q = 0; q < NC; q++) store4(a.mout[q] + o, out[q]);
            if (trim_z && trim_y) {
#pragma unroll
                for (int q = 0; q < NC; q++) {
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (trim_x[e] && mbyte(kc[r], e) != 0) acc[q] += (double)out[q][e];
                }
            }
        }

        // rotate the register planes
#pragma unroll
        for (int r = 0; r < RPT; r++) {
#pragma unroll
            for (int q = 0; q < NC; q++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    mp[r][q][e] = mc[r][q][e];
                    mc[r][q][e] = mn[r][q][e];
                }
            kp[r] = kc[r];
            kc[r] = kn[r];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                xcl[r][q] = xnl[r][q];
                xcr[r][q] = xnr[r][q];
            }
            kxcl[r] = kxnl[r];
            kxcr[r] = kxnr[r];
            if constexpr (VEL) {
#pragma unroll
                for (int e = 0; e < 4; e++) wc[r][e] = wnx[r][e];
            }
        }
#pragma unroll
        for (int q = 0; q < NC; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) hc[q][e] = hn[q][e];
        hkc = hkn;
    }

    // workgroup reduction of the signal: wave shuffle (FP64), then LDS
#pragma unroll
    for (int q = 0; q < NC; q++) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[q] = v;
    }
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NC; q++) sred[wave][q] = acc[q];
    }
    __syncthreads();
    if (tid < NC) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) s += sred[w][tid];
        a.partials[(size_t)blockIdx.x * NC + tid] = s;
    }
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

// Pore sum of a field over the trimmed interior: one block per z-plane
// writes part[k*NC + q]; nmr_reduce_partials-style second pass sums planes.
template <int NC>
__global__ void __launch_bounds__(256) field_sum_kernel(FieldSumArgs a)
{
    __shared__ double sh[256];
    const int k = blockIdx.x;
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc[q] = 0.0;
    const bool kin = k >= a.bz0 && k < a.bz1;
    if (kin) {
        const size_t base = (size_t)k * (size_t)a.ny * (size_t)a.pitch;
        const int n = a.nx * a.ny;
        for (int t = threadIdx.x; t < n; t += 256) {
            const int y = t / a.nx, x = t - y * a.nx;
            if (y < a.by0 || y >= a.by1 || x < a.bx0 || x >= a.bx1) continue;
            const size_t o = base + (size_t)y * a.pitch + x;
            if (a.mask[o] == 0) continue;
#pragma unroll
            for (int q = 0; q < NC; q++) acc[q] += (double)a.f[q][o];
        }
    }
    for (int q = 0; q < NC; q++) {
        sh[threadIdx.x] = acc[q];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.part[(size_t)k * NC + q] = sh[0];
        __syncthreads();
    }
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

int num_variants() { return kNumVariants; }

bool variant_shape(int v, int* lx, int* rpt, int* tx, int* ty)
{
    if (v < 0 || v >= kNumVariants) return false;
    *lx = kVariants[v].lx;
    *rpt = kVariants[v].rpt;
    *tx = 4 * kVariants[v].lx;
    *ty = (256 / kVariants[v].lx) * kVariants[v].rpt;
    return true;
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

hipError_t launch_sweep(int nc, bool vel, bool exact, int variant, const SweepArgs& a, hipStream_t s)
{
    if (nc == 1) return vel ? launch_nc<1, true, false>(variant, a, s) : launch_nc<1, false, false>(variant, a, s);
    if (nc == 2) {
        if (exact) return vel ? launch_nc<2, true, true>(variant, a, s) : launch_nc<2, false, true>(variant, a, s);
        return vel ? launch_nc<2, true, false>(variant, a, s) : launch_nc<2, false, false>(variant, a, s);
    }
    if (nc == 3) return vel ? launch_nc<3, true, false>(variant, a, s) : launch_nc<3, false, false>(variant, a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_reduce_partials(const double* part, int nsteps, int nwg, int nc, double* out, hipStream_t s)
{
    if (nsteps <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(nsteps), dim3(256), 0, s, part, nwg, nc, out);
    return hipGetLastError();
}

hipError_t launch_field_sum(int nc, const FieldSumArgs& a, int nz, hipStream_t s)
{
    switch (nc) {
    case 1: hipLaunchKernelGGL(field_sum_kernel<1>, dim3(nz), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(field_sum_kernel<2>, dim3(nz), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(field_sum_kernel<3>, dim3(nz), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace nmr
