/*
 * nmr_oracle.c -- FP64 CPU restatement of the reference's per-timestep
 * Bloch-Torrey update.  TEST INFRASTRUCTURE ONLY: this file is the checker.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product (initializator_amd/, libnmr_hip.so) never links it.
 *
 * Layout: the reference's padded SoA layout, x fastest, one ghost cell per
 * side:  idx = i + (ik+2)*(j + (jk+2)*k)   (NmrRelaxationSolve.cpp:3284-3287).
 * Masks are the reference's ASCII '0'/'1' chars.
 *
 * Schemes: or_params.scheme 0 is ExplicitSimple (upwind, the host path),
 * 1 is the MacCormack step of the GPU path (ExplicitMakUpgrade); bc[] selects
 * the ghost fill per axis (zero derivative, periodic, zero value).
 *
 * Parity status of each function (how it is pinned without a reference
 * build -- the reference's host sources include cuda_runtime.h, which this
 * image does not provide (an MPI toolchain is present, MPICH under
 * /opt/conda); building them would need a stand-in for that header, so the
 * reference counts as unbuildable here):
 *   or_explicit_simple   pinned: bulk T1/T2 closed forms, enclosed-voxel
 *                        surface rate, Neumann cosine eigenmode, 1-D upwind
 *                        lines along x, y and z (both signs, zero faces),
 *                        periodic upwind Fourier modes along x, y and z,
 *                        precession (1-i*theta)^n under x, y and z gradients,
 *                        invariance under every permutation of the axes of a
 *                        random anisotropic problem, conservation /
 *                        fixed-point invariants; the
 *                        advect_mask variant: hand values of a pore next to
 *                        a flowing solid face, and agreement with the plain
 *                        step where no pore-solid face carries velocity
 *   or_maccormack        pinned: Heun factors of the relaxation-only,
 *                        precession-only (z and y gradients) and
 *                        diffusion-only (Neumann cosine and zero-value sine
 *                        eigenmodes) problems; the Lax-Wendroff amplification
 *                        factor of periodic 1-D advection along x, y and z;
 *                        the two-pore channel along x, y and z; axis
 *                        permutation invariance; uniform fixed point under flow next to
 *                        solids of arbitrary value (the race resolution);
 *                        solids zeroed after a step
 *   or_bc_field          pinned: Neumann invariants + brute copy check;
 *                        periodic advection (upwind amplification factor);
 *                        zero-value diffusion sine eigenmode
 *   or_bc_mask           pinned: brute copy check
 *   or_fill_coefficients pinned: hand-computed rates on a 3^3 pore in 5^3
 *   or_fill_coefficients_2f, or_phase, or_output_2f  pinned: hand-computed
 *                        two-phase fractions, penalty, rates with distinct
 *                        face relaxivities per side and fluid; gate and
 *                        weights of the two-phase sums by brute force
 *   or_fill_gradient     pinned: closed form G*d*(idx-1/2)
 *   or_glibc_rand        pinned: the C library's own rand() after srand(1)
 *                        (glibc, first 10^5 outputs)
 *   or_fill_noise        pinned: the reference's expression evaluated with
 *                        the C library's rand() in the same cell order
 *   or_output            pinned: brute-force numpy sum with trimming
 *   or_step_time         pinned: formula arithmetic
 *   or_add_impulse       pinned: Rodrigues rotation closed form (sign
 *                        convention), 2*pi identity, norm preservation,
 *                        gradient phase linear in position (flag 1); flag 3
 *                        angle g|a||(r - 1e13 d).a_hat| measured from the
 *                        rotated vectors (distance, its sign, the index
 *                        offset, the |a| factor); flag 4 angle
 *                        -g|(r - r_c)*a| measured likewise, and its
 *                        unnormalised axis
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_SMALL 1.0e-50 /* NmrRelaxationSolve.cpp:809 */
#define OR_BIG 1.0e50    /* NmrRelaxationSolve.cpp:808 */

typedef struct or_params {
    int32_t ik, jk, kk;      /* interior sizes */
    int32_t precession;      /* nVectorMult */
    int32_t velocity;        /* nVelocity */
    int32_t diffusion;       /* nDiffusion */
    double dx, dy, dz, dt;
    double gamma;            /* g = 4.78941714e7 in the reference */
    double p_eq;             /* equilibriumMagnetization */
    double b0[3];            /* uniform mainField */
    int32_t scheme;          /* 0: ExplicitSimple (upwind), 1: MacCormack (ExplicitMakUpgrade) */
    int32_t bc[3];           /* per axis: 0 zero derivative, 1 periodic, 2 zero value */
    int32_t advect_mask;     /* upwind only: the NMR_2d variant (cu:90-232) -- advection
                                across a face only when the neighbour is active
                                (cu:190-210), indicators v/(v + 1e-20), harmonic
                                means without the 1e-50 guard */
} or_params;

#define OR_BC_NEUMANN 0
#define OR_BC_PERIODIC 1
#define OR_BC_DIRICHLET 2
#define OR_MAK_EPS 1.0e-25 /* harmonic-mean guard of NmrRelaxationSolveKernels.cu:726-731 */

static inline long IX(const or_params* p, int i, int j, int k)
{
    return (long)i + (long)(p->ik + 2) * ((long)j + (long)(p->jk + 2) * (long)k);
}

long or_num_cells(const or_params* p)
{
    return (long)(p->ik + 2) * (long)(p->jk + 2) * (long)(p->kk + 2);
}

/* Ghost layer of the mask = copy of the adjacent interior on all six faces
 * (NmrRelaxationSolve.cpp:880-928, single rank: the z exchange degenerates
 * to the rank-0 / last-rank copies). */
void or_bc_mask(const or_params* p, char* act)
{
    for (int k = 0; k < p->kk + 2; k++)
        for (int j = 0; j < p->jk + 2; j++) {
            act[IX(p, 0, j, k)] = act[IX(p, 1, j, k)];
            act[IX(p, p->ik + 1, j, k)] = act[IX(p, p->ik, j, k)];
        }
    for (int k = 0; k < p->kk + 2; k++)
        for (int i = 0; i < p->ik + 2; i++) {
            act[IX(p, i, 0, k)] = act[IX(p, i, 1, k)];
            act[IX(p, i, p->jk + 1, k)] = act[IX(p, i, p->jk, k)];
        }
    for (int j = 0; j < p->jk + 2; j++)
        for (int i = 0; i < p->ik + 2; i++) {
            act[IX(p, i, j, 0)] = act[IX(p, i, j, 1)];
            act[IX(p, i, j, p->kk + 1)] = act[IX(p, i, j, p->kk)];
        }
}

/* Ghost fill of one field component, axis by axis in the order x, y, z
 * over the full extent of the other two axes.  Per axis (bc[]):
 *   zero derivative: ghost = adjacent interior cell
 *       (BC_ZeroDer_X/Y/Z, NmrRelaxationSolveKernels.cu:969-1000; the host
 *       path NmrRelaxationSolve.cpp:1664-1754 at one rank);
 *   periodic: ghost 0 = cell n, ghost n+1 = cell 1 (BC_Inf_X/Y/Z, cu:1002-1033);
 *   zero value: both ghosts 0 (BC_Advanced_XM/XP with Val = 0, cu:1035-1053,
 *       which the reference offers on x only). */
void or_bc_field(const or_params* p, double* f)
{
    const int ik = p->ik, jk = p->jk, kk = p->kk;
    for (int k = 0; k < kk + 2; k++)
        for (int j = 0; j < jk + 2; j++) {
            double lo = f[IX(p, 1, j, k)], hi = f[IX(p, ik, j, k)];
            if (p->bc[0] == OR_BC_PERIODIC) { const double t = lo; lo = hi; hi = t; }
            if (p->bc[0] == OR_BC_DIRICHLET) lo = hi = 0.0;
            f[IX(p, 0, j, k)] = lo;
            f[IX(p, ik + 1, j, k)] = hi;
        }
    for (int k = 0; k < kk + 2; k++)
        for (int i = 0; i < ik + 2; i++) {
            double lo = f[IX(p, i, 1, k)], hi = f[IX(p, i, jk, k)];
            if (p->bc[1] == OR_BC_PERIODIC) { const double t = lo; lo = hi; hi = t; }
            if (p->bc[1] == OR_BC_DIRICHLET) lo = hi = 0.0;
            f[IX(p, i, 0, k)] = lo;
            f[IX(p, i, jk + 1, k)] = hi;
        }
    for (int j = 0; j < jk + 2; j++)
        for (int i = 0; i < ik + 2; i++) {
            double lo = f[IX(p, i, j, 1)], hi = f[IX(p, i, j, kk)];
            if (p->bc[2] == OR_BC_PERIODIC) { const double t = lo; lo = hi; hi = t; }
            if (p->bc[2] == OR_BC_DIRICHLET) lo = hi = 0.0;
            f[IX(p, i, j, 0)] = lo;
            f[IX(p, i, j, kk + 1)] = hi;
        }
}

/* Fluid properties of FillInDiffusinAndTime (water = index 0, oil = 1) and
 * the two-phase projection constants AN(1), AN(2), BN(1), BN(2). */
typedef struct or_fluids {
    double d[2], t1[2], t2[2], r1[2], r2[2];
    int32_t ncomp;           /* 1: (c1, c2) = (conc0, conc1); 2: projection */
    int32_t reserved;
    double an1, an2, bn1, bn2;
} or_fluids;

/* Phase fractions of a cell from its two concentrations
 * (NmrRelaxationSolve.cpp:1780-1791): with NCOMP=2 the distances z1, z2 of
 * (x0, x1) to the points (AN1, AN2) and (BN1, BN2) give c1 = z2/(z1+z2),
 * c2 = z1/(z1+z2); with NCOMP=1 the concentrations are the fractions. */
void or_phase(const or_fluids* f, double x0, double x1, double* c1, double* c2)
{
    if (f->ncomp == 2) {
        const double z1 = sqrt((x0 - f->an1) * (x0 - f->an1) + (x1 - f->an2) * (x1 - f->an2));
        const double z2 = sqrt((x0 - f->bn1) * (x0 - f->bn1) + (x1 - f->bn2) * (x1 - f->bn2));
        *c1 = z2 / (z1 + z2);
        *c2 = z1 / (z1 + z2);
    } else {
        *c1 = x0;
        *c2 = x1;
    }
}

/* Per-voxel diffusivity and relaxation times (NmrRelaxationSolve.cpp:1756-1893),
 * followed by the diffusion ghost fill (cpp:1895-1942).  Concentrations
 * conc0/conc1 (padded; NULL: one fluid, water) give the phase fractions;
 * faces[12] are the face relaxivity arrays 1Water x,y,z, 2Water x,y,z,
 * 1Oil x,y,z, 2Oil x,y,z (padded; face of cell i on the minus side stored at
 * i, the plus side at i+1, as cpp:1798-1841 reads them), a NULL entry
 * meaning the uniform R*_WAT / R*_OIL value (cpp:1217-1231).
 * Ghost entries of D copy the interior; T1/T2 ghosts are left untouched
 * (the reference never reads them).  out2[0]=max D, out2[1]=min T. */
void or_fill_coefficients_2f(const or_params* p, const char* act, const or_fluids* fl, const double* conc0,
                             const double* conc1, const double* const* faces, double* D, double* t1, double* t2,
                             double* out2)
{
    double dmax = 0.0, tmin = 100.0;
    for (int k = 1; k < p->kk + 1; k++)
        for (int j = 1; j < p->jk + 1; j++)
            for (int i = 1; i < p->ik + 1; i++) {
                long c = IX(p, i, j, k);
                if (act[c] == '0') {
                    D[c] = OR_SMALL;
                    t1[c] = OR_BIG;
                    t2[c] = OR_BIG;
                    continue;
                }
                if (act[c] != '1') continue;
                double c1 = 1.0, c2 = 0.0;
                if (conc0) or_phase(fl, conc0[c], conc1[c], &c1, &c2);
                const double cd = c1 * fl->d[0] + c2 * fl->d[1];
                D[c] = cd * pow(1.0 - 2.0 * c1 * c2 / (c1 * c1 + c2 * c2), 10.0);
                if (!p->diffusion) D[c] = OR_SMALL;
                double ct1 = c1 / fl->t1[0] + c2 / fl->t1[1], ct2 = c1 / fl->t2[0] + c2 / fl->t2[1];
                /* face relaxivity on each solid side: (axis, neighbour, face index) */
                const long nb[6] = {IX(p, i - 1, j, k), IX(p, i + 1, j, k), IX(p, i, j - 1, k),
                                    IX(p, i, j + 1, k), IX(p, i, j, k - 1), IX(p, i, j, k + 1)};
                const long fc[6] = {c, nb[1], c, nb[3], c, nb[5]};
                const double h[3] = {p->dx, p->dy, p->dz};
                for (int s = 0; s < 6; s++) {
                    if (act[nb[s]] != '0') continue;
                    const int ax = s / 2;
                    const double* w1 = faces ? faces[ax] : NULL;
                    const double* w2 = faces ? faces[3 + ax] : NULL;
                    const double* o1 = faces ? faces[6 + ax] : NULL;
                    const double* o2 = faces ? faces[9 + ax] : NULL;
                    const double q1w = w1 ? w1[fc[s]] : fl->r1[0], q2w = w2 ? w2[fc[s]] : fl->r2[0];
                    const double q1o = o1 ? o1[fc[s]] : fl->r1[1], q2o = o2 ? o2[fc[s]] : fl->r2[1];
                    ct1 += (c1 * q1w + c2 * q1o) / h[ax];
                    ct2 += (c1 * q2w + c2 * q2o) / h[ax];
                }
                t1[c] = 1.0 / ct1;
                t2[c] = 1.0 / ct2;
                if (dmax <= D[c]) dmax = D[c];
                if (tmin >= t1[c]) tmin = t1[c];
                if (tmin >= t2[c]) tmin = t2[c];
            }
    /* diffusion ghosts copy the adjacent interior whatever the field's
     * boundary condition (BoundaryConditionsForDiffusion, cpp:1895-1942) */
    or_params pn = *p;
    pn.bc[0] = pn.bc[1] = pn.bc[2] = OR_BC_NEUMANN;
    or_bc_field(&pn, D);
    if (out2) {
        out2[0] = dmax;
        out2[1] = tmin;
    }
}

/* One fluid (water, NCOMP=1, NFLUID=1: c1 = 1, c2 = 0) with water face
 * relaxivity arrays sr* (NULL: uniform r1_water / r2_water). */
void or_fill_coefficients(const or_params* p, const char* act, double d_water,
                          double t1_water, double t2_water, double r1_water, double r2_water,
                          const double* sr1x, const double* sr1y, const double* sr1z,
                          const double* sr2x, const double* sr2y, const double* sr2z,
                          double* D, double* t1, double* t2, double* out2)
{
    or_fluids fl;
    memset(&fl, 0, sizeof(fl));
    fl.d[0] = d_water;
    fl.t1[0] = t1_water;
    fl.t2[0] = t2_water;
    fl.r1[0] = r1_water;
    fl.r2[0] = r2_water;
    fl.t1[1] = fl.t2[1] = 1.0; /* c2 = 0: oil terms vanish */
    fl.ncomp = 1;
    const double* faces[12] = {sr1x, sr1y, sr1z, sr2x, sr2y, sr2z, NULL, NULL, NULL, NULL, NULL, NULL};
    or_fill_coefficients_2f(p, act, &fl, NULL, NULL, faces, D, t1, t2, out2);
}

/* Two-phase signal (NmrRelaxationSolve.cpp:3324-3350): over pore cells of
 * the trimmed interior whose phase fraction c1 or c2 exceeds 0.8, the sums
 * of f*c1 and f*c2. */
void or_output_2f(const or_params* p, const char* act, int bx, int by, int bz, const or_fluids* fl,
                  const double* conc0, const double* conc1, const double* f, double* s1, double* s2)
{
    double a = 0.0, b = 0.0;
    for (int k = 1 + bz; k < p->kk + 1 - bz; k++)
        for (int j = 1 + by; j < p->jk + 1 - by; j++)
            for (int i = 1 + bx; i < p->ik + 1 - bx; i++) {
                const long c = IX(p, i, j, k);
                if (act[c] != '1') continue;
                double c1, c2;
                or_phase(fl, conc0[c], conc1[c], &c1, &c2);
                if (c1 > 0.8 || c2 > 0.8) {
                    a += f[c] * c1;
                    b += f[c] * c2;
                }
            }
    *s1 = a;
    *s2 = b;
}

/* Constant-gradient field Bg = (Gx*dx*(i-1/2), Gy*dy*(j-1/2), Gz*dz*(k-1/2))
 * on every padded cell (NmrRelaxationSolve.cpp:1966-1972; no random noise). */
void or_fill_gradient(const or_params* p, double gx, double gy, double gz,
                      double* nfx, double* nfy, double* nfz)
{
    for (int k = 0; k < p->kk + 2; k++)
        for (int j = 0; j < p->jk + 2; j++)
            for (int i = 0; i < p->ik + 2; i++) {
                long c = IX(p, i, j, k);
                nfx[c] = gx * p->dx * (i - 0.5);
                nfy[c] = gy * p->dy * (j - 0.5);
                nfz[c] = gz * p->dz * (k - 0.5);
            }
}

/* glibc's rand() from its default state (no srand: seed 1), restated from
 * the published description of the TYPE_3 additive-feedback generator
 * behind random_r: r[0] = seed, r[i] = 16807 r[i-1] mod (2^31 - 1) for
 * i < 31, r[i] = r[i-31] for i = 31..33, r[i] = r[i-31] + r[i-3] (mod 2^32)
 * afterwards; the first 310 of the latter are discarded and output k is
 * r[344 + k] >> 1.  out[n]. */
void or_glibc_rand(int32_t n, int32_t* out)
{
    uint32_t r[34];
    int64_t w = 1;
    uint32_t init[34];
    init[0] = 1u;
    for (int i = 1; i < 31; i++) {
        w = (16807LL * w) % 2147483647LL;
        if (w < 0) w += 2147483647LL;
        init[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; i++) init[i] = init[i - 31];
    for (int i = 0; i < 34; i++) r[i] = init[i];
    /* ring of the last 34 values: slot i % 34 holds r[i] */
    for (long i = 34; i < 344L + n; i++) {
        const uint32_t v = r[(i - 31) % 34] + r[(i - 3) % 34];
        r[i % 34] = v;
        if (i >= 344) out[i - 344] = (int32_t)(v >> 1);
    }
}

/* The noise field of FillInMainField (NmrRelaxationSolve.cpp:1973-1978):
 * with NDIRECT_FIELD = axis + 1 and AMPL_NOISE = ampl != 0, component
 * `axis` of noiseField becomes ampl * (rand() % 2000 + 1 - 1000) / 1000 on
 * every padded cell, k outermost and i innermost, one rand() per cell,
 * replacing the constant-gradient value there.  nf: that component. */
void or_fill_noise(const or_params* p, double ampl, double* nf)
{
    const long n = or_num_cells(p);
    int32_t* r = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!r) return;
    or_glibc_rand((int32_t)n, r);
    long t = 0;
    for (int k = 0; k < p->kk + 2; k++)
        for (int j = 0; j < p->jk + 2; j++)
            for (int i = 0; i < p->ik + 2; i++)
                nf[IX(p, i, j, k)] = ampl * (double)(r[t++] % 2000 + 1 - 1000) / 1000;
    free(r);
}

/* dt from the stability limits (NmrRelaxationSolve.cpp:2519-2534). */
double or_step_time(double dt_noise, double dt_grad, double dt_vel, double dt_diff,
                    double dt_rel, double time1, double time2)
{
    double a = fmin(fmin(dt_noise, dt_grad), fmin(dt_vel, dt_diff));
    double lim = (time1 != 0.0) ? time1 / 100.0 : time2 / 100.0;
    double b = fmin(dt_rel, lim);
    /* the reference's min() returns the first argument on ties */
    return (a <= b) ? a : b;
}

/* One forward-Euler upwind step, NmrRelaxationSolve.cpp:3148-3282.
 * m[3]: magnetization (read), a[3]: scratch written for interior pores, then
 * every interior cell of m is overwritten by a (Jacobi copy-back).  Cells of
 * a that are not pores keep whatever the caller put there (the reference
 * initialises addMagnetization to the initial M, cpp:2511-2513). */
void or_explicit_simple(const or_params* p, const char* act, const double* D,
                        const double* t1, const double* t2,
                        const double* vx, const double* vy, const double* vz,
                        const double* nfx, const double* nfy, const double* nfz,
                        double* mx, double* my, double* mz,
                        double* ax, double* ay, double* az)
{
    const double dt = p->dt;
    const double hx = dt / p->dx, hy = dt / p->dy, hz = dt / p->dz;
    const double hxx = dt / (p->dx * p->dx), hyy = dt / (p->dy * p->dy), hzz = dt / (p->dz * p->dz);
    const double bx0 = p->b0[0], by0 = p->b0[1], bz0 = p->b0[2];
    double* m[3] = {mx, my, mz};
    double* a[3] = {ax, ay, az};

#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->kk + 1; k++) {
        for (int j = 1; j < p->jk + 1; j++) {
            for (int i = 1; i < p->ik + 1; i++) {
                const long c = IX(p, i, j, k);
                if (act[c] != '1') continue;
                const long xm = IX(p, i - 1, j, k), xp = IX(p, i + 1, j, k);
                const long ym = IX(p, i, j - 1, k), yp = IX(p, i, j + 1, k);
                const long zm = IX(p, i, j, k - 1), zp = IX(p, i, j, k + 1);
                const double tau1 = 1.0 / t1[c], tau2 = 1.0 / t2[c];
                /* harmonic-mean face diffusivities (NMR_2d, cu:214-219: no guard) */
                const double dc = D[c];
                const double gs = p->advect_mask ? 0.0 : OR_SMALL;
                const double gxp = 2.0 * (D[xp] * dc) / (D[xp] + dc + gs);
                const double gxm = 2.0 * (D[xm] * dc) / (D[xm] + dc + gs);
                const double gyp = 2.0 * (D[yp] * dc) / (D[yp] + dc + gs);
                const double gym = 2.0 * (D[ym] * dc) / (D[ym] + dc + gs);
                const double gzp = 2.0 * (D[zp] * dc) / (D[zp] + dc + gs);
                const double gzm = 2.0 * (D[zm] * dc) / (D[zm] + dc + gs);
                /* staggered face velocities and wall indicators (NMR_2d: 1e-20,
                 * cu:143-154); with advect_mask, advection only across faces
                 * whose neighbour is active (cu:190-210) */
                const double ie = p->advect_mask ? 1e-20 : OR_SMALL;
                const double um = vx[c], up = vx[xp];
                const double vm = vy[c], vp = vy[yp];
                const double wm = vz[c], wp = vz[zp];
                double umm = um / (um + ie), upp = up / (up + ie);
                double vmm = vm / (vm + ie), vpp = vp / (vp + ie);
                double wmm = wm / (wm + ie), wpp = wp / (wp + ie);
                if (p->advect_mask) {
                    umm *= act[xm] == '0' ? 0.0 : 1.0;
                    upp *= act[xp] == '0' ? 0.0 : 1.0;
                    vmm *= act[ym] == '0' ? 0.0 : 1.0;
                    vpp *= act[yp] == '0' ? 0.0 : 1.0;
                    wmm *= act[zm] == '0' ? 0.0 : 1.0;
                    wpp *= act[zp] == '0' ? 0.0 : 1.0;
                }
                const double u = 0.5 * (um + up), v = 0.5 * (vm + vp), w = 0.5 * (wm + wp);
                const double ud = fabs(u), vd = fabs(v), wd = fabs(w);
                /* unit main-field direction */
                const double nb = sqrt(bx0 * bx0 + by0 * by0 + bz0 * bz0);
                const double n[3] = {bx0 / nb, by0 / nb, bz0 / nb};
                const double m0[3] = {mx[c], my[c], mz[c]};
                const double par = n[0] * m0[0] + n[1] * m0[1] + n[2] * m0[2];
                const double bb = dt * p->gamma;
                const double bg[3] = {bb * nfx[c], bb * nfy[c], bb * nfz[c]};
                double out[3];
                for (int q = 0; q < 3; q++) {
                    const double mpar = n[q] * par;
                    out[q] = m0[q] + dt * tau1 * (p->p_eq * n[q] - mpar) + dt * tau2 * (mpar - m0[q]);
                }
                if (p->precession) {
                    /* + dt*g*(m x Bg) */
                    out[0] += 0.0 - bg[1] * m0[2] + bg[2] * m0[1];
                    out[1] += 0.0 + bg[0] * m0[2] - bg[2] * m0[0];
                    out[2] += 0.0 - bg[0] * m0[1] + bg[1] * m0[0];
                }
                if (p->velocity) {
                    for (int q = 0; q < 3; q++) {
                        const double* f = m[q];
                        const double c0 = m0[q];
                        out[q] -= 0.5 * hx * ((u - ud) * (f[xp] - c0) * upp + (u + ud) * (c0 - f[xm]) * umm)
                                + 0.5 * hy * ((v - vd) * (f[yp] - c0) * vpp + (v + vd) * (c0 - f[ym]) * vmm)
                                + 0.5 * hz * ((w - wd) * (f[zp] - c0) * wpp + (w + wd) * (c0 - f[zm]) * wmm);
                    }
                }
                if (p->diffusion) {
                    for (int q = 0; q < 3; q++) {
                        const double* f = m[q];
                        const double c0 = m0[q];
                        out[q] += hxx * (gxp * (f[xp] - c0) - gxm * (c0 - f[xm]))
                                + hyy * (gyp * (f[yp] - c0) - gym * (c0 - f[ym]))
                                + hzz * (gzp * (f[zp] - c0) - gzm * (c0 - f[zm]));
                    }
                }
                a[0][c] = out[0];
                a[1][c] = out[1];
                a[2][c] = out[2];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->kk + 1; k++)
        for (int j = 1; j < p->jk + 1; j++)
            for (int i = 1; i < p->ik + 1; i++) {
                const long c = IX(p, i, j, k);
                mx[c] = ax[c];
                my[c] = ay[c];
                mz[c] = az[c];
            }
}

/* Relaxation toward pEq*n (T1 along n, T2 across) and the linearised
 * precession dt*g*(m x Bg) of one cell: the increments both schemes add
 * (NmrRelaxationSolve.cpp:3206-3235; NmrRelaxationSolveKernels.cu:701-733). */
static void or_point_terms(const or_params* p, double tau1, double tau2, const double bg[3], const double m0[3],
                           double inc[3])
{
    const double dt = p->dt;
    const double nb = sqrt(p->b0[0] * p->b0[0] + p->b0[1] * p->b0[1] + p->b0[2] * p->b0[2]);
    const double n[3] = {p->b0[0] / nb, p->b0[1] / nb, p->b0[2] / nb};
    const double par = n[0] * m0[0] + n[1] * m0[1] + n[2] * m0[2];
    for (int q = 0; q < 3; q++) {
        const double mpar = n[q] * par;
        inc[q] = dt * tau1 * (p->p_eq * n[q] - mpar) + dt * tau2 * (mpar - m0[q]);
    }
    if (p->precession) {
        inc[0] += 0.0 - bg[1] * m0[2] + bg[2] * m0[1];
        inc[1] += 0.0 + bg[0] * m0[2] - bg[2] * m0[0];
        inc[2] += 0.0 - bg[0] * m0[1] + bg[1] * m0[0];
    }
}

/* One stage of the MacCormack step for every pore (NMR_predictor_upgrade,
 * NmrRelaxationSolveKernels.cu:637-801, and NMR_corrector_upgrade,
 * cu:803-967).  The stage reads field f (M for the predictor, the predictor
 * P for the corrector), adds relaxation, precession, harmonic-mean
 * diffusion and advection with the cell-centred velocity (u = mean of the
 * two x faces, ...): backward differences in the predictor, forward in the
 * corrector.  out = f + increments (predictor) or
 * 0.5*(M + P + increments(P)) (corrector).
 *
 * Race resolution: before differencing, the reference kernels copy the
 * pore's own value into every solid neighbour (cu:765-782, 931-948), in
 * place and without synchronisation, so the value a pore reads from a solid
 * neighbour shared with other pores is whichever thread wrote last.  The
 * value each writer intends is the reading pore's own, so here every read
 * of a solid neighbour (advection and diffusion) returns the pore's own
 * value.  Ghost cells carry the mask of the adjacent interior cell, so a
 * pore on the domain edge reads its ghost as the boundary condition left it. */
static void or_mak_stage(const or_params* p, int corrector, const char* act, const double* D, const double* t1,
                         const double* t2, const double* vx, const double* vy, const double* vz,
                         const double* nfx, const double* nfy, const double* nfz, double* const m[3],
                         double* const f[3], double* const out[3])
{
    const double dt = p->dt;
    const double hx = dt / p->dx, hy = dt / p->dy, hz = dt / p->dz;
    const double hxx = dt / (p->dx * p->dx), hyy = dt / (p->dy * p->dy), hzz = dt / (p->dz * p->dz);
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->kk + 1; k++)
        for (int j = 1; j < p->jk + 1; j++)
            for (int i = 1; i < p->ik + 1; i++) {
                const long c = IX(p, i, j, k);
                if (act[c] != '1') continue;
                const long nbr[6] = {IX(p, i + 1, j, k), IX(p, i - 1, j, k), IX(p, i, j + 1, k),
                                     IX(p, i, j - 1, k), IX(p, i, j, k + 1), IX(p, i, j, k - 1)};
                const double dc = D[c];
                double g[6];
                for (int s = 0; s < 6; s++) g[s] = 2.0 * (D[nbr[s]] * dc) / (D[nbr[s]] + dc + OR_MAK_EPS);
                const double u = 0.5 * (vx[c] + vx[nbr[0]]), v = 0.5 * (vy[c] + vy[nbr[2]]),
                             w = 0.5 * (vz[c] + vz[nbr[4]]);
                const double b = dt * p->gamma;
                const double bg[3] = {b * nfx[c], b * nfy[c], b * nfz[c]};
                const double f0[3] = {f[0][c], f[1][c], f[2][c]};
                double inc[3];
                or_point_terms(p, 1.0 / t1[c], 1.0 / t2[c], bg, f0, inc);
                for (int q = 0; q < 3; q++) {
                    double nv[6];
                    for (int s = 0; s < 6; s++) nv[s] = act[nbr[s]] == '0' ? f0[q] : f[q][nbr[s]];
                    double diff = 0.0, vel = 0.0;
                    if (p->diffusion)
                        diff = hxx * (g[0] * (nv[0] - f0[q]) - g[1] * (f0[q] - nv[1])) +
                               hyy * (g[2] * (nv[2] - f0[q]) - g[3] * (f0[q] - nv[3])) +
                               hzz * (g[4] * (nv[4] - f0[q]) - g[5] * (f0[q] - nv[5]));
                    if (p->velocity) {
                        if (!corrector)
                            vel = 0.0 - u * hx * (f0[q] - nv[1]) - v * hy * (f0[q] - nv[3]) - w * hz * (f0[q] - nv[5]);
                        else
                            vel = 0.0 - u * hx * (nv[0] - f0[q]) - v * hy * (nv[2] - f0[q]) - w * hz * (nv[4] - f0[q]);
                    }
                    if (!corrector)
                        out[q][c] = f0[q] + inc[q] + diff + vel;
                    else
                        out[q][c] = 0.5 * (m[q][c] + f0[q] + inc[q] + diff + vel);
                }
            }
}

/* One MacCormack step (ExplicitMakUpgrade, NmrRelaxationSolveKernels.cu:1473-1501):
 * predictor into P, ghost fill of P, corrector into C, ghost fill of C, then
 * M = C over the whole padded array.  C is only ever written at pores and at
 * ghosts, so solid cells of M become C's solid value; the reference's C starts
 * zeroed (cu:57-62) and so does the caller's. */
void or_maccormack(const or_params* p, const char* act, const double* D, const double* t1, const double* t2,
                   const double* vx, const double* vy, const double* vz, const double* nfx, const double* nfy,
                   const double* nfz, double* mx, double* my, double* mz, double* px, double* py, double* pz,
                   double* cx, double* cy, double* cz)
{
    double* const m[3] = {mx, my, mz};
    double* const P[3] = {px, py, pz};
    double* const C[3] = {cx, cy, cz};
    or_mak_stage(p, 0, act, D, t1, t2, vx, vy, vz, nfx, nfy, nfz, m, m, P);
    for (int q = 0; q < 3; q++) or_bc_field(p, P[q]);
    or_mak_stage(p, 1, act, D, t1, t2, vx, vy, vz, nfx, nfy, nfz, m, P, C);
    const long num = or_num_cells(p);
    for (int q = 0; q < 3; q++) {
        or_bc_field(p, C[q]);
        memcpy(m[q], C[q], sizeof(double) * num);
    }
}

/* Signal: sum of one component over pore cells of the buffer-trimmed
 * interior, one fluid (NmrRelaxationSolve.cpp:3301-3322).  Returns the
 * sum; *count receives the number of cells summed. */
double or_output(const or_params* p, const char* act, int bx, int by, int bz,
                 const double* f, int32_t* count)
{
    double sum = 0.0;
    int32_t n = 0;
    for (int k = 1 + bz; k < p->kk + 1 - bz; k++)
        for (int j = 1 + by; j < p->jk + 1 - by; j++)
            for (int i = 1 + bx; i < p->ik + 1 - bx; i++) {
                const long c = IX(p, i, j, k);
                if (act[c] == '1') {
                    sum += f[c];
                    n += 1;
                }
            }
    if (count) *count = n;
    return sum;
}

/* Time loop in the order of the reference's GPU driver
 * (NmrRelaxationSolveKernels.cu:1187-1199 with no pulses): update, ghost
 * fill, then the signal of the updated field.  signal[(n)*3 + q] holds the
 * sum of component q after n updates, n = 0..nsteps (n = 0 is the initial
 * state).  grad_sched (nullable, [nsteps][3]) gives the constant-gradient
 * amplitudes in force during each step; the gradient field is rebuilt when
 * they change.  All arrays are padded; a* are scratch copies of m*. */
void or_run(const or_params* p, const char* act, const double* D, const double* t1,
            const double* t2, const double* vx, const double* vy, const double* vz,
            int64_t nsteps, const double* grad_sched, const double* grad_const,
            int bx, int by, int bz,
            double* mx, double* my, double* mz, double* signal)
{
    const long num = or_num_cells(p);
    double* ax = (double*)malloc(sizeof(double) * num);
    double* ay = (double*)malloc(sizeof(double) * num);
    double* az = (double*)malloc(sizeof(double) * num);
    double* nfx = (double*)malloc(sizeof(double) * num);
    double* nfy = (double*)malloc(sizeof(double) * num);
    double* nfz = (double*)malloc(sizeof(double) * num);
    double* pc[6] = {NULL, NULL, NULL, NULL, NULL, NULL}; /* MacCormack P and C, zeroed */
    if (p->scheme == 1)
        for (int q = 0; q < 6; q++) pc[q] = (double*)calloc((size_t)num, sizeof(double));
    memcpy(ax, mx, sizeof(double) * num);
    memcpy(ay, my, sizeof(double) * num);
    memcpy(az, mz, sizeof(double) * num);
    double g_cur[3] = {0.0, 0.0, 0.0};
    if (grad_const) {
        g_cur[0] = grad_const[0];
        g_cur[1] = grad_const[1];
        g_cur[2] = grad_const[2];
    }
    or_fill_gradient(p, g_cur[0], g_cur[1], g_cur[2], nfx, nfy, nfz);
    double* ms[3] = {mx, my, mz};
    /* ghosts of the initial state: the boundary condition of the initial field */
    for (int q = 0; q < 3; q++) or_bc_field(p, ms[q]);
    if (signal)
        for (int q = 0; q < 3; q++) signal[q] = or_output(p, act, bx, by, bz, ms[q], NULL);
    for (int64_t n = 0; n < nsteps; n++) {
        if (grad_sched) {
            const double* gs = grad_sched + 3 * n;
            if (gs[0] != g_cur[0] || gs[1] != g_cur[1] || gs[2] != g_cur[2]) {
                g_cur[0] = gs[0];
                g_cur[1] = gs[1];
                g_cur[2] = gs[2];
                or_fill_gradient(p, g_cur[0], g_cur[1], g_cur[2], nfx, nfy, nfz);
            }
        }
        if (p->scheme == 1)
            or_maccormack(p, act, D, t1, t2, vx, vy, vz, nfx, nfy, nfz, mx, my, mz, pc[0], pc[1], pc[2], pc[3], pc[4],
                          pc[5]);
        else
            or_explicit_simple(p, act, D, t1, t2, vx, vy, vz, nfx, nfy, nfz, mx, my, mz, ax, ay, az);
        or_bc_field(p, mx);
        or_bc_field(p, my);
        or_bc_field(p, mz);
        if (signal)
            for (int q = 0; q < 3; q++)
                signal[(n + 1) * 3 + q] = or_output(p, act, bx, by, bz, ms[q], NULL);
    }
    free(ax); free(ay); free(az);
    free(nfx); free(nfy); free(nfz);
    for (int q = 0; q < 6; q++) free(pc[q]);
}

/* Rotation of one cell's M by `angle` about (nx, ny, nz)
 * (NmrRelaxationSolve.cpp:2536-2550). */
static void or_ort(double* mx, double* my, double* mz, long c, double nx, double ny, double nz, double angle)
{
    const double cs = cos(angle), sn = sin(angle);
    const double v0 = mx[c], v1 = my[c], v2 = mz[c];
    const double scal = nx * v0 + ny * v1 + nz * v2;
    mx[c] = v0 * cs + nx * (1.0 - cs) * scal + sn * (ny * v2 - nz * v1);
    my[c] = v1 * cs + ny * (1.0 - cs) * scal + sn * (nz * v0 - nx * v2);
    mz[c] = v2 * cs + nz * (1.0 - cs) * scal + sn * (nx * v1 - ny * v0);
}

/* Instantaneous pulse on every pore cell of the padded domain, ghosts
 * included (NmrRelaxationSolve.cpp:2552-2608); flags 1-4 as there. */
void or_add_impulse(const or_params* p, const char* act, int flag, double grad, double nx, double ny,
                    double nz, double* mx, double* my, double* mz)
{
    double length = pow(nx * nx + ny * ny + nz * nz, 0.5);
    if (length == 0.0) length = 1.0;
    const double nx0 = nx / length, ny0 = ny / length, nz0 = nz / length;
    const double distance = 10e+12;
    for (int k = 0; k < p->kk + 2; k++)
        for (int j = 0; j < p->jk + 2; j++)
            for (int i = 0; i < p->ik + 2; i++) {
                const long c = IX(p, i, j, k);
                if (act[c] != '1') continue;
                if (flag == 1) {
                    const double angle = grad * p->gamma * ((i - 1) * p->dx * nx + (j - 1) * p->dy * ny + (k - 1) * p->dz * nz);
                    or_ort(mx, my, mz, c, nx0, ny0, nz0, angle);
                } else if (flag == 2) {
                    or_ort(mx, my, mz, c, nx, ny, nz, grad);
                } else if (flag == 3) {
                    const double eff = pow(nx * nx + ny * ny + nz * nz, 0.5);
                    const double t = ((double)i - 1.0 - distance) * p->dx * nx0 + ((double)j - 1.0 - distance) * p->dy * ny0 +
                                     ((double)k - 1.0 - distance) * p->dz * nz0;
                    or_ort(mx, my, mz, c, nx0, ny0, nz0, p->gamma * eff * pow(pow(t, 2), 0.5));
                } else if (flag == 4) {
                    const double angle = 0.0 - p->gamma * sqrt(pow((i - 0.5) * p->dx * nx - 0.5 * p->ik * p->dx * nx, 2.0) +
                                                               pow((j - 0.5) * p->dy * ny - 0.5 * p->jk * p->dy * ny, 2.0) +
                                                               pow((k - 0.5) * p->dz * nz - 0.5 * p->kk * p->dz * nz, 2.0));
                    or_ort(mx, my, mz, c, nx, ny, nz, angle);
                }
            }
}

/* Threads of the OpenMP loops from now on (n > 0), and the count in force. */
int or_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
