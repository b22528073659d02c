// nmr_kernels.h -- internal interface between the C ABI (nmr_abi.hip) and
// the gfx950 kernels (nmr_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmr {

// Per-step coefficients of the update, all pre-multiplied by dt (FP32).
struct Coef {
    float cdx, cdy, cdz;    // dt*D/d^2 per axis (0 when diffusion is off)
    float hx, hy, hz;       // dt/d per axis
    float r1, r2;           // dt/T1_bulk, dt/T2_bulk
    float s1x, s1y, s1z;    // dt*rho1/d per solid face
    float s2x, s2y, s2z;    // dt*rho2/d per solid face
    float peq;              // equilibrium magnetization
    float n0, n1, n2;       // unit main-field direction
    float bgx, bgy, bgz;    // dt*g*G_a*d_a: precession angle per index unit
    float dtg;              // dt*g (0 without precession): angle per tesla of a per-voxel field
};

struct SweepArgs {
    const float* min[3];
    float* mout[3];
    const uint8_t* mask;
    const float* adv[3];    // upwind velocity sums per axis (adv_kernel), laid out as the state
    double* partials;       // [nblocks][nc] signal partials of this step
    int nx, ny, nz;
    int pitch, pitch1;
    int ntx, nty, ntz, lz, nblocks;
    int bx0, bx1, by0, by1, bz0, bz1; // signal window [b0, b1) per axis
    int inert;              // no pore-solid face carries velocity
    int nby;                // brick layout: brick rows per plane
    Coef c;
};

// One instance of a batch (device memory).
struct BatchInst {
    float* m[2][3];         // ping-pong magnetization buffers (rows, or bricks with the velocity sums in bricks too)
    const uint8_t* mask;
    const float* adv[3];    // upwind velocity sums, laid out as the state
    int nx, ny, nz, pitch, pitch1;
    int ntx, nty, ntz, lz;
    int bx0, bx1, by0, by1, bz0, bz1;
    int inert;
    int nby;                // brick layout: brick rows per plane
    int nd, nD;             // PGSE lobe length and start of the second lobe (steps)
    float bgz;              // dt*g*G*dz of the +G lobe, rounded once on the host
                            // (the -G lobe negates it, which is exact)
    Coef c;                 // coefficients without the gradient
};

struct BatchArgs {
    const BatchInst* inst;
    const int2* tiles;      // (instance, tile index within the instance)
    int ntiles;
    int parity;             // which ping-pong buffer holds the current state
    int local_step;         // step within the captured chunk
    const int* step_base;   // device step counter (first step of the chunk)
    double* partials;       // [ntiles][nc] of this step
};

// Per-voxel sums of the two face velocities of each axis from the face
// velocities (row layout: vx [nz][ny][pitch1], vy [nz][ny+1][pitch], vz
// [nz+1][ny][pitch]), written as the state's rows [nz][ny][pitch]: the
// upwind sums of the plain sweep (0 where the inflow face carries no
// velocity) or, with centred, the plain sums of the MacCormack stages.
struct AdvArgs {
    const float* vx;
    const float* vy;
    const float* vz;
    float* out[3];
    int nx, ny, nz, pitch, pitch1;
    int centred;
};

struct InertArgs {
    const uint8_t* mask;
    const float* vx;
    const float* vy;
    const float* vz;
    int* flag;
    int nx, ny, nz, pitch, pitch1;
};

// Two-step (temporally blocked) sweep: one x tile (nx <= 4*LX), TY = 512/LX - 2
// rows per tile, z-chunks of lz planes.
struct Sweep2Args {
    const float* min[3];
    float* mout[3];
    const uint8_t* mask;
    const float* adv[3];    // upwind velocity sums (row layout)
    double* partials1;      // [nblocks][nc] signal after the first step
    double* partials2;      // [nblocks][nc] signal after the second step
    int nx, ny, nz;
    int pitch, pitch1;
    int nty, ntz, lz, nblocks;
    int bx0, bx1, by0, by1, bz0, bz1;
    Coef c1, c2;            // coefficients of the two steps
};

// Per-voxel step for the schemes and boundary conditions the z-sweep does
// not cover (MacCormack, periodic and zero-value ghosts).
enum { GEN_UPWIND = 0, GEN_PREDICT = 1, GEN_CORRECT = 2 };
enum { BC_NEUMANN = 0, BC_PERIODIC = 1, BC_DIRICHLET = 2 };
struct GenArgs {
    const float* m[3];      // state at the start of the step
    const float* p[3];      // predictor (corrector stage)
    float* out[3];          // result of the stage
    const uint8_t* mask;
    const float* vx;
    const float* vy;
    const float* vz;
    int nx, ny, nz, pitch, pitch1;
    int bc[3];              // BC_* per axis
    int vel;                // face velocities present
    int amask;              // upwind advection only across faces to pores (NMR_F_ADVECT_MASK)
    const float* dn;        // per-voxel diffusivity / reference (null: uniform, from the mask)
    const float* r1v;       // per-voxel dt/T1 with surface terms (null: from the mask)
    const float* r2v;       // per-voxel dt/T2 with surface terms
    const float* bn;        // per-voxel field (T) along axis bnax, in place of that axis's gradient field; or null
    int bnax;
    Coef c;
};

// General z-sweep: the sweep's tile march (planes in registers, plane k
// with its halo in LDS) for what the plain sweep does not cover -- the
// MacCormack stages, periodic and zero-value ghosts, per-voxel
// coefficients -- with the per-voxel kernel's arithmetic, so that the two
// agree bit for bit.
struct GSweepArgs {
    const float* f[3];      // field the stage differentiates (the state, or the predictor)
    const float* mavg[3];   // corrector stage: the state at the start of the step
    float* out[3];          // result of the stage
    const uint8_t* mask;
    const float* dn;        // per-voxel diffusivity / reference (coefficient variant)
    const float* r1v;       // per-voxel dt/T1
    const float* r2v;       // per-voxel dt/T2
    const float* bn;        // per-voxel field (T) along axis bnax, in place of that axis's gradient field; or null
    int bnax;
    double* partials;       // [nblocks][nc] pore sums of the result, or null
    const float* w[2];      // brick sweep: phase weights (row layout) of fused weighted sums,
                            // partials then [nblocks][2*nc]; or null
    const float* cs[3];     // face-velocity sums per voxel, laid out as the state: centred
                            // (MacCormack stages) or upwind (the upwind stage), AdvArgs
    int nx, ny, nz, pitch, pitch1;
    int ntx, nty, ntz, lz, nblocks;
    int bx0, bx1, by0, by1, bz0, bz1; // signal window [b0, b1) per axis
    int bc[3];              // BC_* per axis
    int skip_m;             // all-solid chunks are never read by a pore: not loaded
    int amask;              // upwind advection only across faces to pores (NMR_F_ADVECT_MASK)
    int nby;                // brick layout (lanes 16): brick rows per plane
    Coef c;
};

struct PulseArgs {
    float* m[3];            // row layout, or bricks when nb8 > 0
    const uint8_t* mask;
    int nx, ny, nz, pitch, nc;
    int nby, nb8;           // brick layout: brick rows per plane, bricks per brick row (0: rows)
    int flag;
    double param, ax, ay, az;
    double dx, dy, dz, gamma;
};

struct FieldSumArgs {
    const float* f[3];      // row layout, or bricks when nb8 > 0
    const uint8_t* mask;
    const float* w[2];      // phase weights (null: plain sums)
    double* part;           // [field_sum_blocks][nc] or, with weights, [..][2*nc]
    int nx, ny, nz, pitch;
    int bx0, bx1, by0, by1, bz0, bz1;
    int nby, nb8;           // brick layout of f (0: rows)
};

// ---- learned corrector (nmr_learn_kernels.hip) ----
// Tensors: BF16 channels-last on the grid padded by one zero voxel per side.
constexpr int kWgradPart = 27 * 256 + 16; // floats per wavefront partial of wgrad

struct ConvArgs {
    const uint16_t* in;     // [(nz+2)(ny+2)(nx+2)][cin]
    const uint16_t* wa;     // packed A fragments [chunks][64][8]
    const float* bias;      // [16] (forward modes)
    const uint16_t* ref;    // gate tensor [..][16] (input-gradient mode)
    uint16_t* out;          // [..][16]
    float* outf;            // [..][2] (linear mode)
    const int* rows;        // row list, row = (z*ny + y)*nseg + seg; null = all rows
    int nrows;
    int nx, ny, nz, nseg;   // nseg = ceil(nx/16)
    int zc;                 // output planes per column of the z-marching kernel (0 = 16)
    const float* smx;       // mode 4: input channels 0, 1 from this state (Mx, My times the pore
    const float* smy;       //   channel), compact with pitch spitch
    int spitch;
};

struct WgradArgs {
    const uint16_t* in;     // layer input [..][cin]
    const uint16_t* gout;   // output gradient [..][16]
    float* part;            // [waves][kWgradPart]
    int nx, ny, nz;
};

struct FeatureArgs {
    const float* mx;
    const float* my;        // nullable
    const uint8_t* mask;
    const float* vx;        // nullable (no velocity)
    const float* vy;
    const float* vz;
    uint16_t* out;          // [..][8]
    int nx, ny, nz, pitch, pitch1;
    float vinv;             // 1 / velocity scale
};

struct LossArgs {
    const float* pred;      // [..][2]
    const float* tx;        // target, compact with pitch
    const float* ty;
    const uint8_t* mask;
    uint16_t* gout;         // [..][16]
    double* lpart;          // [blocks]
    int nx, ny, nz, pitch;
    float tinv;             // 1 / output scale
    float winv;             // 1 / number of pores
};

struct ApplyArgs {
    float* mx;
    float* my;
    const uint8_t* mask;
    const float* pred;      // [..][2] or null (statistics only)
    double* part;           // [blocks][6]
    int nx, ny, nz, pitch;
    int bx0, bx1, by0, by1, bz0, bz1;
    float scale;
};

hipError_t launch_conv(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s);
// tiles of 16 x 8 x 4 output voxels of the LDS-staged convolution (all rows)
int conv_tiles(int nx, int ny, int nz);
hipError_t launch_conv_tiled(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s);
// columns of 16 x 16 output rows by zc planes of the z-marching convolution
int conv_columns(int nx, int ny, int nz, int zc);
hipError_t launch_conv_zm(int cin, int mode, const ConvArgs& a, int grid, hipStream_t s);
hipError_t launch_wgrad(int cin, const WgradArgs& a, int grid, hipStream_t s);
hipError_t launch_wreduce(const float* part, int nwaves, int cin_real, int cout_real, float* grad, hipStream_t s);
hipError_t launch_pack(const float* w, int cin_real, int cout_real, int cin_store, int backward, uint16_t* frag,
                       hipStream_t s);
hipError_t launch_features(const FeatureArgs& a, hipStream_t s);
hipError_t launch_feature_copy(const float* f6, int nx, int ny, int nz, uint16_t* out, hipStream_t s);
hipError_t launch_gout_copy(const float* g2, int nx, int ny, int nz, uint16_t* out, hipStream_t s);
int loss_blocks(int nx, int ny, int nz);
hipError_t launch_loss(const LossArgs& a, hipStream_t s);
hipError_t launch_adam(float* p, float* m, float* v, const float* g, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s);
hipError_t launch_diff(const float* ax, const float* ay, const float* bx, const float* by, float* ox, float* oy,
                       long n, hipStream_t s);
int apply_blocks(int nx, int ny, int nz);
hipError_t launch_apply(const ApplyArgs& a, hipStream_t s);
hipError_t launch_stats_reduce(const double* part, int nblocks, double* out, hipStream_t s);
hipError_t launch_row_flags(const uint8_t* mask, int nx, int ny, int nz, int pitch, uint8_t* flag, hipStream_t s);
// grid workgroups of 256 threads, each wave `iters` x 8 MFMA 16x16x32 BF16; out: grid*256 floats
hipError_t launch_mfma_rate(int iters, int grid, float* out, hipStream_t s);

int num_variants();
// the compacted brick sweep's variant index (the last one)
int compact_variant();
// steps: time steps per pass (1, or 2 for the temporally blocked sweep)
bool variant_shape(int v, int* lx, int* rpt, int* tx, int* ty, int* steps);
// keep: the state's loads and stores go without the non-temporal hint (the
// brick 64 x 16 tile, variant 4; other tiles ignore it)
hipError_t launch_sweep(int nc, bool vel, bool exact, int variant, const SweepArgs& a, hipStream_t s,
                        bool brick = false, bool keep = false);
// whether tile variant v has a brick-layout sweep, and the automatic one
bool brick_variant(int v);
int brick_auto_variant(int nx);
// floats of a field of nc4 4-voxel chunks per row, Y rows, Z planes in
// bricks of 8 x-voxels by 4 rows
size_t brick_floats(int nc4, int Y, int Z);
// row layout [Z][Y][4*nc4] <-> bricks (to_brick: row -> brick)
hipError_t launch_brick(const float* src, float* dst, int nc4, int Y, int Z, bool to_brick, hipStream_t s);
hipError_t launch_sweep2(int nc, bool vel, bool exact, int lx, const Sweep2Args& a, hipStream_t s);
hipError_t launch_batch(int nc, bool vel, bool exact, int variant, int grid, const BatchArgs& b, hipStream_t s,
                        bool brick = false);
int batch_resident_blocks(int nc, bool vel, bool exact, int variant, bool brick = false);
// Workgroups of the one-step sweep (variant 0..6, row or brick layout) and of
// the general sweep's upwind kernel (upwind) or MacCormack corrector (on
// bricks with fused phase-weighted sums: its weighted instantiation)
// resident on the device at once (CUs x occupancy); -1 if the query fails.
int sweep_resident_blocks(int nc, bool vel, bool exact, int variant, bool brick, bool keep = false);
int gsweep_resident_blocks(int nc, bool vel, bool coef, int lanes, bool brick, bool weighted, bool upwind);
hipError_t launch_batch_reduce(const double* part, int ntiles, int nc, const int* off, const int* ids,
                               const int* step_base, int row0, int64_t rows, double* signal, int nsteps, int ninst,
                               hipStream_t s);
hipError_t launch_step_advance(int* step_base, int n, hipStream_t s);
hipError_t launch_reduce_partials(const double* part, int nsteps, int nwg, int nc, double* out, hipStream_t s);
hipError_t launch_inert_check(const InertArgs& a, hipStream_t s);
hipError_t launch_adv(const AdvArgs& a, hipStream_t s);
hipError_t launch_pulse(const PulseArgs& a, hipStream_t s);
hipError_t launch_general(int nc, int stage, const GenArgs& a, hipStream_t s);
// lanes per tile row of the general sweep for a grid of width nx (tile 4*lanes wide)
int gsweep_lanes(int nx);
// keep: as for launch_sweep (the brick sweeps)
hipError_t launch_gsweep(int nc, bool vel, int stage, bool coef, int lanes, const GSweepArgs& a, hipStream_t s,
                         bool brick = false, bool keep = false);
// m[q][i] = 0 where mask[i] == 0, i < n, for q < nc
hipError_t launch_zero_solids(float* const* m, int nc, const uint8_t* mask, size_t n, hipStream_t s);
// the same for fields in bricks (nz planes of ny rows of pitch voxels)
// dst = src over nchunks 16-byte chunks (16-byte aligned) by copy kernel
// shape `kind` (0 .. copy_kinds() - 1) on a device of `cus` compute units
int copy_kinds();
hipError_t launch_copy(int kind, const float* src, float* dst, size_t nchunks, int cus, hipStream_t s);
hipError_t launch_zero_solids_brick(float* const* m, int nc, const uint8_t* mask, int ny, int nz, int pitch,
                                    hipStream_t s);
// workgroups (= partials) of a field sum over the signal window
int field_sum_blocks(const FieldSumArgs& a);
hipError_t launch_field_sum(int nc, const FieldSumArgs& a, hipStream_t s);

} // namespace nmr
