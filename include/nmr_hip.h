/*
 * nmr_hip.h -- C ABI of the MI355X (gfx950) Bloch-Torrey solver library
 * libnmr_hip.so.  Plain C: POD structs, raw host pointers, int status codes.
 * Nothing here throws or exits; every entry point returns NMR_OK (0) or a
 * negative status whose text nmr_hip_strerror() gives, with a detail line
 * from nmr_hip_last_error() (per thread).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   nmr_hip_create          NmrRelaxationSolve::InitializeGpu()
 *                           src/NmrRelaxationSolveKernels.cu:12-88 (27 cudaMalloc,
 *                           21 H2D copies); the coefficient arrays it uploads
 *                           (D, T1, T2, mainField, noiseField) are derived
 *                           on the device from the mask and these scalars
 *                           (src/NmrRelaxationSolve.cpp:1756-1989)
 *   nmr_hip_upload_*        the H2D copies of InitializeGpu (cu:63-87)
 *   nmr_hip_run             the per-step loop body of solve_gpu()
 *                           (cu:1187-1227): ExplicitMakUpgrade (cu:1473-1501) or
 *                           ExplicitSimple (src/NmrRelaxationSolve.cpp:3148-3282),
 *                           BoundaryConditionsForVectorField_gpu (cu:1321-1396),
 *                           OutputX/Y/Z (cpp:3301-3461) every step
 *   nmr_hip_download_magnetization  the D2H copies of solve_gpu (cu:1195-1197)
 *   nmr_hip_destroy         cudaDeviceReset at the end of solve_gpu (cu:1316)
 *   status codes            gpuAssert/gpuErrchk (headers/NmrRelaxationSolve.h:12-27),
 *                           which print and exit(); here the caller decides.
 *
 * Threading: one context per host thread; a context binds the HIP device
 * current at create time and owns one HIP stream; calls return after the
 * work they enqueue has completed.  Not thread-safe per context.
 */
#ifndef NMR_HIP_H
#define NMR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define NMR_OK 0
#define NMR_ERR_INVALID (-1)     /* bad dimension / option / size */
#define NMR_ERR_NULL (-2)        /* required pointer is NULL */
#define NMR_ERR_DEVICE (-3)      /* HIP runtime or launch failure */
#define NMR_ERR_NOMEM (-4)       /* device allocation failed */
#define NMR_ERR_STATE (-5)       /* call out of order (e.g. run before upload) */
#define NMR_ERR_UNSUPPORTED (-6) /* option combination not implemented */

/* ---- physics switches (nVectorMult, nVelocity, nDiffusion of the reference) ---- */
#define NMR_F_DIFFUSION 1
#define NMR_F_VELOCITY 2
#define NMR_F_PRECESSION 4
/* Upwind scheme only: advection across a face only when the neighbour is a
 * pore -- the face masks of NMR_2d (NmrRelaxationSolveKernels.cu:190-210), a
 * kernel the reference defines but never launches.  Without it a pore next
 * to a solid face that carries velocity takes the solid's value into its
 * upwind flux, as ExplicitSimple does. */
#define NMR_F_ADVECT_MASK 8

/* ---- host array layouts accepted by upload/download ---- */
/* COMPACT: interior only, x fastest.  mask uint8 0/1 [nz][ny][nx];
 *   magnetization double [nz][ny][nx];
 *   face velocities double Vx [nz][ny][nx+1], Vy [nz][ny+1][nx], Vz [nz+1][ny][nx]
 *   (Vx[k][j][i] is the face between cells i-1 and i).
 * PADDED: the reference's arrays, (nx+2)(ny+2)(nz+2) with one ghost cell per
 *   side, idx = i + (nx+2)*(j + (ny+2)*k); mask as chars '0'/'1';
 *   velocity.x[idx(i,j,k)] is the face between i-1 and i
 *   (src/NmrRelaxationSolve.cpp:1435-1662, 3284-3287). */
#define NMR_LAYOUT_COMPACT 0
#define NMR_LAYOUT_PADDED 1

/* ---- update scheme ---- */
#define NMR_SCHEME_UPWIND 0      /* ExplicitSimple, cpp:3148-3282 */
#define NMR_SCHEME_MACCORMACK 1  /* ExplicitMakUpgrade, cu:637-967, 1473-1501; a
                                    solid neighbour reads as the pore's own value
                                    (the reference's racy in-place ghost fill,
                                    cu:765-782, 931-948, resolved race-free);
                                    solid voxels end each step at 0 */

/* ---- boundary condition per axis (BoundaryConditionsForVectorField_gpu,
 * cu:1321-1396; NB1/NB3 = x, NB2/NB4 = y, NB5/NB6 = z) ---- */
#define NMR_BC_NEUMANN 0         /* ghost = adjacent interior (BC_ZeroDer, NB = 1) */
#define NMR_BC_PERIODIC 1        /* ghost = opposite interior (BC_Inf, NB = 3) */
#define NMR_BC_DIRICHLET 2       /* ghost = 0 (BC_Advanced with Val 0, NB = 2) */

/* ---- gradient precession form ---- */
#define NMR_PHASE_LINEAR 0       /* M += dt*g*(M x Bg), as cpp:3230-3235 */
#define NMR_PHASE_EXACT 1        /* (Mx + iMy) *= exp(-i*dt*g*Bg_z) (ncomp 2 only) */

typedef struct nmr_problem {
    int32_t nx, ny, nz;     /* interior voxels (ik, jk, kk) */
    int32_t ncomp;          /* 1: Mx; 2: Mx,My; 3: Mx,My,Mz */
    int32_t flags;          /* NMR_F_* */
    int32_t scheme;         /* NMR_SCHEME_* */
    int32_t phase_mode;     /* NMR_PHASE_* */
    int32_t buffer[3];      /* signal trim per side (bufferSizeX/Y/Z) */
    int32_t bc[3];          /* NMR_BC_* per axis x, y, z; ghost cells count as
                               pores for diffusion and surface relaxation, as
                               the reference's mask ghosts copy the interior */
    int32_t reserved;
    double dx, dy, dz;      /* voxel size (m) */
    double dt;              /* time step (s) */
    double diffusivity;     /* pore diffusivity D_WAT (m^2/s) */
    double r1_bulk;         /* 1/T1_WAT (1/s) */
    double r2_bulk;         /* 1/T2_WAT (1/s) */
    double rho1, rho2;      /* surface relaxivities R1_WAT, R2_WAT (m/s) */
    double gamma;           /* g (reference: 4.78941714e7) */
    double p_eq;            /* equilibrium magnetization pEq */
    double b0_dir[3];       /* main field (uniform); normalised internally */
    double grad[3];         /* constant gradient (T/m) used when no schedule */
} nmr_problem;

typedef struct nmr_ctx nmr_ctx;

/* Validate the problem, allocate all device memory, bind the current device. */
int nmr_hip_create(const nmr_problem* prob, nmr_ctx** out);
/* Free everything the context owns.  NULL is accepted (no-op). */
int nmr_hip_destroy(nmr_ctx* ctx);

int nmr_hip_upload_mask(nmr_ctx* ctx, const uint8_t* mask, int32_t layout);
/* Any of vx/vy/vz may be NULL (zero velocity on that axis). */
int nmr_hip_upload_velocity(nmr_ctx* ctx, const double* vx, const double* vy,
                            const double* vz, int32_t layout);
/* Components beyond ncomp are ignored and may be NULL. */
int nmr_hip_upload_magnetization(nmr_ctx* ctx, const double* mx, const double* my,
                                 const double* mz, int32_t layout);
/* Components beyond ncomp are not written (may be NULL).  PADDED output
 * fills the ghost layer with the zero-derivative copy the reference's
 * BoundaryConditionsForVectorField leaves there. */
int nmr_hip_download_magnetization(nmr_ctx* ctx, double* mx, double* my, double* mz,
                                   int32_t layout);

/* Per-voxel coefficients for mixtures and face-wise surface relaxivity
 * (FillInDiffusinAndTime, src/NmrRelaxationSolve.cpp:1756-1893): pore
 * diffusivity D (m^2/s) and relaxation rates r1 = 1/T1, r2 = 1/T2 (1/s,
 * surface terms included) of every voxel (r1 may be NULL for ncomp < 3);
 * solid voxels are ignored.  They replace diffusivity, r1/r2_bulk and
 * rho1/rho2 of the problem from then on; diffusion between two pores uses
 * the harmonic mean 2 Da Db / (Da + Db) (cpp:3248-3259).  Linear phase only. */
int nmr_hip_upload_coefficients(nmr_ctx* ctx, const double* D, const double* r1, const double* r2, int32_t layout);

/* Per-voxel field b (T) along the main-field axis (the main field must lie
 * along one axis), in place of the constant-gradient field on that axis:
 * the reference's noise field, noiseField.<axis> = AMPL_NOISE * (rand() %
 * 2000 + 1 - 1000) / 1000 on every cell when AMPL_NOISE and NDIRECT_FIELD
 * are set (FillInMainField, src/NmrRelaxationSolve.cpp:1973-1978), which
 * ExplicitSimple turns into precession (cpp:3218-3235).  NULL removes it.
 * Runs in the general z-sweep; linear phase only. */
int nmr_hip_upload_field(nmr_ctx* ctx, const double* b, int32_t layout);
/* Phase weights of the two-phase outputs (OutputX/Y/Z with NCOMP=2,
 * cpp:3324-3350; w = c * [c1 > 0.8 or c2 > 0.8]).  With weights uploaded
 * every signal row holds 2*ncomp sums: sum(w1*M_q) for q < ncomp, then
 * sum(w2*M_q), over pores of the trimmed interior. */
int nmr_hip_upload_phase_weights(nmr_ctx* ctx, const double* w1, const double* w2, int32_t layout);

/* Advance nsteps time steps.  grad_sched (nullable): [nsteps][3] gradient
 * amplitudes (T/m) in force during each step; NULL uses prob->grad.
 * signal (nullable): [(nsteps+1)][ncomp] pore sums over the trimmed
 * interior; row 0 is the state before the first step, row n the state
 * after n updates. */
int nmr_hip_run(nmr_ctx* ctx, int64_t nsteps, const double* grad_sched, double* signal);

/* nmr_hip_run without device allocation.  No counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same steps,
 * fields and signal rows as nmr_hip_run, but the per-workgroup partial sums
 * and the signal rows live in a fixed 8 MiB device array of the library
 * (one per GPU, held until the process ends) instead of buffers of the
 * context: a run is split into sub-runs whose partial sums and rows fit it.
 * Every other buffer must exist already (the uploads create them; a buffer
 * that a later layout or scheme switch would need gives NMR_ERR_STATE), and
 * a grid whose one step needs more partial sums than the array holds gives
 * NMR_ERR_INVALID.  All contexts on one GPU share the array: calls that
 * overlap on one GPU (two host threads) race. */
int nmr_hip_run_static(nmr_ctx* ctx, int64_t nsteps, const double* grad_sched, double* signal);

/* Instantaneous pulse on every pore voxel, AddImpulse semantics
 * (src/NmrRelaxationSolve.cpp:2552-2608, rotation OrtTransform cpp:2536-2550):
 *   flag 1: gradient pulse, angle = param*g*((i-1)dx*ax + (j-1)dy*ay + (k-1)dz*az)
 *           about the unit vector of (ax, ay, az);
 *   flag 2: RF pulse, angle = param about (ax, ay, az) as given;
 *   flag 3: multi-gradient pulse (angle g*|a|*|(r - 1e13 d).a_hat|);
 *   flag 4: radial phase -g*|(r - r_centre) * a| about (ax, ay, az).
 * Replaces the host-side AddImpulse and its 6 PCIe copies in Sequence_gpu
 * (cpp:2681-3146).  ncomp 3, or ncomp 2 with a rotation axis along z. */
int nmr_hip_pulse(nmr_ctx* ctx, int32_t flag, double param, double ax, double ay, double az);

/* Device selection before nmr_hip_create (main.cpp:32 picks (rank+shift)%NGPU). */
int nmr_hip_device_count(int32_t* count);
int nmr_hip_set_device(int32_t device);

/* Pore sums of the current state: out[ncomp] (out[2*ncomp] with phase
 * weights); the last row of the last run when nothing has changed the state
 * since, else a sum over the state (OutputX/Y/Z, cpp:3301-3461). */
int nmr_hip_signal(nmr_ctx* ctx, double* out);

/* Measurement, no reference counterpart: average GPU time (ms, HIP events)
 * of `reps` device-to-device copies of `bytes` bytes by the library's own
 * 16-byte-per-lane copy kernels, the fastest of its four shapes (grid-stride
 * or one pass, plain or non-temporal) -- the streaming ceiling the sweeps are
 * compared with.  Allocates and frees its two buffers. */
int nmr_hip_copy_bandwidth(int64_t bytes, int32_t reps, double* ms);

/* Measurement, no reference counterpart: dense BF16 matrix-core rate
 * (TFLOP/s, the fastest of `reps` timed launches) of the instruction the
 * learned corrector's convolutions issue, v_mfma_f32_16x16x32_bf16, with
 * every wave running `iters` rounds of 8 independent accumulation chains
 * on register operands (four waves per SIMD) -- the ceiling the
 * convolutions' MFMA work is compared with.  Allocates and frees a small
 * output buffer.  iters in [1, 2^24]. */
int nmr_hip_mfma_rate(int64_t iters, int32_t reps, double* tflops);

/* GPU time (ms) between HIP events recorded on the context's stream around
 * the kernels of the last nmr_hip_run, and the number of sweep launches. */
int nmr_hip_last_run_ms(nmr_ctx* ctx, double* ms, int64_t* sweeps);

/* Device layout of the sweeps (no reference counterpart): 1 (the default)
 * keeps the state, the velocity (as the per-voxel face-velocity sums the
 * sweeps read) and the per-voxel coefficients, field and phase weights in
 * bricks of 128 bytes (8 x-voxels by 4 rows of one plane)
 * -- for the one-step sweep with any one-step tile variant and for the
 * general sweep -- so that all-solid bricks are never fetched.  The uploads
 * convert their arrays once; the state stays in bricks from run to run
 * (pulses rotate it there) and is converted back only for a download or a
 * learned solve.  0 always sweeps the row layout.  The fields are the same
 * bits either way. */
int nmr_hip_set_layout(nmr_ctx* ctx, int32_t brick);
/* Tile shape of the sweep kernel: variant index (-1 = automatic, -2 = a
 * one-thread-per-voxel kernel with the arithmetic of the sweeps, a
 * cross-check) and z-chunk length (0 = automatic: the length in 6..64
 * that best fills whole rounds of the workgroups the device holds at
 * once, 8 on ties).  nmr_hip_num_variants() gives the variant count;
 * nmr_hip_variant_info() describes one: lanes per row, rows per thread, tile
 * width and height, and time steps per pass (2 = temporally blocked sweep,
 * usable when nx <= tile_x).  Every variant gives bit-identical fields. */
int nmr_hip_set_tile(nmr_ctx* ctx, int32_t variant, int32_t lz);
int nmr_hip_num_variants(void);
int nmr_hip_variant_info(int32_t variant, int32_t* lanes_x, int32_t* rows_per_thread,
                         int32_t* tile_x, int32_t* tile_y, int32_t* steps_per_pass);
/* Bytes of device memory the context holds. */
int nmr_hip_device_bytes(nmr_ctx* ctx, int64_t* bytes);

/* ---- batched solves (many uneven instances per launch) ----
 * No reference counterpart: the reference solves one domain per process.
 * All instances share the physics of `common` (ncomp, flags, phase mode,
 * dx/dy/dz, dt, D, bulk relaxation, g, pEq, B0 direction, signal trim);
 * each has its own grid, surface relaxivities and pulsed-gradient spin
 * echo along z (+G on steps [0, nd), -G on [nD, nD+nd), the second lobe
 * taking precedence where they overlap; nd = lround(delta/dt),
 * nD = lround(big_delta/dt)).  One persistent launch advances every
 * instance by one step; chunks of steps are captured in a HIP graph and
 * replayed.  Arrays are in the COMPACT layout. */
typedef struct nmr_batch_instance {
    int32_t nx, ny, nz, reserved;
    double rho1, rho2;      /* surface relaxivities (m/s) */
    double grad_amp;        /* pulsed gradient amplitude along z (T/m) */
    double delta, big_delta;/* lobe length and lobe spacing (s) */
} nmr_batch_instance;

typedef struct nmr_batch nmr_batch;

int nmr_hip_batch_create(const nmr_problem* common, int32_t ninst, const nmr_batch_instance* inst,
                         nmr_batch** out);
int nmr_hip_batch_destroy(nmr_batch* b);
/* Mask (0/1), face velocities (nullable when the flags have no velocity)
 * and magnetization (components beyond ncomp may be NULL) of instance i. */
int nmr_hip_batch_upload(nmr_batch* b, int32_t i, const uint8_t* mask, const double* vx, const double* vy,
                         const double* vz, const double* mx, const double* my, const double* mz);
int nmr_hip_batch_download(nmr_batch* b, int32_t i, double* mx, double* my, double* mz);
/* Device layout of the batch (no reference counterpart): 1 (the default)
 * keeps every instance's state and per-voxel face-velocity sums in bricks
 * of 8 x-voxels by 4 rows from run to run, as nmr_hip_set_layout does for one domain; 0 sweeps
 * the row layout.  The fields are the same bits either way. */
int nmr_hip_batch_set_layout(nmr_batch* b, int32_t brick);
/* signal (nullable): [ninst][nsteps+1][ncomp]; step counting continues
 * across calls (the gradient schedule sees steps 0, 1, ... since create). */
int nmr_hip_batch_run(nmr_batch* b, int64_t nsteps, double* signal);
/* variant (-1 automatic / autotuned; -3 each instance its own one-step
 * tile shape, the least padded one, with one launch per shape), z-chunk
 * (0 automatic), steps per captured graph (0: direct launches, no graph;
 * must be even), grid (1: as many workgroups as fit on the device at once,
 * each looping over tiles; 0: one workgroup per tile). */
int nmr_hip_batch_set_tile(nmr_batch* b, int32_t variant, int32_t lz, int32_t graph_steps, int32_t persistent);
/* Times every one-step tile variant (and -3), z-chunk and grid kind on this batch
 * (repeated single steps from the current state, which stays unchanged),
 * times the four fastest again over three times the steps on the warm
 * device, and keeps the fastest of those; the choice is cached per set of
 * instance shapes.  Any output pointer may be NULL. */
int nmr_hip_batch_autotune(nmr_batch* b, int32_t trial_steps, int32_t* variant, int32_t* lz,
                           int32_t* persistent);
/* GPU ms of the last run's steps and the voxel-steps they advanced. */
int nmr_hip_batch_last_run_ms(nmr_batch* b, double* ms, int64_t* voxel_steps);

/* ---- learned corrector ----
 * No reference counterpart.  A 3-layer 3-D convolution net (6 -> 16 -> 16 -> 2
 * channels, 3x3x3 kernels, zero padding, ReLU; BF16 weights and activations,
 * FP32 accumulation on the matrix cores) that corrects one explicit step of
 * factor*dt toward the result of factor plain steps.  Inputs per voxel:
 * (Mx, My) of the state before the step times the pore mask, the pore mask,
 * and the cell-centred velocity / vscale.  Output: the correction of
 * (Mx, My) in units of out_scale, applied at pores.
 * Parameters (FP32, NMR_CORRECTOR_PARAMS of them) in the order of
 * conv3d weights: W1[16][6][3][3][3], b1[16], W2[16][16][3][3][3], b2[16],
 * W3[2][16][3][3][3], b3[2] (kernel index kz, ky, kx; x fastest). */
#define NMR_CORRECTOR_PARAMS 10402
typedef struct nmr_corrector nmr_corrector;

typedef struct nmr_train_opts {
    int32_t factor;         /* plain steps per coarse step */
    int32_t rounds;         /* sampling rounds: one sample per instance each */
    int32_t stride;         /* plain steps of the batch before each round */
    int32_t reserved;
    double lr;              /* Adam step size */
} nmr_train_opts;

/* He-normal initial weights from seed (deterministic), zero biases. */
int nmr_hip_corrector_create(uint64_t seed, nmr_corrector** out);
int nmr_hip_corrector_destroy(nmr_corrector* net);
int nmr_hip_corrector_num_params(void);
int nmr_hip_corrector_get_params(nmr_corrector* net, float* params);
/* Also resets the Adam moments. */
int nmr_hip_corrector_set_params(nmr_corrector* net, const float* params);
/* vscale: velocity unit of the input (m/s, default 1e-3); out_scale: unit of
 * the output correction (0: set by the first training sample to the RMS of
 * its targets). */
int nmr_hip_corrector_set_scales(nmr_corrector* net, double vscale, double out_scale);
int nmr_hip_corrector_get_scales(nmr_corrector* net, double* vscale, double* out_scale);
/* Convolution kernel of the corrector (all give the same bits):
 * NMR_CONV_ZMARCH (default) marches columns of 16 x 16 output rows along z
 * through a ring of LDS planes, zc planes per column (0: 16);
 * NMR_CONV_TILES stages 16 x 8 x 4-voxel tiles; NMR_CONV_ROWS gathers each
 * row's operands from memory; NMR_CONV_ZMARCH_SPARSE is NMR_CONV_ZMARCH with
 * the learned solve's last layer computed only on rows that hold a pore
 * (as NMR_CONV_TILES and NMR_CONV_ROWS do). */
#define NMR_CONV_ZMARCH 0
#define NMR_CONV_TILES 1
#define NMR_CONV_ROWS 2
#define NMR_CONV_ZMARCH_SPARSE 3
int nmr_hip_corrector_set_conv(nmr_corrector* net, int32_t kind, int32_t zc);
/* The net on host data: features [nz][ny][nx][6] -> out [nz][ny][nx][2]. */
int nmr_hip_corrector_forward(nmr_corrector* net, int32_t nx, int32_t ny, int32_t nz, const float* features,
                              float* out);
/* Gradient of sum(grad_out * out) with respect to the parameters. */
int nmr_hip_corrector_backward(nmr_corrector* net, int32_t nx, int32_t ny, int32_t nz, const float* features,
                               const float* grad_out, float* grad_params);
/* Trains on samples the batch generates (ncomp 2, exact phase): each round
 * advances the batch `stride` steps, keeps every instance's state, advances
 * `factor` steps, and takes one Adam step per instance on the mean squared
 * error over its pores between the net's output and (plain result - coarse
 * step) / out_scale.  losses (nullable): [rounds][ninst], before each step. */
int nmr_hip_corrector_train(nmr_corrector* net, nmr_batch* b, const nmr_train_opts* opts, double* losses);
/* ncoarse coarse steps of factor*dt with the corrector on a context (ncomp
 * 2, exact phase, upwind, zero-derivative boundaries).  A step is accepted
 * when nothing is non-finite, the pore energy sum(Mx^2 + My^2) has not
 * grown by more than a factor (1 + tol), and no correction exceeds tol
 * times the largest |M| of the previous state; otherwise it is redone as
 * factor plain steps.  grad_sched (nullable): [ncoarse*factor][3] z gradients per plain
 * step.  signal (nullable): [ncoarse+1][2] trimmed pore sums after each
 * coarse step.  fallbacks (nullable): [4] = coarse steps redone, and of
 * those the ones that failed on a non-finite value, on energy, on size. */
int nmr_hip_run_learned(nmr_ctx* ctx, nmr_corrector* net, int64_t ncoarse, int32_t factor, const double* grad_sched,
                        double tol, double* signal, int64_t* fallbacks);

const char* nmr_hip_strerror(int code);
const char* nmr_hip_last_error(void);
int nmr_hip_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NMR_HIP_H */
