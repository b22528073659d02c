// nmr_abi.hip -- host side of the C ABI declared in include/nmr_hip.h.
//
// Owns the device state of one solve: ping-pong FP32 magnetization, the
// uint8 pore mask, FP32 staggered face velocities, the per-step signal
// partials and the signal curve.  Replaces the reference's InitializeGpu
// (src/NmrRelaxationSolveKernels.cu:12-88) and the per-step loop body of
// solve_gpu (cu:1187-1227); every HIP call is checked and turned into a
// status code instead of gpuAssert's exit() (headers/NmrRelaxationSolve.h:12-27).

#include "nmr_internal.h"

#include <map>
#include <mutex>

namespace nmri {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

} // namespace nmri

namespace {

using nmri::fail;
using nmri::g_last_error;
using nmri::round_up4;

constexpr int kChunkSteps = 512; // steps whose partials are held before a reduction
constexpr size_t kPinnedSignal = (size_t)1 << 17; // doubles of pinned host staging for signal rows (1 MiB)

// The single-domain sweeps load and store the state without the
// non-temporal hint while its two buffers take at most this many bytes, so
// that the device's memory-side cache keeps what one pass writes for the
// next.  Measured per plain / MacCormack step with and without the hint:
// 128^3 16.3 / 15.7 us; 256^3 65.9 / 60.8 us (phi 0.35: 79.8 / 68.3),
// MacCormack 135.8 / 123.5 us; 512x512x130 (545 MB) 125.1 / 125.2 us; 384^3
// 221 / 257 us (MacCormack 479 / 579); 512x512x1024 1047 / 1234 us
// (MacCormack 2260 / 2683).  The batch keeps the hint.
constexpr size_t kKeepStateBytes = (size_t)512 << 20;
constexpr size_t kBatchStage = (size_t)1 << 20;   // doubles of pinned staging for a batch run's signal (8 MiB)

} // namespace



// Fixed scratch of nmr_hip_run_static: the per-workgroup partial sums and
// the signal rows of a run in a file-scope device array (kScratchDoubles
// doubles, 8 MiB on every device the library runs on, held until the process
// ends), for callers that cannot allocate device memory.
namespace {
constexpr size_t kScratchDoubles = (size_t)1 << 20;
}
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr; // device -> address of its copy of scratch

// Address of `device`'s copy of the scratch array, with that device current;
// taken once per device (by the first context created there) and kept.
int scratch_address(int device, double** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        double* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        it = g_scratch_ptr.emplace(device, ptr).first;
    }
    *out = it->second;
    return NMR_OK;
}

int check_layout(int32_t layout)
{
    if (layout != NMR_LAYOUT_COMPACT && layout != NMR_LAYOUT_PADDED)
        return fail(NMR_ERR_INVALID, "unknown layout %d", (int)layout);
    return NMR_OK;
}

template <typename T>
int dev_alloc(nmr_ctx* c, T** p, size_t count)
{
    if (c->no_alloc) return fail(NMR_ERR_STATE, "a device buffer is missing and this entry point allocates nothing");
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_TRY(hipMalloc((void**)p, bytes));
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, c->stream));
    c->device_bytes += (int64_t)bytes;
    return NMR_OK;
}

void free_all(nmr_ctx* c)
{
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 3; q++)
            if (c->m[b][q]) (void)hipFree(c->m[b][q]);
    for (int q = 0; q < 3; q++)
        if (c->pred[q]) (void)hipFree(c->pred[q]);
    for (float* q : {c->dn, c->r1v, c->r2v, c->bn, c->w[0], c->w[1]})
        if (q) (void)hipFree(q);
    for (int q = 0; q < 3; q++)
        for (float* f : {c->adv[q], c->advb[q], c->csum[q], c->csumb[q]})
            if (f) (void)hipFree(f);
    for (float* q : c->pvb)
        if (q) (void)hipFree(q);
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 3; q++)
            if (c->mb[b][q]) (void)hipFree(c->mb[b][q]);
    for (int q = 0; q < 3; q++)
        if (c->pb[q]) (void)hipFree(c->pb[q]);
    if (c->mask) (void)hipFree(c->mask);
    if (c->vx) (void)hipFree(c->vx);
    if (c->vy) (void)hipFree(c->vy);
    if (c->vz) (void)hipFree(c->vz);
    if (c->partials) (void)hipFree(c->partials);
    if (c->dsignal) (void)hipFree(c->dsignal);
    if (c->hsignal) (void)hipHostFree(c->hsignal);
    if (c->dflag) (void)hipFree(c->dflag);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Padded (reference) index of interior-or-ghost cell (i, j, k), each 0..n+1.
inline size_t pidx(const nmr_problem& p, int i, int j, int k)
{
    return (size_t)i + (size_t)(p.nx + 2) * ((size_t)j + (size_t)(p.ny + 2) * (size_t)k);
}

// Tile shape picked when the caller leaves it automatic.
struct Launch {
    int variant, lx, rpt, tx, ty, steps, lz;
    int ntx, nty, ntz, nwg;
};

void fill_launch(const nmr_problem& p, Launch* L)
{
    nmr::variant_shape(L->variant, &L->lx, &L->rpt, &L->tx, &L->ty, &L->steps);
    L->lz = std::min(std::max(L->lz, 1), p.nz);
    L->ntx = (p.nx + L->tx - 1) / L->tx;
    L->nty = (p.ny + L->ty - 1) / L->ty;
    L->ntz = (p.nz + L->lz - 1) / L->lz;
    const long n = (long)L->ntx * L->nty * L->ntz;
    L->nwg = n > (1L << 30) ? -1 : (int)n;
}

// z-chunk length (planes per workgroup) when the caller leaves it
// automatic.  The workgroups of a sweep run in rounds of `resident` (CUs x
// workgroups per CU), a round lasting about lz + 2 plane-times (the chunk and
// its two halo planes), so a step takes about ceil(nwg / resident) * (lz + 2)
// of them: the lz in 6..64 that minimises this, 8 on ties, else the
// shortest.  At 256^3 (64 brick tiles per chunk layer, 768 resident) it
// picks lz 22, one full round: 70.9 us per step against 73.0 at lz 11 (two
// rounds), 76.8 at lz 8 (2 2/3) and 85.9 at lz 16 (1 1/3); on 512x512x130
// lz 44 / 22 / 11 / 33 (1, 2, 4, 1 1/3 rounds) measured 138.7 / 144.5 /
// 148.1 / 168.8 us, the model's order.
int round_lz(long tiles_xy, int nz, long resident)
{
    if (resident <= 0 || tiles_xy <= 0 || nz <= 0) return std::max(1, std::min(8, nz));
    auto cost = [&](int lz) {
        const long ntz = (nz + lz - 1) / lz;
        return ((ntz * tiles_xy + resident - 1) / resident) * (long)(lz + 2);
    };
    int best = 8;
    long bc = cost(8);
    for (int lz = 6; lz <= 64; lz++) {
        const long c = cost(lz);
        if (c < bc) {
            bc = c;
            best = lz;
        }
    }
    return std::min(best, nz);
}

bool keep_state(const nmr_ctx* c);

void auto_launch(const nmr_ctx* c, int variant, int lz, Launch* L, bool brick = false)
{
    const nmr_problem& p = c->prob;
    const int n1 = 7; // first two-step variant
    (void)n1;
    if (variant < 0) { // measured fastest one-step tiles (two-step variants are opt-in)
        if (p.nx > 128) variant = 0;      // TX 256, TY 4
        else if (p.nx > 64) variant = 2;  // TX 128, TY 8
        else if (p.nx > 32) variant = 4;  // TX 64, TY 16
        else variant = 6;                 // TX 32, TY 32
    }
    L->variant = variant;
    L->lz = 1;
    fill_launch(p, L);
    if (lz <= 0) {
        // one-step: the z-chunk that best fills whole rounds of the
        // workgroups the device holds at once (of the kernel the run
        // launches: with the state kept cached, 4 per CU instead of 3);
        // two-step (one 8-wave workgroup per CU): >= 512 workgroups
        if (L->steps == 2) {
            const long tiles_xy = (long)L->ntx * L->nty;
            const long ntz = std::max<long>(1, (512 + tiles_xy - 1) / tiles_xy);
            lz = (int)std::max<long>(4, (p.nz + ntz - 1) / ntz);
        } else {
            const int res = nmr::sweep_resident_blocks(c->nc, c->vel, c->exact, variant, brick,
                                                       brick && keep_state(c));
            lz = round_lz((long)L->ntx * L->nty, p.nz, res);
        }
    }
    L->lz = lz;
    fill_launch(p, L);
}

nmr::Coef make_coef(const nmr_ctx* c, const double g[3])
{
    const nmr_problem& p = c->prob;
    nmr::Coef k{};
    const double dt = p.dt;
    const double D = (p.flags & NMR_F_DIFFUSION) ? (c->hetero ? c->d_ref : p.diffusivity) : 0.0;
    k.cdx = (float)(dt * D / (p.dx * p.dx));
    k.cdy = (float)(dt * D / (p.dy * p.dy));
    k.cdz = (float)(dt * D / (p.dz * p.dz));
    k.hx = (float)(dt / p.dx);
    k.hy = (float)(dt / p.dy);
    k.hz = (float)(dt / p.dz);
    k.r1 = (float)(dt * p.r1_bulk);
    k.r2 = (float)(dt * p.r2_bulk);
    k.s1x = (float)(dt * p.rho1 / p.dx);
    k.s1y = (float)(dt * p.rho1 / p.dy);
    k.s1z = (float)(dt * p.rho1 / p.dz);
    k.s2x = (float)(dt * p.rho2 / p.dx);
    k.s2y = (float)(dt * p.rho2 / p.dy);
    k.s2z = (float)(dt * p.rho2 / p.dz);
    k.peq = (float)p.p_eq;
    const double nb = std::sqrt(p.b0_dir[0] * p.b0_dir[0] + p.b0_dir[1] * p.b0_dir[1] + p.b0_dir[2] * p.b0_dir[2]);
    k.n0 = (float)(p.b0_dir[0] / nb);
    k.n1 = (float)(p.b0_dir[1] / nb);
    k.n2 = (float)(p.b0_dir[2] / nb);
    if (p.flags & NMR_F_PRECESSION) {
        const double b = dt * p.gamma;
        k.bgx = (float)(b * g[0] * p.dx);
        k.bgy = (float)(b * g[1] * p.dy);
        k.bgz = (float)(b * g[2] * p.dz);
        k.dtg = (float)b;
    }
    return k;
}

void signal_window(const nmr_problem& p, int w[6])
{
    w[0] = p.buffer[0];
    w[1] = p.nx - p.buffer[0];
    w[2] = p.buffer[1];
    w[3] = p.ny - p.buffer[1];
    w[4] = p.buffer[2];
    w[5] = p.nz - p.buffer[2];
}

// The partial and signal buffers grow to at least twice their capacity, so
// that runs of growing length reallocate rarely; uploads pre-size them for
// the automatic launch (presize_buffers).
int ensure_partials(nmr_ctx* c, size_t doubles)
{
    if (doubles <= c->partials_cap) return NMR_OK;
    if (c->no_alloc)
        return fail(NMR_ERR_INVALID, "partial sums need %zu doubles, the fixed scratch holds %zu", doubles,
                    c->partials_cap);
    doubles = std::max(doubles, 2 * c->partials_cap);
    if (c->partials) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipFree(c->partials));
        c->device_bytes -= (int64_t)(c->partials_cap * sizeof(double));
        c->partials = nullptr;
        c->partials_cap = 0;
    }
    HIP_TRY(hipMalloc((void**)&c->partials, doubles * sizeof(double)));
    c->partials_cap = doubles;
    c->device_bytes += (int64_t)(doubles * sizeof(double));
    return NMR_OK;
}

int ensure_signal(nmr_ctx* c, size_t doubles)
{
    if (doubles <= c->dsignal_cap) return NMR_OK;
    if (c->no_alloc)
        return fail(NMR_ERR_INVALID, "signal rows need %zu doubles, the fixed scratch holds %zu", doubles,
                    c->dsignal_cap);
    doubles = std::max(doubles, 2 * c->dsignal_cap);
    if (c->dsignal) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipFree(c->dsignal));
        c->device_bytes -= (int64_t)(c->dsignal_cap * sizeof(double));
        c->dsignal = nullptr;
        c->dsignal_cap = 0;
    }
    HIP_TRY(hipMalloc((void**)&c->dsignal, doubles * sizeof(double)));
    c->dsignal_cap = doubles;
    c->device_bytes += (int64_t)(doubles * sizeof(double));
    return NMR_OK;
}

// Signal of the current state into dsignal[row*nc ...] (enqueued on the stream).
int enqueue_field_sum(nmr_ctx* c, size_t row)
{
    const nmr_problem& p = c->prob;
    if ((row + 1) * (size_t)c->nsig() > c->dsignal_cap)
        return fail(NMR_ERR_STATE, "signal buffer holds %zu values, row %zu needs %zu", c->dsignal_cap, row,
                    (row + 1) * (size_t)c->nsig());
    nmr::FieldSumArgs a{};
    // the state where it lives (rows preferred)
    const bool rows = (c->home & nmri::kHomeRows) != 0;
    for (int q = 0; q < 3; q++) a.f[q] = rows ? c->m[c->cur][q] : c->mb[c->bpar][q];
    if (!rows) {
        a.nby = (p.ny + 3) / 4;
        a.nb8 = (c->pitch / 4 + 1) / 2;
    }
    a.mask = c->mask;
    if (c->weighted) {
        a.w[0] = c->w[0];
        a.w[1] = c->w[1];
    }
    a.part = c->partials;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.pitch = c->pitch;
    int w[6];
    signal_window(p, w);
    a.bx0 = w[0]; a.bx1 = w[1]; a.by0 = w[2]; a.by1 = w[3]; a.bz0 = w[4]; a.bz1 = w[5];
    const int nfb = nmr::field_sum_blocks(a);
    int rc = ensure_partials(c, (size_t)nfb * c->nsig());
    if (rc) return rc;
    a.part = c->partials;
    HIP_TRY(nmr::launch_field_sum(c->nc, a, c->stream));
    // one "step" whose workgroups are the field sum's
    HIP_TRY(nmr::launch_reduce_partials(c->partials, 1, nfb, c->nsig(), c->dsignal + row * c->nsig(), c->stream));
    return NMR_OK;
}

int refresh_inert(nmr_ctx* c);
int velocity_sums(nmr_ctx* c, bool brick, bool centred);
bool bricks_fit(const nmr_ctx* c);
int ensure_bricks(nmr_ctx* c);
int ensure_pv_bricks(nmr_ctx* c);
int state_bricks(nmr_ctx* c);

// Launch geometry of the general sweep: 256 threads, tiles of 4*lanes x
// (256/lanes) voxels, z-chunks of lz planes (the round model unless set).
struct GGeom {
    int lanes, lz;
    long ntx, nty, ntz, nwg;
};
GGeom general_geometry(const nmr_ctx* c, bool gbrick)
{
    const nmr_problem& p = c->prob;
    GGeom g{};
    g.lanes = gbrick ? 16 : nmr::gsweep_lanes(p.nx);
    const int tx = 4 * g.lanes, ty = 256 / g.lanes;
    g.ntx = (p.nx + tx - 1) / tx;
    g.nty = (p.ny + ty - 1) / ty;
    int lz = c->lz > 0 ? c->lz
                       : round_lz(g.ntx * g.nty, p.nz,
                                  nmr::gsweep_resident_blocks(c->nc, c->vel, c->hetero, g.lanes, gbrick,
                                                              c->weighted, p.scheme != NMR_SCHEME_MACCORMACK));
    g.lz = std::min(std::max(lz, 1), p.nz);
    g.ntz = (p.nz + g.lz - 1) / g.lz;
    g.nwg = g.ntx * g.nty * g.ntz;
    return g;
}

// Launch geometry of the plain sweep: L1 for the run, L2 for an odd last
// step of a two-step run; *brick: whether the sweep runs on the bricks
// (one-step tiles, the layout switch on, every brick array addressable).
void plain_geometry(const nmr_ctx* c, Launch* L1, Launch* L2, bool* brick)
{
    const nmr_problem& p = c->prob;
    auto_launch(c, c->variant, c->lz, L1);
    const bool want = c->brick_on && bricks_fit(c);
    if (want) auto_launch(c, c->variant < 0 ? nmr::brick_auto_variant(p.nx) : c->variant, c->lz, L1, true);
    auto_launch(c, L1->steps == 2 ? 1 : L1->variant, L1->steps == 2 ? 0 : L1->lz, L2);
    *brick = want && L1->steps == 1 && nmr::brick_variant(L1->variant);
}

// Copies a run's signal rows to the host -- row 0 from the kept signal of
// the current state when there was one (sig0) -- and keeps the last row as
// the signal of the new current state.  Waits for the stream.
// Pinned host staging for at least `doubles` signal values (grown by
// doubling, at most kPinnedSignal); sized by the uploads so that runs of up
// to kChunkSteps steps never allocate it.
bool keep_state(const nmr_ctx* c)
{
    const nmr_problem& p = c->prob;
    return 2.0 * c->nc * 4.0 * (double)p.nx * (double)p.ny * (double)p.nz <= (double)kKeepStateBytes;
}

int ensure_host_signal(nmr_ctx* c, size_t doubles)
{
    if (doubles <= c->hsignal_cap) return NMR_OK;
    doubles = std::min(kPinnedSignal, std::max(doubles, 2 * c->hsignal_cap));
    if (c->hsignal) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipHostFree(c->hsignal));
        c->hsignal = nullptr;
        c->hsignal_cap = 0;
    }
    HIP_TRY(hipHostMalloc((void**)&c->hsignal, doubles * sizeof(double), hipHostMallocDefault));
    c->hsignal_cap = doubles;
    return NMR_OK;
}

int finish_signal(nmr_ctx* c, int64_t nsteps, bool have0, const double* sig0, double* signal)
{
    const size_t ns = (size_t)c->nsig(), first = have0 ? 1 : 0;
    // rows [r0, r0 + n) of dsignal go to dst: the caller's rows, or only the
    // last row (the signal of the new state) when the caller takes none
    size_t r0 = 0, n = 0;
    double* dst = nullptr;
    if (signal && (size_t)nsteps + 1 > first) {
        r0 = first;
        n = (size_t)nsteps + 1 - first;
        dst = signal + first * ns;
    } else if (!signal && (nsteps > 0 || !have0)) {
        r0 = (size_t)nsteps;
        n = 1;
        dst = c->sig_now;
    }
    // Through a pinned staging buffer the copy is one DMA that the stream
    // synchronisation below waits for; into pageable memory the runtime
    // blocks for the stream, stages the copy and blocks again (about 20 us
    // more per run).  Not under the no-allocation entry point, and not for
    // long signals (over 1 MiB), where that cost is noise.
    const size_t cnt = n * ns;
    double* stage = nullptr;
    if (cnt && !c->no_alloc && cnt <= kPinnedSignal) {
        int rc = ensure_host_signal(c, cnt);
        if (rc) return rc;
        stage = c->hsignal;
    }
    if (cnt)
        HIP_TRY(hipMemcpyAsync(stage ? stage : dst, c->dsignal + r0 * ns, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                               c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (stage) std::copy(stage, stage + cnt, dst);
    if (signal) {
        if (have0) std::copy(sig0, sig0 + ns, signal);
        std::copy(signal + (size_t)nsteps * ns, signal + ((size_t)nsteps + 1) * ns, c->sig_now);
    }
    c->sig_valid = true;
    return NMR_OK;
}

// Steps of the general problems (MacCormack, non-zero-derivative boundaries,
// per-voxel coefficients, two-phase sums): the general z-sweep, or with
// variant -2 the per-voxel kernel; each step's signal is fused into the sweep
// (plain sums; phase-weighted sums on bricks) or taken by the field sum.
int run_general(nmr_ctx* ctx, int64_t nsteps, const double* grad_sched, double* signal)
{
    const nmr_problem& p = ctx->prob;
    const int nc = ctx->nc, ns = ctx->nsig();
    const int64_t rows = nsteps + 1;
    int rc = ensure_signal(ctx, (size_t)rows * ns);
    if (rc) return rc;
    const bool mak = p.scheme == NMR_SCHEME_MACCORMACK;
    const bool pervoxel = ctx->variant == -2;
    if (!pervoxel && (rc = refresh_inert(ctx))) return rc;
    auto coef_of = [&](int64_t step) {
        double g[3] = {p.grad[0], p.grad[1], p.grad[2]};
        if (grad_sched)
            for (int q = 0; q < 3; q++) g[q] = grad_sched[3 * step + q];
        return make_coef(ctx, g);
    };
    // the general sweep runs on the bricks (state, velocity sums, per-voxel
    // arrays and the MacCormack predictor) whenever they are addressable
    const bool gbrick = !pervoxel && ctx->brick_on && bricks_fit(ctx);
    GGeom g{};
    if (!pervoxel) {
        g = general_geometry(ctx, gbrick);
        if (g.nwg > (1L << 30)) return fail(NMR_ERR_INVALID, "too many workgroups");
    }
    // each step's signal: fused into the sweep (plain sums; weighted sums
    // on bricks), else a field sum after the step
    const bool fused = !pervoxel && (!ctx->weighted || gbrick);
    if (fused && (rc = ensure_partials(ctx, (size_t)std::min<int64_t>(nsteps, kChunkSteps) * g.nwg * ns))) return rc;
    if ((rc = gbrick ? state_bricks(ctx) : nmri::state_rows(ctx))) return rc;
    if (gbrick && (rc = ensure_pv_bricks(ctx))) return rc;
    if (!pervoxel && (rc = velocity_sums(ctx, gbrick, mak))) return rc;
    const bool have0 = ctx->sig_valid;
    double sig0[6];
    std::copy(ctx->sig_now, ctx->sig_now + 6, sig0);
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if (!have0 && (rc = enqueue_field_sum(ctx, 0))) return rc;
    // from here on the sweeps write one layout only
    ctx->home = gbrick ? nmri::kHomeBricks : nmri::kHomeRows;
    if (pervoxel) {
        nmr::GenArgs a{};
        a.mask = ctx->mask;
        a.vx = ctx->vx;
        a.vy = ctx->vy;
        a.vz = ctx->vz;
        a.nx = p.nx;
        a.ny = p.ny;
        a.nz = p.nz;
        a.pitch = ctx->pitch;
        a.pitch1 = ctx->pitch1;
        for (int q = 0; q < 3; q++) a.bc[q] = p.bc[q];
        a.vel = ctx->vel ? 1 : 0;
        a.amask = (p.flags & NMR_F_ADVECT_MASK) ? 1 : 0;
        a.dn = ctx->dn;
        a.r1v = ctx->r1v;
        a.r2v = ctx->r2v;
        a.bn = ctx->bn;
        a.bnax = ctx->bnax;
        for (int64_t s = 0; s < nsteps; s++) {
            a.c = coef_of(s);
            for (int q = 0; q < 3; q++) {
                a.m[q] = ctx->m[ctx->cur][q];
                a.p[q] = ctx->pred[q];
                a.out[q] = mak ? ctx->pred[q] : ctx->m[ctx->cur ^ 1][q];
            }
            HIP_TRY(nmr::launch_general(nc, mak ? nmr::GEN_PREDICT : nmr::GEN_UPWIND, a, ctx->stream));
            if (mak) {
                for (int q = 0; q < 3; q++) a.out[q] = ctx->m[ctx->cur ^ 1][q];
                HIP_TRY(nmr::launch_general(nc, nmr::GEN_CORRECT, a, ctx->stream));
            }
            ctx->cur ^= 1;
            if ((rc = enqueue_field_sum(ctx, (size_t)s + 1))) return rc;
        }
    } else {
        // The corrector leaves 0 at solids but skips all-solid chunks: the
        // first step's output buffer is cleared at solids before the step and
        // its input buffer after it (the first step still reads the input's
        // solid values, e.g. through periodic ghosts).
        auto out_buf = [&]() { return gbrick ? ctx->mb[ctx->bpar ^ 1] : ctx->m[ctx->cur ^ 1]; };
        auto zero_solids = [&](float* const* m) {
            return gbrick ? nmr::launch_zero_solids_brick(m, nc, ctx->mask, p.ny, p.nz, ctx->pitch, ctx->stream)
                          : nmr::launch_zero_solids(m, nc, ctx->mask, ctx->n_m, ctx->stream);
        };
        const bool zero_first = mak && nsteps > 0 && !ctx->solids_zeroed;
        if (zero_first) HIP_TRY(zero_solids(out_buf()));
        int w[6];
        signal_window(p, w);
        nmr::GSweepArgs a{};
        a.mask = ctx->mask;
        // per-voxel arrays: their brick copies on bricks
        a.dn = ctx->dn ? (gbrick ? ctx->pvb[0] : ctx->dn) : nullptr;
        a.r1v = ctx->r1v ? (gbrick ? ctx->pvb[1] : ctx->r1v) : nullptr;
        a.r2v = ctx->r2v ? (gbrick ? ctx->pvb[2] : ctx->r2v) : nullptr;
        a.bn = ctx->bn ? (gbrick ? ctx->pvb[3] : ctx->bn) : nullptr;
        a.bnax = ctx->bnax;
        if (gbrick && ctx->weighted) {
            a.w[0] = ctx->pvb[4];
            a.w[1] = ctx->pvb[5];
        }
        a.nx = p.nx;
        a.ny = p.ny;
        a.nz = p.nz;
        a.pitch = ctx->pitch;
        a.pitch1 = ctx->pitch1;
        a.ntx = (int)g.ntx;
        a.nty = (int)g.nty;
        a.ntz = (int)g.ntz;
        a.lz = g.lz;
        a.nblocks = (int)g.nwg;
        a.bx0 = w[0]; a.bx1 = w[1]; a.by0 = w[2]; a.by1 = w[3]; a.bz0 = w[4]; a.bz1 = w[5];
        for (int q = 0; q < 3; q++) a.bc[q] = p.bc[q];
        a.skip_m = ctx->inert;
        a.amask = (p.flags & NMR_F_ADVECT_MASK) ? 1 : 0;
        a.nby = (p.ny + 3) / 4;
        for (int q = 0; q < 3; q++)
            a.cs[q] = mak ? (gbrick ? ctx->csumb[q] : ctx->csum[q]) : (gbrick ? ctx->advb[q] : ctx->adv[q]);
        const bool coef = ctx->hetero;
        const bool keep = gbrick && keep_state(ctx);
        int64_t done = 0;
        while (done < nsteps) {
            const int chunk = (int)std::min<int64_t>(kChunkSteps, nsteps - done);
            for (int s = 0; s < chunk; s++) {
                a.c = coef_of(done + s);
                float* const* cin = gbrick ? ctx->mb[ctx->bpar] : ctx->m[ctx->cur];
                float* const* cout = out_buf();
                float* const* pred = gbrick ? ctx->pb : ctx->pred;
                if (mak) {
                    for (int q = 0; q < 3; q++) {
                        a.f[q] = cin[q];
                        a.out[q] = pred[q];
                    }
                    a.partials = nullptr;
                    HIP_TRY(nmr::launch_gsweep(nc, ctx->vel, nmr::GEN_PREDICT, coef, g.lanes, a, ctx->stream, gbrick,
                                               keep));
                }
                for (int q = 0; q < 3; q++) {
                    a.f[q] = mak ? pred[q] : cin[q];
                    a.mavg[q] = cin[q];
                    a.out[q] = cout[q];
                }
                a.partials = fused ? ctx->partials + (size_t)s * g.nwg * ns : nullptr;
                HIP_TRY(nmr::launch_gsweep(nc, ctx->vel, mak ? nmr::GEN_CORRECT : nmr::GEN_UPWIND, coef, g.lanes, a,
                                           ctx->stream, gbrick, keep));
                if (gbrick) ctx->bpar ^= 1;
                else ctx->cur ^= 1;
                if (zero_first && done + s == 0) { // the next output buffer: the first step's input
                    HIP_TRY(zero_solids(out_buf()));
                    ctx->solids_zeroed = true;
                }
                if (!fused && (rc = enqueue_field_sum(ctx, (size_t)(done + s) + 1))) return rc;
            }
            if (fused)
                HIP_TRY(nmr::launch_reduce_partials(ctx->partials, chunk, (int)g.nwg, ns,
                                                    ctx->dsignal + (size_t)(done + 1) * ns, ctx->stream));
            done += chunk;
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    if ((rc = finish_signal(ctx, nsteps, have0, sig0, signal))) return rc;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    ctx->last_sweeps = nsteps * (mak ? 2 : 1);
    return NMR_OK;
}

// Whether every brick array stays inside the sweep's 32-bit byte offsets.
bool bricks_fit(const nmr_ctx* c)
{
    const nmr_problem& p = c->prob;
    const int nbx = c->pitch / 4, nbx1 = c->pitch1 / 4;
    const size_t big = std::max({nmr::brick_floats(nbx, p.ny, p.nz), nmr::brick_floats(nbx1, p.ny, p.nz),
                                 nmr::brick_floats(nbx, p.ny + 1, p.nz), nmr::brick_floats(nbx, p.ny, p.nz + 1)});
    return big * 4 < ((size_t)1 << 32) - 256;
}

// Brick buffers of the state (both ping-pong buffers and, for MacCormack,
// the predictor), allocated by the uploads (or the first brick run after
// the layout switch); the velocity enters the brick sweeps as per-voxel
// sums (velocity_sums).
int ensure_bricks(nmr_ctx* c)
{
    const nmr_problem& p = c->prob;
    const int nbx = c->pitch / 4;
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < c->nc; q++)
            if (!c->mb[b][q]) {
                int rc = dev_alloc(c, &c->mb[b][q], nmr::brick_floats(nbx, p.ny, p.nz));
                if (rc) return rc;
            }
    if (p.scheme == NMR_SCHEME_MACCORMACK)
        for (int q = 0; q < c->nc; q++)
            if (!c->pb[q]) {
                int rc = dev_alloc(c, &c->pb[q], nmr::brick_floats(nbx, p.ny, p.nz));
                if (rc) return rc;
            }
    return NMR_OK;
}

// Per-voxel face-velocity sums (upwind: those of the plain sweep; centred:
// those of the MacCormack stages) in rows (also the source of the brick
// copy) and, for brick sweeps, in bricks; computed once per velocity upload.
int velocity_sums(nmr_ctx* c, bool brick, bool centred)
{
    if (!c->vel) return NMR_OK;
    const nmr_problem& p = c->prob;
    float** rows = centred ? c->csum : c->adv;
    float** bricks = centred ? c->csumb : c->advb;
    bool& rows_ok = centred ? c->csum_valid : c->adv_valid;
    bool& bricks_ok = centred ? c->csumb_valid : c->advb_valid;
    int rc;
    if (!rows_ok) {
        nmr::AdvArgs a{};
        for (int q = 0; q < 3; q++) {
            if (!rows[q] && (rc = dev_alloc(c, &rows[q], c->n_m))) return rc;
            a.out[q] = rows[q];
        }
        a.vx = c->vx;
        a.vy = c->vy;
        a.vz = c->vz;
        a.nx = p.nx;
        a.ny = p.ny;
        a.nz = p.nz;
        a.pitch = c->pitch;
        a.pitch1 = c->pitch1;
        a.centred = centred ? 1 : 0;
        HIP_TRY(nmr::launch_adv(a, c->stream));
        rows_ok = true;
        bricks_ok = false;
    }
    if (brick && !bricks_ok) {
        const int nbx = c->pitch / 4;
        for (int q = 0; q < 3; q++) {
            if (!bricks[q] && (rc = dev_alloc(c, &bricks[q], nmr::brick_floats(nbx, p.ny, p.nz)))) return rc;
            HIP_TRY(nmr::launch_brick(rows[q], bricks[q], nbx, p.ny, p.nz, true, c->stream));
        }
        bricks_ok = true;
    }
    return NMR_OK;
}

int ensure_adv(nmr_ctx* c, bool brick) { return velocity_sums(c, brick, false); }

// Brick copies of the per-voxel arrays the general sweep reads (those
// uploaded), converted once after each upload.
int ensure_pv_bricks(nmr_ctx* c)
{
    if (c->brick_pv_valid) return NMR_OK;
    const nmr_problem& p = c->prob;
    const int nbx = c->pitch / 4;
    float* const src[6] = {c->dn, c->r1v, c->r2v, c->bn, c->w[0], c->w[1]};
    for (int i = 0; i < 6; i++) {
        if (!src[i]) continue;
        if (!c->pvb[i]) {
            int rc = dev_alloc(c, &c->pvb[i], nmr::brick_floats(nbx, p.ny, p.nz));
            if (rc) return rc;
        }
        HIP_TRY(nmr::launch_brick(src[i], c->pvb[i], nbx, p.ny, p.nz, true, c->stream));
    }
    c->brick_pv_valid = true;
    return NMR_OK;
}

// The current state into both brick buffers, from the row buffer m[cur],
// unless the bricks already hold it.
int state_bricks(nmr_ctx* c)
{
    if (c->home & nmri::kHomeBricks) return NMR_OK;
    int rc = ensure_bricks(c);
    if (rc) return rc;
    const nmr_problem& p = c->prob;
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < c->nc; q++)
            HIP_TRY(nmr::launch_brick(c->m[c->cur][q], c->mb[b][q], c->pitch / 4, p.ny, p.nz, true, c->stream));
    c->bpar = 0;
    c->home |= nmri::kHomeBricks;
    return NMR_OK;
}

// Partial-sum and signal buffers sized for the launch the next run takes
// (chunks of kChunkSteps steps, and a field sum of the current state).
int presize_buffers(nmr_ctx* c)
{
    const nmr_problem& p = c->prob;
    const size_t ns = (size_t)c->nsig();
    size_t need = 0;
    if (c->general) {
        if (c->variant != -2) need = (size_t)kChunkSteps * general_geometry(c, c->brick_on && bricks_fit(c)).nwg * ns;
    } else {
        Launch L1{}, L2{};
        bool brick = false;
        plain_geometry(c, &L1, &L2, &brick);
        if (L1.nwg > 0 && L2.nwg > 0)
            need = (size_t)kChunkSteps * std::max(L1.nwg, L2.nwg) * ns;
    }
    nmr::FieldSumArgs f{};
    int w[6];
    signal_window(p, w);
    f.bx0 = w[0]; f.bx1 = w[1]; f.by0 = w[2]; f.by1 = w[3]; f.bz0 = w[4]; f.bz1 = w[5];
    need = std::max(need, (size_t)nmr::field_sum_blocks(f) * ns);
    int rc = ensure_partials(c, need);
    if (!rc) rc = ensure_signal(c, (size_t)(kChunkSteps + 1) * ns);
    if (!rc) rc = ensure_host_signal(c, (size_t)(kChunkSteps + 1) * ns);
    return rc;
}

// The one-time work of a solve, done by the uploads so that runs start at
// once: the inert check, the brick buffers with the velocity and per-voxel
// arrays converted, the partial-sum and signal buffers.
int prepare(nmr_ctx* c)
{
    if (!c->has_mask) return NMR_OK;
    int rc = refresh_inert(c);
    if (rc) return rc;
    if (c->brick_on && bricks_fit(c)) {
        if ((rc = ensure_bricks(c))) return rc;
        if ((rc = ensure_pv_bricks(c))) return rc;
    }
    // the per-voxel velocity sums the sweeps read: upwind (the plain sweep,
    // the general sweep's upwind stage) or centred (MacCormack)
    if ((rc = velocity_sums(c, c->brick_on && bricks_fit(c), c->prob.scheme == NMR_SCHEME_MACCORMACK))) return rc;
    if ((rc = presize_buffers(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NMR_OK;
}

// Whether solid magnetization can reach a pore (a face touching a solid
// carries velocity); recomputed on the device after mask/velocity uploads.
int refresh_inert(nmr_ctx* c)
{
    if (!c->inert_dirty) return NMR_OK;
    if (!c->vel) {
        c->inert = 1;
    } else {
        HIP_TRY(hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
        nmr::InertArgs a{};
        a.mask = c->mask;
        a.vx = c->vx;
        a.vy = c->vy;
        a.vz = c->vz;
        a.flag = c->dflag;
        a.nx = c->prob.nx;
        a.ny = c->prob.ny;
        a.nz = c->prob.nz;
        a.pitch = c->pitch;
        a.pitch1 = c->pitch1;
        HIP_TRY(nmr::launch_inert_check(a, c->stream));
        int h = 0;
        HIP_TRY(hipMemcpyAsync(&h, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->inert = h == 0;
    }
    c->inert_dirty = false;
    return NMR_OK;
}

} // namespace

extern "C" {

int nmr_hip_abi_version(void) { return 2; }

const char* nmr_hip_strerror(int code)
{
    switch (code) {
    case NMR_OK: return "success";
    case NMR_ERR_INVALID: return "invalid argument";
    case NMR_ERR_NULL: return "null pointer";
    case NMR_ERR_DEVICE: return "HIP device error";
    case NMR_ERR_NOMEM: return "device out of memory";
    case NMR_ERR_STATE: return "call out of order";
    case NMR_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown status";
    }
}

const char* nmr_hip_last_error(void) { return g_last_error.c_str(); }

int nmr_hip_num_variants(void) { return nmr::num_variants(); }

int nmr_hip_variant_info(int32_t variant, int32_t* lanes_x, int32_t* rows_per_thread, int32_t* tile_x,
                         int32_t* tile_y, int32_t* steps_per_pass)
{
    int lx, rpt, tx, ty, st;
    if (!nmr::variant_shape(variant, &lx, &rpt, &tx, &ty, &st))
        return fail(NMR_ERR_INVALID, "no variant %d", (int)variant);
    if (lanes_x) *lanes_x = lx;
    if (rows_per_thread) *rows_per_thread = rpt;
    if (tile_x) *tile_x = tx;
    if (tile_y) *tile_y = ty;
    if (steps_per_pass) *steps_per_pass = st;
    return NMR_OK;
}

int nmr_hip_create(const nmr_problem* prob, nmr_ctx** out)
{
    if (!prob || !out) return fail(NMR_ERR_NULL, "nmr_hip_create: null argument");
    *out = nullptr;
    const nmr_problem& p = *prob;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) return fail(NMR_ERR_INVALID, "grid %d x %d x %d", p.nx, p.ny, p.nz);
    if ((long)p.nx * (long)p.ny * (long)p.nz > (1L << 31) - 1 || p.nx > (1 << 20) || p.ny > (1 << 20))
        return fail(NMR_ERR_INVALID, "grid %d x %d x %d exceeds 2^31 voxels", p.nx, p.ny, p.nz);
    if (p.ncomp < 1 || p.ncomp > 3) return fail(NMR_ERR_INVALID, "ncomp %d", p.ncomp);
    if (p.scheme != NMR_SCHEME_UPWIND && p.scheme != NMR_SCHEME_MACCORMACK)
        return fail(NMR_ERR_UNSUPPORTED, "scheme %d", p.scheme);
    bool general = p.scheme == NMR_SCHEME_MACCORMACK;
    for (int a = 0; a < 3; a++) {
        if (p.bc[a] < NMR_BC_NEUMANN || p.bc[a] > NMR_BC_DIRICHLET) return fail(NMR_ERR_INVALID, "bc[%d] = %d", a, p.bc[a]);
        general = general || p.bc[a] != NMR_BC_NEUMANN;
    }
    // masked-face advection runs in the general sweep (MacCormack masks its
    // advection by construction; the flag changes nothing there)
    general = general || ((p.flags & NMR_F_ADVECT_MASK) && (p.flags & NMR_F_VELOCITY));
    if (general && p.phase_mode == NMR_PHASE_EXACT)
        return fail(NMR_ERR_UNSUPPORTED, "exact phase needs the upwind scheme with zero-derivative boundaries");
    if (p.phase_mode != NMR_PHASE_LINEAR && p.phase_mode != NMR_PHASE_EXACT)
        return fail(NMR_ERR_INVALID, "phase_mode %d", p.phase_mode);
    if (p.flags & ~(NMR_F_DIFFUSION | NMR_F_VELOCITY | NMR_F_PRECESSION | NMR_F_ADVECT_MASK))
        return fail(NMR_ERR_INVALID, "flags 0x%x", p.flags);
    if (!(p.dx > 0) || !(p.dy > 0) || !(p.dz > 0) || !(p.dt > 0))
        return fail(NMR_ERR_INVALID, "dx, dy, dz, dt must be > 0");
    for (int a = 0; a < 3; a++)
        if (p.buffer[a] < 0) return fail(NMR_ERR_INVALID, "buffer[%d] < 0", a);
    const double nb = std::sqrt(p.b0_dir[0] * p.b0_dir[0] + p.b0_dir[1] * p.b0_dir[1] + p.b0_dir[2] * p.b0_dir[2]);
    if (!(nb > 0)) return fail(NMR_ERR_INVALID, "b0_dir is zero");
    if (p.ncomp < 3 && (p.b0_dir[0] != 0.0 || p.b0_dir[1] != 0.0))
        return fail(NMR_ERR_UNSUPPORTED, "ncomp < 3 needs the main field along z");
    if (p.ncomp == 1 && (p.flags & NMR_F_PRECESSION))
        return fail(NMR_ERR_UNSUPPORTED, "precession needs ncomp >= 2");
    if (p.ncomp == 2 && (p.grad[0] != 0.0 || p.grad[1] != 0.0))
        return fail(NMR_ERR_UNSUPPORTED, "ncomp 2 takes a z gradient only");
    if (p.phase_mode == NMR_PHASE_EXACT && p.ncomp != 2)
        return fail(NMR_ERR_UNSUPPORTED, "exact phase needs ncomp 2");

    nmr_ctx* c = new (std::nothrow) nmr_ctx();
    if (!c) return fail(NMR_ERR_NOMEM, "host allocation");
    c->prob = p;
    c->general = general;
    c->nc = p.ncomp;
    c->vel = (p.flags & NMR_F_VELOCITY) != 0;
    c->exact = p.phase_mode == NMR_PHASE_EXACT;
    c->pitch = round_up4(p.nx);
    c->pitch1 = round_up4(p.nx + 1);
    c->n_m = (size_t)p.nz * p.ny * c->pitch;
    // the sweep addresses every array with 32-bit byte offsets
    const size_t biggest = nmri::biggest_array_bytes(p.nx, p.ny, p.nz);
    if (biggest >= ((size_t)1 << 32) - 64) {
        delete c;
        return fail(NMR_ERR_UNSUPPORTED, "arrays of %zu bytes exceed the 4 GiB addressing of the sweep", biggest);
    }
    auto bail = [&](int rc) {
        free_all(c);
        delete c;
        return rc;
    };
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) return bail(fail(NMR_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e)));
    int rc = scratch_address(c->device, &c->scratch);
    if (rc) return bail(rc);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail(NMR_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e)));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(NMR_ERR_DEVICE, "hipEventCreate failed"));
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < c->nc; q++)
            if ((rc = dev_alloc(c, &c->m[b][q], c->n_m))) return bail(rc);
    if (p.scheme == NMR_SCHEME_MACCORMACK)
        for (int q = 0; q < c->nc; q++)
            if ((rc = dev_alloc(c, &c->pred[q], c->n_m))) return bail(rc);
    if ((rc = dev_alloc(c, &c->mask, c->n_m))) return bail(rc);
    if ((rc = dev_alloc(c, &c->dflag, 64))) return bail(rc);
    if (c->vel) {
        c->n_vx = (size_t)p.nz * p.ny * c->pitch1;
        c->n_vy = (size_t)p.nz * (p.ny + 1) * c->pitch;
        c->n_vz = (size_t)(p.nz + 1) * p.ny * c->pitch;
        if ((rc = dev_alloc(c, &c->vx, c->n_vx))) return bail(rc);
        if ((rc = dev_alloc(c, &c->vy, c->n_vy))) return bail(rc);
        if ((rc = dev_alloc(c, &c->vz, c->n_vz))) return bail(rc);
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(fail(NMR_ERR_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e)));
    *out = c;
    return NMR_OK;
}

int nmr_hip_destroy(nmr_ctx* ctx)
{
    if (!ctx) return NMR_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return NMR_OK;
}

int nmr_hip_device_bytes(nmr_ctx* ctx, int64_t* bytes)
{
    if (!ctx || !bytes) return fail(NMR_ERR_NULL, "nmr_hip_device_bytes: null argument");
    *bytes = ctx->device_bytes;
    return NMR_OK;
}

int nmr_hip_upload_mask(nmr_ctx* ctx, const uint8_t* mask, int32_t layout)
{
    if (!ctx || !mask) return fail(NMR_ERR_NULL, "nmr_hip_upload_mask: null argument");
    int rc = check_layout(layout);
    if (rc) return rc;
    const nmr_problem& p = ctx->prob;
    std::vector<uint8_t> h(ctx->n_m, 0);
    for (int k = 0; k < p.nz; k++)
        for (int j = 0; j < p.ny; j++)
            for (int i = 0; i < p.nx; i++) {
                uint8_t v;
                if (layout == NMR_LAYOUT_COMPACT) {
                    v = mask[(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k)];
                    if (v > 1) return fail(NMR_ERR_INVALID, "compact mask value %d (want 0/1)", (int)v);
                } else {
                    const uint8_t ch = mask[pidx(p, i + 1, j + 1, k + 1)];
                    if (ch != '0' && ch != '1') return fail(NMR_ERR_INVALID, "padded mask char 0x%02x (want '0'/'1')", ch);
                    v = (ch == '1');
                }
                h[(size_t)k * p.ny * ctx->pitch + (size_t)j * ctx->pitch + i] = v;
            }
    HIP_TRY(hipMemcpyAsync(ctx->mask, h.data(), h.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->has_mask = true;
    ctx->solids_zeroed = false;
    ctx->sig_valid = false;
    ctx->mask_version += 1;
    ctx->inert_dirty = true;
    return prepare(ctx);
}

int nmr_hip_upload_velocity(nmr_ctx* ctx, const double* vx, const double* vy, const double* vz, int32_t layout)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_upload_velocity: null context");
    int rc = check_layout(layout);
    if (rc) return rc;
    if (!ctx->vel) return fail(NMR_ERR_STATE, "velocity upload without NMR_F_VELOCITY");
    const nmr_problem& p = ctx->prob;
    const int P = ctx->pitch, P1 = ctx->pitch1;
    {
        std::vector<float> h(ctx->n_vx, 0.f);
        if (vx)
            for (int k = 0; k < p.nz; k++)
                for (int j = 0; j < p.ny; j++)
                    for (int i = 0; i <= p.nx; i++) {
                        const double v = (layout == NMR_LAYOUT_COMPACT)
                                             ? vx[(size_t)i + (size_t)(p.nx + 1) * ((size_t)j + (size_t)p.ny * k)]
                                             : vx[pidx(p, i + 1, j + 1, k + 1)];
                        h[(size_t)k * p.ny * P1 + (size_t)j * P1 + i] = (float)v;
                    }
        HIP_TRY(hipMemcpyAsync(ctx->vx, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    {
        std::vector<float> h(ctx->n_vy, 0.f);
        if (vy)
            for (int k = 0; k < p.nz; k++)
                for (int j = 0; j <= p.ny; j++)
                    for (int i = 0; i < p.nx; i++) {
                        const double v = (layout == NMR_LAYOUT_COMPACT)
                                             ? vy[(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)(p.ny + 1) * k)]
                                             : vy[pidx(p, i + 1, j + 1, k + 1)];
                        h[(size_t)k * (p.ny + 1) * P + (size_t)j * P + i] = (float)v;
                    }
        HIP_TRY(hipMemcpyAsync(ctx->vy, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    {
        std::vector<float> h(ctx->n_vz, 0.f);
        if (vz)
            for (int k = 0; k <= p.nz; k++)
                for (int j = 0; j < p.ny; j++)
                    for (int i = 0; i < p.nx; i++) {
                        const double v = (layout == NMR_LAYOUT_COMPACT)
                                             ? vz[(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k)]
                                             : vz[pidx(p, i + 1, j + 1, k + 1)];
                        h[(size_t)k * p.ny * P + (size_t)j * P + i] = (float)v;
                    }
        HIP_TRY(hipMemcpyAsync(ctx->vz, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    ctx->inert_dirty = true;
    ctx->adv_valid = ctx->advb_valid = false;
    ctx->csum_valid = ctx->csumb_valid = false;
    return prepare(ctx);
}

int nmr_hip_upload_magnetization(nmr_ctx* ctx, const double* mx, const double* my, const double* mz, int32_t layout)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_upload_magnetization: null context");
    int rc = check_layout(layout);
    if (rc) return rc;
    const double* src[3] = {mx, my, mz};
    for (int q = 0; q < ctx->nc; q++)
        if (!src[q]) return fail(NMR_ERR_NULL, "magnetization component %d is NULL", q);
    const nmr_problem& p = ctx->prob;
    std::vector<float> h(ctx->n_m, 0.f);
    for (int q = 0; q < ctx->nc; q++) {
        for (int k = 0; k < p.nz; k++)
            for (int j = 0; j < p.ny; j++)
                for (int i = 0; i < p.nx; i++) {
                    const double v = (layout == NMR_LAYOUT_COMPACT)
                                         ? src[q][(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k)]
                                         : src[q][pidx(p, i + 1, j + 1, k + 1)];
                    h[(size_t)k * p.ny * ctx->pitch + (size_t)j * ctx->pitch + i] = (float)v;
                }
        // both ping-pong buffers: the sweeps do not rewrite all-solid chunks,
        // whose values therefore have to agree in the two buffers
        for (int b = 0; b < 2; b++)
            HIP_TRY(hipMemcpyAsync(ctx->m[b][q], h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                                   ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    ctx->has_mag = true;
    ctx->solids_zeroed = false;
    ctx->sig_valid = false;
    ctx->home = nmri::kHomeRows;
    ctx->cur = 0;
    if ((rc = prepare(ctx))) return rc;
    // the sweeps keep the state in bricks: converted here, once per upload
    if (ctx->brick_on && bricks_fit(ctx) && (rc = state_bricks(ctx))) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NMR_OK;
}

// Reads one per-voxel field of the problem grid (compact or padded).
static int read_voxels(const nmr_ctx* ctx, const double* src, int32_t layout, std::vector<double>& out)
{
    const nmr_problem& p = ctx->prob;
    out.assign((size_t)p.nx * p.ny * p.nz, 0.0);
    for (int k = 0; k < p.nz; k++)
        for (int j = 0; j < p.ny; j++)
            for (int i = 0; i < p.nx; i++) {
                const size_t o = (size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k);
                const double v = layout == NMR_LAYOUT_COMPACT ? src[o] : src[pidx(p, i + 1, j + 1, k + 1)];
                if (!std::isfinite(v)) return fail(NMR_ERR_INVALID, "non-finite value at (%d, %d, %d)", i, j, k);
                out[o] = v;
            }
    return NMR_OK;
}

static int put_voxels(nmr_ctx* ctx, float** dst, const std::vector<double>& v, double scale)
{
    const nmr_problem& p = ctx->prob;
    if (!*dst) {
        int rc = dev_alloc(ctx, dst, ctx->n_m);
        if (rc) return rc;
    }
    ctx->brick_pv_valid = false;
    std::vector<float> h(ctx->n_m, 0.f);
    for (int k = 0; k < p.nz; k++)
        for (int j = 0; j < p.ny; j++)
            for (int i = 0; i < p.nx; i++)
                h[(size_t)k * p.ny * ctx->pitch + (size_t)j * ctx->pitch + i] =
                    (float)(v[(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k)] * scale);
    HIP_TRY(hipMemcpyAsync(*dst, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return NMR_OK;
}

int nmr_hip_upload_coefficients(nmr_ctx* ctx, const double* D, const double* r1, const double* r2, int32_t layout)
{
    if (!ctx || !D || !r2 || (ctx->nc == 3 && !r1)) return fail(NMR_ERR_NULL, "nmr_hip_upload_coefficients: null argument");
    int rc = check_layout(layout);
    if (rc) return rc;
    if (ctx->exact) return fail(NMR_ERR_UNSUPPORTED, "per-voxel coefficients use the linear phase");
    std::vector<double> d, a, b;
    if ((rc = read_voxels(ctx, D, layout, d)) || (rc = read_voxels(ctx, r2, layout, b))) return rc;
    if (r1 && (rc = read_voxels(ctx, r1, layout, a))) return rc;
    double dref = 0.0;
    for (double v : d) {
        if (v < 0.0) return fail(NMR_ERR_INVALID, "negative diffusivity");
        dref = std::max(dref, v);
    }
    if (!(dref > 0.0)) dref = 1.0; // no diffusion anywhere
    const double dt = ctx->prob.dt;
    if ((rc = put_voxels(ctx, &ctx->dn, d, 1.0 / dref)) || (rc = put_voxels(ctx, &ctx->r2v, b, dt))) return rc;
    if (r1 && (rc = put_voxels(ctx, &ctx->r1v, a, dt))) return rc;
    ctx->d_ref = dref;
    ctx->hetero = true;
    ctx->general = true;
    return prepare(ctx);
}

int nmr_hip_upload_field(nmr_ctx* ctx, const double* b, int32_t layout)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_upload_field: null context");
    int rc = check_layout(layout);
    if (rc) return rc;
    if (!b) { // back to the gradient field
        if (ctx->bn) (void)hipFree(ctx->bn);
        ctx->bn = nullptr;
        ctx->bnax = -1;
        return NMR_OK;
    }
    if (ctx->exact) return fail(NMR_ERR_UNSUPPORTED, "per-voxel fields use the linear phase");
    const nmr_problem& p = ctx->prob;
    int ax = -1, naxes = 0;
    for (int q = 0; q < 3; q++)
        if (p.b0_dir[q] != 0.0) {
            ax = q;
            naxes += 1;
        }
    if (naxes != 1) return fail(NMR_ERR_UNSUPPORTED, "a per-voxel field needs the main field along one axis");
    std::vector<double> v;
    if ((rc = read_voxels(ctx, b, layout, v))) return rc;
    for (double x : v)
        if (!std::isfinite(x)) return fail(NMR_ERR_INVALID, "per-voxel field values must be finite");
    if ((rc = put_voxels(ctx, &ctx->bn, v, 1.0))) return rc;
    ctx->bnax = ax;
    ctx->general = true;
    return prepare(ctx);
}

int nmr_hip_upload_phase_weights(nmr_ctx* ctx, const double* w1, const double* w2, int32_t layout)
{
    if (!ctx || !w1 || !w2) return fail(NMR_ERR_NULL, "nmr_hip_upload_phase_weights: null argument");
    int rc = check_layout(layout);
    if (rc) return rc;
    std::vector<double> a, b;
    if ((rc = read_voxels(ctx, w1, layout, a)) || (rc = read_voxels(ctx, w2, layout, b))) return rc;
    if ((rc = put_voxels(ctx, &ctx->w[0], a, 1.0)) || (rc = put_voxels(ctx, &ctx->w[1], b, 1.0))) return rc;
    ctx->weighted = true;
    ctx->general = true;
    ctx->sig_valid = false; // signal rows now hold 2*ncomp sums
    return prepare(ctx);
}

int nmr_hip_download_magnetization(nmr_ctx* ctx, double* mx, double* my, double* mz, int32_t layout)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_download_magnetization: null context");
    int rc = check_layout(layout);
    if (rc) return rc;
    double* dst[3] = {mx, my, mz};
    const nmr_problem& p = ctx->prob;
    if ((rc = nmri::state_rows(ctx))) return rc;
    std::vector<float> h(ctx->n_m);
    for (int q = 0; q < ctx->nc; q++) {
        if (!dst[q]) continue;
        HIP_TRY(hipMemcpyAsync(h.data(), ctx->m[ctx->cur][q], h.size() * sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (layout == NMR_LAYOUT_COMPACT) {
            for (int k = 0; k < p.nz; k++)
                for (int j = 0; j < p.ny; j++)
                    for (int i = 0; i < p.nx; i++)
                        dst[q][(size_t)i + (size_t)p.nx * ((size_t)j + (size_t)p.ny * k)] =
                            h[(size_t)k * p.ny * ctx->pitch + (size_t)j * ctx->pitch + i];
        } else {
            // interior plus the ghost layer each axis's boundary condition
            // leaves (copy of the adjacent cell, opposite cell, or 0)
            auto ghost = [](int i, int n, int bc, bool* zero) {
                if (i >= 1 && i <= n) return i - 1;
                if (bc == NMR_BC_DIRICHLET) *zero = true;
                if (bc == NMR_BC_PERIODIC) return i == 0 ? n - 1 : 0;
                return i == 0 ? 0 : n - 1;
            };
            for (int k = 0; k < p.nz + 2; k++)
                for (int j = 0; j < p.ny + 2; j++)
                    for (int i = 0; i < p.nx + 2; i++) {
                        bool zero = false;
                        const int ii = ghost(i, p.nx, p.bc[0], &zero);
                        const int jj = ghost(j, p.ny, p.bc[1], &zero);
                        const int kk = ghost(k, p.nz, p.bc[2], &zero);
                        dst[q][pidx(p, i, j, k)] =
                            zero ? 0.0 : h[(size_t)kk * p.ny * ctx->pitch + (size_t)jj * ctx->pitch + ii];
                    }
        }
    }
    return NMR_OK;
}

int nmr_hip_set_layout(nmr_ctx* ctx, int32_t brick)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_set_layout: null context");
    if (brick != 0 && brick != 1) return fail(NMR_ERR_INVALID, "layout %d", (int)brick);
    ctx->brick_on = brick != 0;
    return NMR_OK;
}

int nmr_hip_set_tile(nmr_ctx* ctx, int32_t variant, int32_t lz)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_set_tile: null context");
    if (variant < -2 || variant >= nmr::num_variants()) return fail(NMR_ERR_INVALID, "variant %d", (int)variant);
    if (variant == -2 && ctx->exact) return fail(NMR_ERR_UNSUPPORTED, "the per-voxel kernel has no exact phase");
    if (lz < 0) return fail(NMR_ERR_INVALID, "lz %d", (int)lz);
    ctx->variant = variant;
    ctx->lz = lz;
    return NMR_OK;
}

int nmr_hip_signal(nmr_ctx* ctx, double* out)
{
    if (!ctx || !out) return fail(NMR_ERR_NULL, "nmr_hip_signal: null argument");
    if (!ctx->has_mask || !ctx->has_mag) return fail(NMR_ERR_STATE, "signal before mask and magnetization upload");
    const int ns = ctx->nsig();
    if (!ctx->sig_valid) { // a field sum of the state where it lives
        int rc = ensure_signal(ctx, (size_t)ns);
        if (rc) return rc;
        if ((rc = enqueue_field_sum(ctx, 0))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->sig_now, ctx->dsignal, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->sig_valid = true;
    }
    std::copy(ctx->sig_now, ctx->sig_now + ns, out);
    return NMR_OK;
}

int nmr_hip_pulse(nmr_ctx* ctx, int32_t flag, double param, double ax, double ay, double az)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_pulse: null context");
    if (flag < 1 || flag > 4) return fail(NMR_ERR_INVALID, "pulse flag %d", (int)flag);
    if (!std::isfinite(param) || !std::isfinite(ax) || !std::isfinite(ay) || !std::isfinite(az))
        return fail(NMR_ERR_INVALID, "pulse parameters must be finite");
    if (!ctx->has_mask || !ctx->has_mag) return fail(NMR_ERR_STATE, "pulse before mask and magnetization upload");
    if (ctx->nc == 1) return fail(NMR_ERR_UNSUPPORTED, "pulses need ncomp >= 2");
    if (ctx->nc == 2 && (ax != 0.0 || ay != 0.0))
        return fail(NMR_ERR_UNSUPPORTED, "ncomp 2 takes rotations about z only");
    const nmr_problem& p = ctx->prob;
    nmr::PulseArgs a{};
    // on the layout that holds the state (bricks preferred); pores only, so
    // the other buffer of that layout keeps agreeing at all-solid chunks
    const bool brick = (ctx->home & nmri::kHomeBricks) != 0;
    for (int q = 0; q < 3; q++) a.m[q] = brick ? ctx->mb[ctx->bpar][q] : ctx->m[ctx->cur][q];
    if (brick) {
        a.nby = (p.ny + 3) / 4;
        a.nb8 = (ctx->pitch / 4 + 1) / 2;
    }
    a.mask = ctx->mask;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.pitch = ctx->pitch;
    a.nc = ctx->nc;
    a.flag = flag;
    a.param = param;
    a.ax = ax;
    a.ay = ay;
    a.az = az;
    a.dx = p.dx;
    a.dy = p.dy;
    a.dz = p.dz;
    a.gamma = p.gamma;
    HIP_TRY(nmr::launch_pulse(a, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->home = brick ? nmri::kHomeBricks : nmri::kHomeRows;
    ctx->sig_valid = false;
    return NMR_OK;
}

int nmr_hip_copy_bandwidth(int64_t bytes, int32_t reps, double* ms)
{
    if (!ms) return fail(NMR_ERR_NULL, "nmr_hip_copy_bandwidth: null argument");
    if (bytes < 4096 || reps < 1) return fail(NMR_ERR_INVALID, "bytes %lld, reps %d", (long long)bytes, (int)reps);
    const size_t n = (size_t)bytes / 16;
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cus = std::max(cus, 1);
    float *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NMR_OK;
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == NMR_OK)
            rc = fail(e == hipErrorOutOfMemory ? NMR_ERR_NOMEM : NMR_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
        return rc == NMR_OK;
    };
    double best = 0.0;
    if (check(hipMalloc((void**)&src, n * 16), "hipMalloc") && check(hipMalloc((void**)&dst, n * 16), "hipMalloc") &&
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") &&
        check(hipEventCreate(&e0), "hipEventCreate") && check(hipEventCreate(&e1), "hipEventCreate") &&
        check(hipMemsetAsync(src, 0, n * 16, st), "hipMemsetAsync")) {
        // every copy shape: two untimed copies, then reps timed ones; the fastest shape's time
        for (int kind = 0; kind < nmr::copy_kinds() && rc == NMR_OK; kind++) {
            for (int r = 0; r < 2 && rc == NMR_OK; r++) check(nmr::launch_copy(kind, src, dst, n, cus, st), "copy");
            check(hipEventRecord(e0, st), "hipEventRecord");
            for (int r = 0; r < reps && rc == NMR_OK; r++) check(nmr::launch_copy(kind, src, dst, n, cus, st), "copy");
            float t = 0.f;
            if (check(hipEventRecord(e1, st), "hipEventRecord") && check(hipEventSynchronize(e1), "hipEventSynchronize") &&
                check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime"))
                if (best == 0.0 || t / reps < best) best = t / reps;
        }
    }
    if (rc == NMR_OK) *ms = best;
    if (st) (void)hipStreamSynchronize(st);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int nmr_hip_mfma_rate(int64_t iters, int32_t reps, double* tflops)
{
    if (!tflops) return fail(NMR_ERR_NULL, "nmr_hip_mfma_rate: null argument");
    if (iters < 1 || iters > (1 << 24) || reps < 1)
        return fail(NMR_ERR_INVALID, "iters %lld, reps %d", (long long)iters, (int)reps);
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cus = std::max(cus, 1);
    const int grid = 4 * cus; // four workgroups of four waves per CU: four waves per SIMD
    float* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NMR_OK;
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == NMR_OK)
            rc = fail(e == hipErrorOutOfMemory ? NMR_ERR_NOMEM : NMR_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
        return rc == NMR_OK;
    };
    double best = 0.0;
    if (check(hipMalloc((void**)&out, (size_t)grid * 256 * sizeof(float)), "hipMalloc") &&
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") &&
        check(hipEventCreate(&e0), "hipEventCreate") && check(hipEventCreate(&e1), "hipEventCreate") &&
        check(nmr::launch_mfma_rate((int)iters, grid, out, st), "mfma")) {
        for (int r = 0; r < reps && rc == NMR_OK; r++) {
            float t = 0.f;
            if (check(hipEventRecord(e0, st), "hipEventRecord") &&
                check(nmr::launch_mfma_rate((int)iters, grid, out, st), "mfma") &&
                check(hipEventRecord(e1, st), "hipEventRecord") && check(hipEventSynchronize(e1), "hipEventSynchronize") &&
                check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime"))
                if (best == 0.0 || t < best) best = t;
        }
    }
    if (rc == NMR_OK) {
        const double flop = (double)grid * 4.0 * (double)iters * 8.0 * 16384.0;
        *tflops = best > 0.0 ? flop / (best * 1e-3) / 1e12 : 0.0;
    }
    if (st) (void)hipStreamSynchronize(st);
    if (out) (void)hipFree(out);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int nmr_hip_device_count(int32_t* count)
{
    if (!count) return fail(NMR_ERR_NULL, "nmr_hip_device_count: null argument");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    *count = n;
    return NMR_OK;
}

int nmr_hip_set_device(int32_t device)
{
    HIP_TRY(hipSetDevice(device));
    return NMR_OK;
}

int nmr_hip_last_run_ms(nmr_ctx* ctx, double* ms, int64_t* sweeps)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_last_run_ms: null context");
    if (ms) *ms = ctx->last_ms;
    if (sweeps) *sweeps = ctx->last_sweeps;
    return NMR_OK;
}

int nmr_hip_run(nmr_ctx* ctx, int64_t nsteps, const double* grad_sched, double* signal)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_run: null context");
    if (nsteps < 0) return fail(NMR_ERR_INVALID, "nsteps %lld", (long long)nsteps);
    if (!ctx->has_mask || !ctx->has_mag) return fail(NMR_ERR_STATE, "run before mask and magnetization upload");
    const nmr_problem& p = ctx->prob;
    if (grad_sched) {
        for (int64_t s = 0; s < nsteps; s++) {
            const double* g = grad_sched + 3 * s;
            if (!std::isfinite(g[0]) || !std::isfinite(g[1]) || !std::isfinite(g[2]))
                return fail(NMR_ERR_INVALID, "gradient schedule step %lld is not finite", (long long)s);
            if (ctx->nc == 2 && (g[0] != 0.0 || g[1] != 0.0))
                return fail(NMR_ERR_UNSUPPORTED, "ncomp 2 takes a z gradient only (step %lld)", (long long)s);
            if (ctx->nc == 1 && (p.flags & NMR_F_PRECESSION))
                return fail(NMR_ERR_UNSUPPORTED, "precession needs ncomp >= 2");
        }
    }
    if (ctx->general || ctx->variant == -2) return run_general(ctx, nsteps, grad_sched, signal);
    int rc = refresh_inert(ctx);
    if (rc) return rc;
    Launch L1{}, L2{};
    bool brick = false;
    plain_geometry(ctx, &L1, &L2, &brick);
    if (L1.nwg < 0 || L2.nwg < 0) return fail(NMR_ERR_INVALID, "too many workgroups");
    if (L1.steps == 2 && p.nx > L1.tx)
        return fail(NMR_ERR_INVALID, "variant %d needs nx <= %d", L1.variant, L1.tx);
    const int nc = ctx->nc;
    const bool keep = brick && keep_state(ctx);

    const int64_t rows = nsteps + 1;
    if ((rc = ensure_signal(ctx, (size_t)rows * nc))) return rc;
    if ((rc = ensure_partials(ctx, (size_t)std::min<int64_t>(nsteps, kChunkSteps) * std::max(L1.nwg, L2.nwg) * nc)))
        return rc;
    // the one-step sweep keeps the state in 8x4 bricks from run to run
    if ((rc = brick ? state_bricks(ctx) : nmri::state_rows(ctx))) return rc;
    if ((rc = ensure_adv(ctx, brick))) return rc;

    int w[6];
    signal_window(p, w);
    nmr::SweepArgs a{};
    a.mask = ctx->mask;
    for (int q = 0; q < 3; q++) a.adv[q] = brick ? ctx->advb[q] : ctx->adv[q];
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.pitch = ctx->pitch;
    a.pitch1 = ctx->pitch1;
    a.bx0 = w[0]; a.bx1 = w[1]; a.by0 = w[2]; a.by1 = w[3]; a.bz0 = w[4]; a.bz1 = w[5];
    a.inert = ctx->inert;
    a.nby = (p.ny + 3) / 4;
    nmr::Sweep2Args b{};
    b.mask = ctx->mask;
    for (int q = 0; q < 3; q++) b.adv[q] = ctx->adv[q];
    b.nx = p.nx;
    b.ny = p.ny;
    b.nz = p.nz;
    b.pitch = ctx->pitch;
    b.pitch1 = ctx->pitch1;
    b.bx0 = w[0]; b.bx1 = w[1]; b.by0 = w[2]; b.by1 = w[3]; b.bz0 = w[4]; b.bz1 = w[5];

    auto coef_of = [&](int64_t step) {
        double g[3] = {p.grad[0], p.grad[1], p.grad[2]};
        if (grad_sched) {
            g[0] = grad_sched[3 * step];
            g[1] = grad_sched[3 * step + 1];
            g[2] = grad_sched[3 * step + 2];
        }
        return make_coef(ctx, g);
    };

    const bool have0 = ctx->sig_valid;
    double sig0[6];
    std::copy(ctx->sig_now, ctx->sig_now + 6, sig0);
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if (!have0 && (rc = enqueue_field_sum(ctx, 0))) return rc;
    ctx->home = brick ? nmri::kHomeBricks : nmri::kHomeRows; // the sweeps write this layout only
    int64_t done = 0;
    while (done < nsteps) {
        const int64_t left = nsteps - done;
        const bool two = L1.steps == 2 && left >= 2;
        const Launch& L = two ? L1 : L2;
        int chunk = (int)std::min<int64_t>(kChunkSteps, left);
        if (two) chunk &= ~1;
        if (two) {
            b.nty = L.nty;
            b.ntz = L.ntz;
            b.lz = L.lz;
            b.nblocks = L.nwg;
            for (int s = 0; s < chunk; s += 2) {
                b.c1 = coef_of(done + s);
                b.c2 = coef_of(done + s + 1);
                for (int q = 0; q < 3; q++) {
                    b.min[q] = ctx->m[ctx->cur][q];
                    b.mout[q] = ctx->m[ctx->cur ^ 1][q];
                }
                b.partials1 = ctx->partials + (size_t)s * L.nwg * nc;
                b.partials2 = ctx->partials + (size_t)(s + 1) * L.nwg * nc;
                HIP_TRY(nmr::launch_sweep2(nc, ctx->vel, ctx->exact, L.lx, b, ctx->stream));
                ctx->cur ^= 1;
            }
        } else {
            a.ntx = L.ntx;
            a.nty = L.nty;
            a.ntz = L.ntz;
            a.lz = L.lz;
            a.nblocks = L.nwg;
            for (int s = 0; s < chunk; s++) {
                a.c = coef_of(done + s);
                for (int q = 0; q < 3; q++) {
                    a.min[q] = brick ? ctx->mb[ctx->bpar][q] : ctx->m[ctx->cur][q];
                    a.mout[q] = brick ? ctx->mb[ctx->bpar ^ 1][q] : ctx->m[ctx->cur ^ 1][q];
                }
                a.partials = ctx->partials + (size_t)s * L.nwg * nc;
                HIP_TRY(nmr::launch_sweep(nc, ctx->vel, ctx->exact, L.variant, a, ctx->stream, brick, keep));
                if (brick) ctx->bpar ^= 1;
                else ctx->cur ^= 1;
            }
        }
        HIP_TRY(nmr::launch_reduce_partials(ctx->partials, chunk, L.nwg, nc, ctx->dsignal + (size_t)(done + 1) * nc,
                                            ctx->stream));
        done += chunk;
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    if ((rc = finish_signal(ctx, nsteps, have0, sig0, signal))) return rc;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    ctx->last_sweeps = L1.steps == 2 ? nsteps / 2 + (nsteps & 1) : nsteps;
    return NMR_OK;
}

int nmr_hip_run_static(nmr_ctx* ctx, int64_t nsteps, const double* grad_sched, double* signal)
{
    if (!ctx) return fail(NMR_ERR_NULL, "nmr_hip_run_static: null context");
    if (nsteps < 0) return fail(NMR_ERR_INVALID, "nsteps %lld", (long long)nsteps);
    if (!ctx->has_mask || !ctx->has_mag) return fail(NMR_ERR_STATE, "run before mask and magnetization upload");
    if (!ctx->scratch) return fail(NMR_ERR_STATE, "no scratch address on this device");
    const nmr_problem& p = ctx->prob;
    const size_t ns = (size_t)ctx->nsig();
    // partial sums per step of the sweep this run takes (none for the
    // per-voxel kernel, whose signal is a field sum after each step)
    size_t per_step = 0;
    if (ctx->general) {
        if (ctx->variant != -2) per_step = (size_t)general_geometry(ctx, ctx->brick_on && bricks_fit(ctx)).nwg * ns;
    } else {
        Launch L1{}, L2{};
        bool brick = false;
        plain_geometry(ctx, &L1, &L2, &brick);
        if (L1.nwg < 0 || L2.nwg < 0) return fail(NMR_ERR_INVALID, "too many workgroups");
        per_step = (size_t)std::max(L1.nwg, L2.nwg) * ns;
    }
    nmr::FieldSumArgs f{};
    int w[6];
    signal_window(p, w);
    f.bx0 = w[0]; f.bx1 = w[1]; f.by0 = w[2]; f.by1 = w[3]; f.bz0 = w[4]; f.bz1 = w[5];
    const size_t fsum = (size_t)nmr::field_sum_blocks(f) * ns;
    // steps per sub-run: their partials (or one field sum's) and their
    // signal rows must fit the scratch
    int64_t C = std::min<int64_t>(kChunkSteps, (int64_t)((kScratchDoubles - ns) / (per_step + ns)));
    while (C >= 1 && std::max((size_t)C * per_step, fsum) + (size_t)(C + 1) * ns > kScratchDoubles) C--;
    if (C < 1)
        return fail(NMR_ERR_INVALID,
                    "grid too large for the fixed scratch: one step needs %zu doubles of partial sums and a field sum "
                    "%zu, the scratch holds %zu",
                    per_step + 2 * ns, fsum + 2 * ns, kScratchDoubles);
    const size_t P = std::max((size_t)C * per_step, fsum);
    double* const keep_p = ctx->partials;
    double* const keep_s = ctx->dsignal;
    const size_t keep_pc = ctx->partials_cap, keep_sc = ctx->dsignal_cap;
    ctx->partials = ctx->scratch;
    ctx->partials_cap = P;
    ctx->dsignal = ctx->scratch + P;
    ctx->dsignal_cap = (size_t)(C + 1) * ns;
    ctx->no_alloc = true;
    double ms = 0.0;
    int64_t sweeps = 0, done = 0;
    int rc = NMR_OK;
    do { // sub-runs of at most C steps; each waits for its stream before the next reuses the scratch
        const int64_t n = std::min<int64_t>(C, nsteps - done);
        rc = nmr_hip_run(ctx, n, grad_sched ? grad_sched + 3 * done : nullptr, signal ? signal + done * ns : nullptr);
        ms += ctx->last_ms;
        sweeps += ctx->last_sweeps;
        done += n;
    } while (rc == NMR_OK && done < nsteps);
    ctx->partials = keep_p;
    ctx->partials_cap = keep_pc;
    ctx->dsignal = keep_s;
    ctx->dsignal_cap = keep_sc;
    ctx->no_alloc = false;
    ctx->last_ms = ms;
    ctx->last_sweeps = sweeps;
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batched solves: one persistent launch per step over all instances.

namespace {

std::mutex g_tune_mu;
struct Tuned { int variant, lz, persistent; std::array<double, 7> cost; };
std::map<std::vector<int>, Tuned> g_tune_cache; // shapes -> fastest launch

void batch_free(nmr_batch* b)
{
    for (auto& I : b->inst) {
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 3; q++)
                if (I.m[p][q]) (void)hipFree(I.m[p][q]);
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 3; q++)
                if (I.mb[p][q]) (void)hipFree(I.mb[p][q]);
        for (int q = 0; q < 3; q++)
            for (float* f : {I.adv[q], I.advb[q]})
                if (f) (void)hipFree(f);
        if (I.mask) (void)hipFree(I.mask);
        if (I.vx) (void)hipFree(I.vx);
        if (I.vy) (void)hipFree(I.vy);
        if (I.vz) (void)hipFree(I.vz);
    }
    for (void* p : {(void*)b->d_inst, (void*)b->d_tiles, (void*)b->d_off, (void*)b->d_ids, (void*)b->d_step,
                    (void*)b->d_part, (void*)b->d_signal, (void*)b->d_fpart, (void*)b->d_flag})
        if (p) (void)hipFree(p);
    if (b->h_signal) (void)hipHostFree(b->h_signal);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// A cleared device buffer.  The clear runs on the batch's stream and is
// waited for: that stream is non-blocking, so a clear on the null stream
// (hipMemset) would not be ordered before the batch's own work on the
// buffer and could land after it.
template <typename T>
int balloc(hipStream_t s, T** p, size_t count)
{
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_TRY(hipMalloc((void**)p, bytes));
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NMR_OK;
}

constexpr int kMixedVariant = -3; // every instance takes its own tile shape

// Whether every brick array of every instance stays inside the sweep's
// 32-bit byte offsets.
bool batch_bricks_fit(const nmr_batch* b)
{
    for (size_t i = 0; i < b->inst.size(); i++) {
        const auto& s = b->spec[i];
        const size_t big = nmr::brick_floats(b->inst[i].pitch / 4, s.ny, s.nz);
        if (big * 4 >= ((size_t)1 << 32) - 256) return false;
    }
    return true;
}

// Brick buffers of instance i (allocated once) with its velocity sums converted.
int batch_inst_bricks(nmr_batch* b, int i)
{
    const auto& s = b->spec[i];
    auto& I = b->inst[i];
    const int nbx = I.pitch / 4;
    int rc;
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < b->nc; q++)
            if (!I.mb[p][q] && (rc = balloc(b->stream, &I.mb[p][q], nmr::brick_floats(nbx, s.ny, s.nz)))) return rc;
    if (b->vel && !I.adv_bricks) {
        for (int q = 0; q < 3; q++) {
            if (!I.advb[q] && (rc = balloc(b->stream, &I.advb[q], nmr::brick_floats(nbx, s.ny, s.nz)))) return rc;
            HIP_TRY(nmr::launch_brick(I.adv[q], I.advb[q], nbx, s.ny, s.nz, true, b->stream));
        }
        I.adv_bricks = true;
    }
    return NMR_OK;
}

// Instance i's current row state into both of its brick buffers.
int batch_inst_to_bricks(nmr_batch* b, int i)
{
    int rc = batch_inst_bricks(b, i);
    if (rc) return rc;
    const auto& s = b->spec[i];
    auto& I = b->inst[i];
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < b->nc; q++)
            HIP_TRY(nmr::launch_brick(I.m[b->parity][q], I.mb[p][q], I.pitch / 4, s.ny, s.nz, true, b->stream));
    return NMR_OK;
}

// Every instance's state into the bricks unless they hold it already.
int batch_bricks(nmr_batch* b)
{
    if (b->home & nmri::kHomeBricks) return NMR_OK;
    for (int i = 0; i < b->ninst; i++) {
        int rc = batch_inst_to_bricks(b, i);
        if (rc) return rc;
    }
    b->home |= nmri::kHomeBricks;
    return NMR_OK;
}

// One-step tile shape for an nx x ny instance under variant -3: the shape
// that pads the instance least, weighed by each shape's time per padded
// voxel (cost[v]: measured on the batch by the autotuner; before that, the
// relative costs on a 256^3 grid, tile 256x4 fastest, 64x32 slowest).
int mixed_variant(int nx, int ny, const double* cost)
{
    int best = 0;
    double bc = 1e300;
    for (int v = 0; v < 7; v++) {
        int lx, rpt, tx, ty, st;
        if (!nmr::variant_shape(v, &lx, &rpt, &tx, &ty, &st) || st != 1) continue;
        const double area = (double)((nx + tx - 1) / tx * tx) * (double)((ny + ty - 1) / ty * ty);
        const double c = area * cost[v];
        if (c < bc) {
            bc = c;
            best = v;
        }
    }
    return best;
}

// Tile list, descriptors and grid for (variant, lz, grid kind); variant -3
// gives each instance its own tile shape and groups the tiles by shape.
int batch_prepare(nmr_batch* b, int variant, int lz, int persistent, bool brick)
{
    if (variant == b->cur_variant && lz == b->cur_lz && persistent == b->cur_persistent && brick == b->cur_brick)
        return NMR_OK;
    const bool mixed = variant == kMixedVariant;
    int lx, rpt, tx, ty, st;
    if (!mixed && (!nmr::variant_shape(variant, &lx, &rpt, &tx, &ty, &st) || st != 1))
        return fail(NMR_ERR_INVALID, "batch variant %d is not a one-step tile", variant);
    const nmr_problem& p = b->prob;
    std::vector<nmr::BatchInst> hi(b->ninst);
    struct T { int inst, t, variant; long cost; };
    std::vector<T> tiles;
    for (int i = 0; i < b->ninst; i++) {
        const nmr_batch_instance& s = b->spec[i];
        auto& I = b->inst[i];
        nmr::BatchInst& d = hi[i];
        const int vi = mixed ? mixed_variant(s.nx, s.ny, b->tile_cost.data()) : variant;
        nmr::variant_shape(vi, &lx, &rpt, &tx, &ty, &st);
        for (int q = 0; q < 3; q++) {
            d.m[0][q] = brick ? I.mb[0][q] : I.m[0][q];
            d.m[1][q] = brick ? I.mb[1][q] : I.m[1][q];
        }
        d.mask = I.mask;
        for (int q = 0; q < 3; q++) d.adv[q] = brick ? I.advb[q] : I.adv[q];
        d.nby = (s.ny + 3) / 4;
        d.nx = s.nx;
        d.ny = s.ny;
        d.nz = s.nz;
        d.pitch = I.pitch;
        d.pitch1 = I.pitch1;
        d.lz = std::min(std::max(lz, 1), s.nz);
        d.ntx = (s.nx + tx - 1) / tx;
        d.nty = (s.ny + ty - 1) / ty;
        d.ntz = (s.nz + d.lz - 1) / d.lz;
        d.bx0 = p.buffer[0]; d.bx1 = s.nx - p.buffer[0];
        d.by0 = p.buffer[1]; d.by1 = s.ny - p.buffer[1];
        d.bz0 = p.buffer[2]; d.bz1 = s.nz - p.buffer[2];
        d.inert = I.inert;
        nmr_problem q = p;
        q.rho1 = s.rho1;
        q.rho2 = s.rho2;
        q.flags &= ~NMR_F_PRECESSION; // the gradient is chosen per step in the kernel
        nmr_ctx tmp;
        tmp.prob = q;
        const double g0[3] = {0.0, 0.0, 0.0};
        d.c = make_coef(&tmp, g0);
        d.nd = (int)std::lround(s.delta / p.dt);
        d.nD = (int)std::lround(s.big_delta / p.dt);
        // same rounding as make_coef for a gradient (0, 0, G)
        d.bgz = (p.flags & NMR_F_PRECESSION) ? (float)(p.dt * p.gamma * s.grad_amp * p.dz) : 0.f;
        const int n = d.ntx * d.nty * d.ntz;
        for (int t = 0; t < n; t++) {
            const int tz = t / (d.ntx * d.nty);
            const long planes = std::min(d.lz, s.nz - tz * d.lz);
            tiles.push_back({i, t, vi, planes * (long)tx * ty});
        }
    }
    // by tile shape, then longest tiles first (stable: ties keep
    // instance/tile order)
    std::stable_sort(tiles.begin(), tiles.end(), [](const T& a, const T& c) {
        return a.variant != c.variant ? a.variant < c.variant : a.cost > c.cost;
    });
    b->groups.clear();
    for (size_t k = 0; k < tiles.size(); k++) {
        if (b->groups.empty() || b->groups.back().variant != tiles[k].variant)
            b->groups.push_back({tiles[k].variant, (int)k, 0, 0});
        b->groups.back().count += 1;
    }
    // XCD runs: the dispatcher hands workgroup w to XCD w % 8, so list
    // position s0 + 8i + j takes sorted tile s0 + jR + i.  Each XCD then
    // works through R consecutive sorted tiles (x-, then y-neighbours of one
    // instance, whose halo lines stay in that XCD's L2) while the cost order
    // holds from one run of 8R tiles to the next.  256 instances of edge
    // 32-128, tile 64x16, z-chunk 32: 1052.5 us per step in sorted order,
    // 1039.2 / 1023.2 / 1017.3 / 1016.7 / 1024.6 us at R = 2 / 4 / 8 / 16 / 32.
    {
        constexpr int kXcdRun = 16;
        const std::vector<T> src(tiles);
        for (const auto& gr : b->groups)
            for (int s0 = gr.first, end = gr.first + gr.count; end - s0 >= 16;) {
                const int R = std::min(kXcdRun, (end - s0) / 8);
                for (int i = 0; i < R; i++)
                    for (int j = 0; j < 8; j++) tiles[s0 + 8 * i + j] = src[s0 + j * R + i];
                s0 += 8 * R;
            }
    }
    std::vector<int2> ht(tiles.size());
    std::vector<std::vector<int>> per(b->ninst);
    for (size_t k = 0; k < tiles.size(); k++) {
        ht[k] = make_int2(tiles[k].inst, tiles[k].t);
        per[tiles[k].inst].push_back((int)k);
    }
    std::vector<int> off(b->ninst + 1, 0), ids;
    for (int i = 0; i < b->ninst; i++) {
        // fixed summation order: by tile index
        std::sort(per[i].begin(), per[i].end(), [&](int u, int v) { return ht[u].y < ht[v].y; });
        off[i + 1] = off[i] + (int)per[i].size();
        ids.insert(ids.end(), per[i].begin(), per[i].end());
    }
    for (void* q : {(void*)b->d_inst, (void*)b->d_tiles, (void*)b->d_off, (void*)b->d_ids})
        if (q) HIP_TRY(hipFree(q));
    b->d_inst = nullptr;
    b->d_tiles = nullptr;
    b->d_off = nullptr;
    b->d_ids = nullptr;
    int rc;
    if ((rc = balloc(b->stream, &b->d_inst, hi.size())) || (rc = balloc(b->stream, &b->d_tiles, ht.size())) ||
        (rc = balloc(b->stream, &b->d_off, off.size())) || (rc = balloc(b->stream, &b->d_ids, ids.size())))
        return rc;
    HIP_TRY(hipMemcpy(b->d_inst, hi.data(), hi.size() * sizeof(nmr::BatchInst), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_tiles, ht.data(), ht.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice));
    b->ntiles = (int)tiles.size();
    for (auto& gr : b->groups) {
        if (persistent) {
            const int res = nmr::batch_resident_blocks(b->nc, b->vel, b->exact, gr.variant, brick);
            if (res <= 0) return fail(NMR_ERR_DEVICE, "occupancy query failed");
            gr.grid = std::min(res, gr.count);
        } else {
            gr.grid = gr.count;
        }
    }
    b->grid = b->groups.empty() ? 0 : b->groups[0].grid;
    b->cur_variant = variant;
    b->cur_lz = lz;
    b->cur_persistent = persistent;
    b->cur_brick = brick;
    return NMR_OK;
}

int batch_partials(nmr_batch* b, size_t steps)
{
    const size_t need = steps * (size_t)b->ntiles * b->nc;
    if (need <= b->part_cap) return NMR_OK;
    if (b->d_part) HIP_TRY(hipFree(b->d_part));
    b->d_part = nullptr;
    int rc = balloc(b->stream, &b->d_part, need);
    if (rc) return rc;
    b->part_cap = need;
    return NMR_OK;
}

int batch_launch_step(nmr_batch* b, int local, int parity, double* part)
{
    for (const auto& gr : b->groups) {
        nmr::BatchArgs a{};
        a.inst = b->d_inst;
        a.tiles = b->d_tiles + gr.first;
        a.ntiles = gr.count;
        a.parity = parity;
        a.local_step = local;
        a.step_base = b->d_step;
        a.partials = part + (size_t)gr.first * b->nc;
        HIP_TRY(nmr::launch_batch(b->nc, b->vel, b->exact, gr.variant, gr.grid, a, b->stream, b->cur_brick));
    }
    return NMR_OK;
}

void batch_auto(const nmr_batch* b, int* variant, int* lz)
{
    int mx = 0;
    for (const auto& s : b->spec) mx = std::max(mx, s.nx);
    *variant = mx > 64 ? 2 : (mx > 32 ? 4 : 6);
    *lz = 8;
}

std::vector<int> batch_key(const nmr_batch* b)
{
    std::vector<int> k = {b->nc, (int)b->vel, (int)b->exact};
    for (const auto& s : b->spec) {
        k.push_back(s.nx);
        k.push_back(s.ny);
        k.push_back(s.nz);
    }
    return k;
}

} // namespace

extern "C" {

int nmr_hip_batch_create(const nmr_problem* common, int32_t ninst, const nmr_batch_instance* inst, nmr_batch** out)
{
    if (!common || !inst || !out) return fail(NMR_ERR_NULL, "nmr_hip_batch_create: null argument");
    *out = nullptr;
    if (ninst < 1 || ninst > 65535) return fail(NMR_ERR_INVALID, "ninst %d (1..65535)", (int)ninst);
    nmr_problem p = *common;
    if (p.ncomp < 1 || p.ncomp > 3) return fail(NMR_ERR_INVALID, "ncomp %d", p.ncomp);
    if (p.scheme != NMR_SCHEME_UPWIND) return fail(NMR_ERR_UNSUPPORTED, "batched solves use the upwind scheme");
    for (int a = 0; a < 3; a++)
        if (p.bc[a] != NMR_BC_NEUMANN) return fail(NMR_ERR_UNSUPPORTED, "batched solves use zero-derivative boundaries");
    if (p.flags & ~(NMR_F_DIFFUSION | NMR_F_VELOCITY | NMR_F_PRECESSION)) return fail(NMR_ERR_INVALID, "flags");
    if (!(p.dx > 0) || !(p.dy > 0) || !(p.dz > 0) || !(p.dt > 0)) return fail(NMR_ERR_INVALID, "dx, dy, dz, dt");
    if (p.ncomp < 3 && (p.b0_dir[0] != 0.0 || p.b0_dir[1] != 0.0 || p.b0_dir[2] == 0.0))
        return fail(NMR_ERR_UNSUPPORTED, "ncomp < 3 needs the main field along z");
    if (p.ncomp == 1 && (p.flags & NMR_F_PRECESSION)) return fail(NMR_ERR_UNSUPPORTED, "precession needs ncomp >= 2");
    if (p.phase_mode == NMR_PHASE_EXACT && p.ncomp != 2) return fail(NMR_ERR_UNSUPPORTED, "exact phase needs ncomp 2");
    for (int i = 0; i < ninst; i++) {
        const nmr_batch_instance& s = inst[i];
        if (s.nx < 1 || s.ny < 1 || s.nz < 1 || (long)s.nx * s.ny * s.nz > (1L << 30))
            return fail(NMR_ERR_INVALID, "instance %d grid %d x %d x %d", i, s.nx, s.ny, s.nz);
        // the batched sweep addresses every array with 32-bit byte offsets
        if (nmri::biggest_array_bytes(s.nx, s.ny, s.nz) >= ((size_t)1 << 32) - 64)
            return fail(NMR_ERR_UNSUPPORTED, "instance %d: arrays above the 4 GiB addressing of the sweep", i);
        if (!(s.delta >= 0) || !(s.big_delta >= 0) || !std::isfinite(s.grad_amp))
            return fail(NMR_ERR_INVALID, "instance %d gradient schedule", i);
        for (int a = 0; a < 3; a++)
            if (p.buffer[a] < 0) return fail(NMR_ERR_INVALID, "buffer");
    }
    nmr_batch* b = new (std::nothrow) nmr_batch();
    if (!b) return fail(NMR_ERR_NOMEM, "host allocation");
    auto bail = [&](int rc) {
        batch_free(b);
        delete b;
        return rc;
    };
    b->prob = p;
    b->ninst = ninst;
    b->nc = p.ncomp;
    b->vel = (p.flags & NMR_F_VELOCITY) != 0;
    b->exact = p.phase_mode == NMR_PHASE_EXACT;
    b->spec.assign(inst, inst + ninst);
    b->inst.resize(ninst);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return bail(fail(NMR_ERR_DEVICE, "stream/event creation failed"));
    int rc;
    size_t maxrows = 1; // y*z rows of the largest instance (field-sum partials)
    for (int i = 0; i < ninst; i++) {
        const nmr_batch_instance& s = inst[i];
        auto& I = b->inst[i];
        I.pitch = round_up4(s.nx);
        I.pitch1 = round_up4(s.nx + 1);
        I.n_m = (size_t)s.nz * s.ny * I.pitch;
        maxrows = std::max(maxrows, (size_t)s.ny * (size_t)s.nz);
        for (int par = 0; par < 2; par++)
            for (int q = 0; q < b->nc; q++)
                if ((rc = balloc(b->stream, &I.m[par][q], I.n_m))) return bail(rc);
        if ((rc = balloc(b->stream, &I.mask, I.n_m))) return bail(rc);
        if (b->vel) {
            I.n_vx = (size_t)s.nz * s.ny * I.pitch1;
            I.n_vy = (size_t)s.nz * (s.ny + 1) * I.pitch;
            I.n_vz = (size_t)(s.nz + 1) * s.ny * I.pitch;
            hipStream_t st = b->stream;
            if ((rc = balloc(st, &I.vx, I.n_vx)) || (rc = balloc(st, &I.vy, I.n_vy)) || (rc = balloc(st, &I.vz, I.n_vz)) ||
                (rc = balloc(st, &I.adv[0], I.n_m)) || (rc = balloc(st, &I.adv[1], I.n_m)) ||
                (rc = balloc(st, &I.adv[2], I.n_m)))
                return bail(rc);
        }
    }
    if ((rc = balloc(b->stream, &b->d_step, 1)) || (rc = balloc(b->stream, &b->d_fpart, (maxrows / 16 + 2) * 3)) ||
        (rc = balloc(b->stream, &b->d_flag, 16)))
        return bail(rc);
    *out = b;
    return NMR_OK;
}

int nmr_hip_batch_destroy(nmr_batch* b)
{
    if (!b) return NMR_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
    return NMR_OK;
}

int nmr_hip_batch_upload(nmr_batch* b, int32_t i, const uint8_t* mask, const double* vx, const double* vy,
                         const double* vz, const double* mx, const double* my, const double* mz)
{
    if (!b || !mask) return fail(NMR_ERR_NULL, "nmr_hip_batch_upload: null argument");
    if (i < 0 || i >= b->ninst) return fail(NMR_ERR_INVALID, "instance %d", (int)i);
    const double* src[3] = {mx, my, mz};
    for (int q = 0; q < b->nc; q++)
        if (!src[q]) return fail(NMR_ERR_NULL, "magnetization component %d is NULL", q);
    const nmr_batch_instance& s = b->spec[i];
    auto& I = b->inst[i];
    const int P = I.pitch, P1 = I.pitch1;
    int rc = NMR_OK;
    {
        std::vector<uint8_t> h(I.n_m, 0);
        int64_t np = 0;
        for (int k = 0; k < s.nz; k++)
            for (int j = 0; j < s.ny; j++)
                for (int x = 0; x < s.nx; x++) {
                    const uint8_t v = mask[(size_t)x + (size_t)s.nx * ((size_t)j + (size_t)s.ny * k)];
                    if (v > 1) return fail(NMR_ERR_INVALID, "mask value %d (want 0/1)", (int)v);
                    h[(size_t)k * s.ny * P + (size_t)j * P + x] = v;
                    np += v;
                }
        I.npore = np;
        HIP_TRY(hipMemcpy(I.mask, h.data(), h.size(), hipMemcpyHostToDevice));
    }
    if (b->vel) {
        std::vector<float> h(I.n_vx, 0.f);
        if (vx)
            for (int k = 0; k < s.nz; k++)
                for (int j = 0; j < s.ny; j++)
                    for (int x = 0; x <= s.nx; x++)
                        h[(size_t)k * s.ny * P1 + (size_t)j * P1 + x] =
                            (float)vx[(size_t)x + (size_t)(s.nx + 1) * ((size_t)j + (size_t)s.ny * k)];
        HIP_TRY(hipMemcpy(I.vx, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        h.assign(I.n_vy, 0.f);
        if (vy)
            for (int k = 0; k < s.nz; k++)
                for (int j = 0; j <= s.ny; j++)
                    for (int x = 0; x < s.nx; x++)
                        h[(size_t)k * (s.ny + 1) * P + (size_t)j * P + x] =
                            (float)vy[(size_t)x + (size_t)s.nx * ((size_t)j + (size_t)(s.ny + 1) * k)];
        HIP_TRY(hipMemcpy(I.vy, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        h.assign(I.n_vz, 0.f);
        if (vz)
            for (int k = 0; k <= s.nz; k++)
                for (int j = 0; j < s.ny; j++)
                    for (int x = 0; x < s.nx; x++)
                        h[(size_t)k * s.ny * P + (size_t)j * P + x] =
                            (float)vz[(size_t)x + (size_t)s.nx * ((size_t)j + (size_t)s.ny * k)];
        HIP_TRY(hipMemcpy(I.vz, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    std::vector<float> h(I.n_m, 0.f);
    for (int q = 0; q < b->nc; q++) {
        for (int k = 0; k < s.nz; k++)
            for (int j = 0; j < s.ny; j++)
                for (int x = 0; x < s.nx; x++)
                    h[(size_t)k * s.ny * P + (size_t)j * P + x] =
                        (float)src[q][(size_t)x + (size_t)s.nx * ((size_t)j + (size_t)s.ny * k)];
        for (int par = 0; par < 2; par++) HIP_TRY(hipMemcpy(I.m[par][q], h.data(), h.size() * 4, hipMemcpyHostToDevice));
    }
    // inertness of this instance (no velocity on faces touching solid)
    if (!b->vel) {
        I.inert = 1;
    } else {
        HIP_TRY(hipMemsetAsync(b->d_flag, 0, sizeof(int), b->stream));
        nmr::InertArgs a{};
        a.mask = I.mask;
        a.vx = I.vx;
        a.vy = I.vy;
        a.vz = I.vz;
        a.flag = b->d_flag;
        a.nx = s.nx;
        a.ny = s.ny;
        a.nz = s.nz;
        a.pitch = P;
        a.pitch1 = P1;
        HIP_TRY(nmr::launch_inert_check(a, b->stream));
        nmr::AdvArgs v{};
        v.vx = I.vx;
        v.vy = I.vy;
        v.vz = I.vz;
        for (int q = 0; q < 3; q++) v.out[q] = I.adv[q];
        v.nx = s.nx;
        v.ny = s.ny;
        v.nz = s.nz;
        v.pitch = P;
        v.pitch1 = P1;
        HIP_TRY(nmr::launch_adv(v, b->stream));
        int f = 0;
        HIP_TRY(hipMemcpyAsync(&f, b->d_flag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        I.inert = f == 0;
    }
    I.uploaded = true;
    I.adv_bricks = false;
    // the rows hold the new state (both buffers); so do the bricks if they
    // are the home of the other instances
    if ((b->home & nmri::kHomeBricks) && (rc = batch_inst_to_bricks(b, i))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if ((size_t)i < b->sig_ok.size()) b->sig_ok[i] = 0;
    b->cur_variant = -2; // descriptors carry the inert flag: rebuild
    return NMR_OK;
}

int nmr_hip_batch_download(nmr_batch* b, int32_t i, double* mx, double* my, double* mz)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_download: null batch");
    if (i < 0 || i >= b->ninst) return fail(NMR_ERR_INVALID, "instance %d", (int)i);
    double* dst[3] = {mx, my, mz};
    const nmr_batch_instance& s = b->spec[i];
    auto& I = b->inst[i];
    if (!(b->home & nmri::kHomeRows)) // the instance's current buffer from the bricks
        for (int q = 0; q < b->nc; q++)
            HIP_TRY(nmr::launch_brick(I.mb[b->parity][q], I.m[b->parity][q], I.pitch / 4, s.ny, s.nz, false,
                                      b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<float> h(I.n_m);
    for (int q = 0; q < b->nc; q++) {
        if (!dst[q]) continue;
        HIP_TRY(hipMemcpy(h.data(), I.m[b->parity][q], h.size() * 4, hipMemcpyDeviceToHost));
        for (int k = 0; k < s.nz; k++)
            for (int j = 0; j < s.ny; j++)
                for (int x = 0; x < s.nx; x++)
                    dst[q][(size_t)x + (size_t)s.nx * ((size_t)j + (size_t)s.ny * k)] =
                        h[(size_t)k * s.ny * I.pitch + (size_t)j * I.pitch + x];
    }
    return NMR_OK;
}

int nmr_hip_batch_set_layout(nmr_batch* b, int32_t brick)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_set_layout: null batch");
    if (brick != 0 && brick != 1) return fail(NMR_ERR_INVALID, "layout %d", (int)brick);
    b->brick_on = brick != 0;
    return NMR_OK;
}

int nmr_hip_batch_set_tile(nmr_batch* b, int32_t variant, int32_t lz, int32_t graph_steps, int32_t persistent)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_set_tile: null batch");
    int lx, rpt, tx, ty, st;
    if (variant != -1 && variant != kMixedVariant &&
        (!nmr::variant_shape(variant, &lx, &rpt, &tx, &ty, &st) || st != 1 || variant > 6))
        return fail(NMR_ERR_INVALID, "batch variant %d (one-step tiles 0..6)", (int)variant);
    if (lz < 0 || graph_steps < 0 || (graph_steps & 1) || (persistent != 0 && persistent != 1))
        return fail(NMR_ERR_INVALID, "lz %d / graph_steps %d / persistent %d", lz, graph_steps, persistent);
    b->variant = variant;
    b->lz = lz;
    b->graph_steps = graph_steps;
    b->persistent = persistent;
    return NMR_OK;
}

int nmr_hip_batch_autotune(nmr_batch* b, int32_t trial_steps, int32_t* variant, int32_t* lz, int32_t* persistent)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_autotune: null batch");
    for (const auto& I : b->inst)
        if (!I.uploaded) return fail(NMR_ERR_STATE, "autotune before every instance is uploaded");
    const std::vector<int> key = batch_key(b);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            b->variant = it->second.variant;
            b->lz = it->second.lz;
            b->persistent = it->second.persistent;
            b->tile_cost = it->second.cost;
            b->cur_variant = -2; // the per-instance shapes may have changed
            if (variant) *variant = b->variant;
            if (lz) *lz = b->lz;
            if (persistent) *persistent = b->persistent;
            return NMR_OK;
        }
    }
    // trials on the layout the runs take
    const bool brick = b->brick_on && batch_bricks_fit(b);
    {
        int rc = brick ? batch_bricks(b) : nmri::batch_rows(b);
        if (rc) return rc;
    }
    const int trials = std::max(1, (int)trial_steps);
    // GPU time per launch of (variant, lz, grid kind) over n launches after
    // one untimed: every trial reads the current buffer and writes the
    // other one, so the state is unchanged
    auto time_cfg = [&](int v, int l, int per, int n, double* ms_out) -> int {
        int rc = batch_prepare(b, v, l, per, brick);
        if (rc) return rc;
        if ((rc = batch_partials(b, 1))) return rc;
        if ((rc = batch_launch_step(b, 0, b->parity, b->d_part))) return rc;
        HIP_TRY(hipEventRecord(b->ev0, b->stream));
        for (int r = 0; r < n; r++)
            if ((rc = batch_launch_step(b, 0, b->parity, b->d_part))) return rc;
        HIP_TRY(hipEventRecord(b->ev1, b->stream));
        HIP_TRY(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *ms_out = (double)ms / n;
        return NMR_OK;
    };
    struct Trial { int v, l, per; double ms; };
    std::vector<Trial> done;
    // every one-step tile shape on the whole batch, then per-instance shapes
    // (variant -3) weighed by what the uniform trials measured per padded voxel
    std::vector<int> order;
    for (int v = 0; v < nmr::num_variants(); v++) {
        int lx, rpt, tx, ty, st;
        if (nmr::variant_shape(v, &lx, &rpt, &tx, &ty, &st) && st == 1 && v < 7) order.push_back(v);
    }
    order.push_back(kMixedVariant);
    std::array<double, 7> vbest;
    vbest.fill(1e300);
    for (int v : order) {
        if (v == kMixedVariant) {
            // time per padded voxel of each shape on this batch
            bool all = true;
            for (int u = 0; u < 7; u++) all = all && vbest[u] < 1e299;
            if (all) {
                for (int u = 0; u < 7; u++) {
                    int lx, rpt, tx, ty, st;
                    nmr::variant_shape(u, &lx, &rpt, &tx, &ty, &st);
                    double area = 0.0;
                    for (const auto& sp : b->spec)
                        area += (double)((sp.nx + tx - 1) / tx * tx) * (double)((sp.ny + ty - 1) / ty * ty) * sp.nz;
                    b->tile_cost[u] = vbest[u] / area;
                }
                b->cur_variant = -2;
            }
        }
        for (int cfg = 0; cfg < 10; cfg++) {
            const int l = 4 << (cfg >> 1), per = cfg & 1; // lz 4..64, grid kind
            double ms = 0.0;
            int rc = time_cfg(v, l, per, trials, &ms);
            if (rc) return rc;
            if (v >= 0 && v < 7) vbest[v] = std::min(vbest[v], ms);
            done.push_back({v, l, per, ms});
        }
    }
    // The first trials run while the clocks still rise, which favours the
    // later candidates: the four fastest are timed again, over three times
    // the steps, on the warm device, and the fastest of that pass is kept.
    std::stable_sort(done.begin(), done.end(), [](const Trial& x, const Trial& y) { return x.ms < y.ms; });
    double best = 1e300;
    int bv = 0, bl = 8, bp = 1;
    for (size_t k = 0; k < std::min<size_t>(4, done.size()); k++) {
        double ms = 0.0;
        int rc = time_cfg(done[k].v, done[k].l, done[k].per, 3 * trials, &ms);
        if (rc) return rc;
        if (ms < best) {
            best = ms;
            bv = done[k].v;
            bl = done[k].l;
            bp = done[k].per;
        }
    }
    b->variant = bv;
    b->lz = bl;
    b->persistent = bp;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = {bv, bl, bp, b->tile_cost};
    }
    if (variant) *variant = bv;
    if (lz) *lz = bl;
    if (persistent) *persistent = bp;
    return NMR_OK;
}

int nmr_hip_batch_last_run_ms(nmr_batch* b, double* ms, int64_t* voxel_steps)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_last_run_ms: null batch");
    if (ms) *ms = b->last_ms;
    if (voxel_steps) *voxel_steps = b->last_voxel_steps;
    return NMR_OK;
}

int nmr_hip_batch_run(nmr_batch* b, int64_t nsteps, double* signal)
{
    if (!b) return fail(NMR_ERR_NULL, "nmr_hip_batch_run: null batch");
    if (nsteps < 0 || b->steps_done + nsteps > (int64_t)1 << 30) return fail(NMR_ERR_INVALID, "nsteps");
    for (const auto& I : b->inst)
        if (!I.uploaded) return fail(NMR_ERR_STATE, "run before every instance is uploaded");
    int v = b->variant, l = b->lz;
    {
        int av, al;
        batch_auto(b, &av, &al);
        if (v < 0) v = av;
        if (l <= 0) l = al;
    }
    // the state stays in bricks from run to run when the layout allows
    const bool brick = b->brick_on && batch_bricks_fit(b);
    int rc = brick ? batch_bricks(b) : nmri::batch_rows(b);
    if (rc) return rc;
    if ((rc = batch_prepare(b, v, l, b->persistent, brick))) return rc;
    const int nc = b->nc;
    const int64_t rows = nsteps + 1;
    const int S = b->graph_steps;
    if ((rc = batch_partials(b, (size_t)std::max(S, 2)))) return rc;
    const size_t need = (size_t)b->ninst * rows * nc;
    if (need > b->signal_cap) {
        // room for runs of up to kChunkSteps steps, so that a short warm-up
        // run leaves nothing to allocate for the next one
        const size_t cap = std::max({need, 2 * b->signal_cap, (size_t)b->ninst * (kChunkSteps + 1) * nc});
        if (b->d_signal) HIP_TRY(hipFree(b->d_signal));
        b->d_signal = nullptr;
        b->signal_cap = 0;
        if ((rc = balloc(b->stream, &b->d_signal, cap))) return rc;
        b->signal_cap = cap;
    }
    // row 0: signal of every instance before the run (kept from the last
    // run, else summed over the state where it lives)
    b->sig_ok.resize(b->ninst, 0);
    b->sig_now.resize((size_t)b->ninst * nc, 0.0);
    const std::vector<uint8_t> have0 = b->sig_ok;
    const std::vector<double> sig0 = b->sig_now;
    for (int i = 0; i < b->ninst; i++) {
        if (have0[i]) continue;
        const nmr_batch_instance& s = b->spec[i];
        nmr::FieldSumArgs a{};
        for (int q = 0; q < 3; q++) a.f[q] = brick ? b->inst[i].mb[b->parity][q] : b->inst[i].m[b->parity][q];
        if (brick) {
            a.nby = (s.ny + 3) / 4;
            a.nb8 = (b->inst[i].pitch / 4 + 1) / 2;
        }
        a.mask = b->inst[i].mask;
        a.part = b->d_fpart;
        a.nx = s.nx;
        a.ny = s.ny;
        a.nz = s.nz;
        a.pitch = b->inst[i].pitch;
        a.bx0 = b->prob.buffer[0]; a.bx1 = s.nx - b->prob.buffer[0];
        a.by0 = b->prob.buffer[1]; a.by1 = s.ny - b->prob.buffer[1];
        a.bz0 = b->prob.buffer[2]; a.bz1 = s.nz - b->prob.buffer[2];
        const int nfb = nmr::field_sum_blocks(a);
        HIP_TRY(nmr::launch_field_sum(nc, a, b->stream));
        HIP_TRY(nmr::launch_reduce_partials(b->d_fpart, 1, nfb, nc, b->d_signal + (size_t)i * rows * nc, b->stream));
    }
    // the device step counter holds the global step (the gradient schedule
    // reads it); signal rows count from this run's first step
    const int step0 = (int)b->steps_done;
    HIP_TRY(hipMemcpyAsync(b->d_step, &step0, sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int64_t done = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    auto enqueue_chunk = [&](int n) -> int {
        for (int s = 0; s < n; s++) {
            int r = batch_launch_step(b, s, (b->parity + s) & 1, b->d_part + (size_t)s * b->ntiles * nc);
            if (r) return r;
        }
        HIP_TRY(nmr::launch_batch_reduce(b->d_part, b->ntiles, nc, b->d_off, b->d_ids, b->d_step, step0, rows,
                                         b->d_signal, n, b->ninst, b->stream));
        HIP_TRY(nmr::launch_step_advance(b->d_step, n, b->stream));
        return NMR_OK;
    };
    if (S > 0 && nsteps >= S) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int r = enqueue_chunk(S);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        HIP_TRY(e);
        graph = g;
        HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        while (nsteps - done >= S) {
            HIP_TRY(hipGraphLaunch(exec, b->stream));
            done += S; // S is even: the parity is unchanged
        }
    }
    while (done < nsteps) {
        const int n = (int)std::min<int64_t>(std::max(S, 2), nsteps - done);
        if ((rc = enqueue_chunk(n))) return rc;
        b->parity = (b->parity + n) & 1;
        done += n;
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    b->home = brick ? nmri::kHomeBricks : nmri::kHomeRows;
    // The signal rows go to the host in one copy of the whole
    // [ninst][rows][nc] block: into the caller's buffer, else into pinned
    // staging (up to kBatchStage doubles), and the last row of every
    // instance -- the signal of its new state -- is read from there.  Only
    // past that size are the last rows copied alone: a strided copy, which
    // the runtime splits into one small transfer per instance (256 of them
    // took 2-11 ms).  Row 0 of an instance that kept its signal is not
    // written on the device.
    double* host = signal;
    if (!host && need <= kBatchStage) {
        if (need > b->h_signal_cap) {
            const size_t cap = std::min(kBatchStage, std::max(need, 2 * b->h_signal_cap));
            if (b->h_signal) HIP_TRY(hipHostFree(b->h_signal));
            b->h_signal = nullptr;
            b->h_signal_cap = 0;
            HIP_TRY(hipHostMalloc((void**)&b->h_signal, cap * sizeof(double), hipHostMallocDefault));
            b->h_signal_cap = cap;
        }
        host = b->h_signal;
    }
    std::vector<double> last;
    if (host) {
        HIP_TRY(hipMemcpyAsync(host, b->d_signal, sizeof(double) * need, hipMemcpyDeviceToHost, b->stream));
    } else {
        last.resize((size_t)b->ninst * nc);
        HIP_TRY(hipMemcpy2DAsync(last.data(), sizeof(double) * nc, b->d_signal + (size_t)nsteps * nc,
                                 sizeof(double) * rows * nc, sizeof(double) * nc, b->ninst, hipMemcpyDeviceToHost,
                                 b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int i = 0; i < b->ninst; i++) {
        const size_t o = (size_t)i * nc;
        const double* lr = host ? host + ((size_t)i * rows + (size_t)nsteps) * nc : last.data() + o;
        if (nsteps > 0 || !have0[i]) std::copy(lr, lr + nc, b->sig_now.begin() + o);
        if (signal && have0[i]) std::copy(sig0.begin() + o, sig0.begin() + o + nc, signal + (size_t)i * rows * nc);
    }
    std::fill(b->sig_ok.begin(), b->sig_ok.end(), 1);
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    int64_t vox = 0;
    for (const auto& s : b->spec) vox += (int64_t)s.nx * s.ny * s.nz;
    b->last_voxel_steps = vox * nsteps;
    b->steps_done += nsteps;
    return NMR_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Entry points for the other host files of the library.

namespace nmri {

int batch_rows(nmr_batch* b)
{
    if (b->home & kHomeRows) return NMR_OK;
    for (int i = 0; i < b->ninst; i++) {
        const auto& s = b->spec[i];
        auto& I = b->inst[i];
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < b->nc; q++)
                HIP_TRY(nmr::launch_brick(I.mb[b->parity][q], I.m[p][q], I.pitch / 4, s.ny, s.nz, false, b->stream));
    }
    b->home |= kHomeRows;
    return NMR_OK;
}

int state_rows(nmr_ctx* c)
{
    if (c->home & kHomeRows) return NMR_OK;
    const nmr_problem& p = c->prob;
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < c->nc; q++)
            HIP_TRY(nmr::launch_brick(c->mb[c->bpar][q], c->m[b][q], c->pitch / 4, p.ny, p.nz, false, c->stream));
    c->home |= kHomeRows;
    return NMR_OK;
}

nmr::Coef make_coef(const nmr_ctx* c, const double g[3]) { return ::make_coef(c, g); }

int ensure_partials(nmr_ctx* c, size_t doubles) { return ::ensure_partials(c, doubles); }

int sweep_step(nmr_ctx* c, const nmr::Coef& cf, bool exact, int in, int out)
{
    const nmr_problem& p = c->prob;
    Launch L{};
    auto_launch(c, -1, 0, &L);
    if (L.nwg < 0) return fail(NMR_ERR_INVALID, "too many workgroups");
    int rc = refresh_inert(c);
    if (rc) return rc;
    if ((rc = ::ensure_partials(c, (size_t)L.nwg * c->nc))) return rc;
    if ((rc = ensure_adv(c, false))) return rc;
    int w[6];
    signal_window(p, w);
    nmr::SweepArgs a{};
    for (int q = 0; q < 3; q++) {
        a.min[q] = c->m[in][q];
        a.mout[q] = c->m[out][q];
        a.adv[q] = c->adv[q];
    }
    a.mask = c->mask;
    a.partials = c->partials;
    a.nx = p.nx;
    a.ny = p.ny;
    a.nz = p.nz;
    a.pitch = c->pitch;
    a.pitch1 = c->pitch1;
    a.ntx = L.ntx;
    a.nty = L.nty;
    a.ntz = L.ntz;
    a.lz = L.lz;
    a.nblocks = L.nwg;
    a.bx0 = w[0]; a.bx1 = w[1]; a.by0 = w[2]; a.by1 = w[3]; a.bz0 = w[4]; a.bz1 = w[5];
    a.inert = c->inert;
    a.c = cf;
    HIP_TRY(nmr::launch_sweep(c->nc, c->vel, exact, L.variant, a, c->stream));
    return NMR_OK;
}

} // namespace nmri
