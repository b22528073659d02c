// nmr_internal.h -- state shared by the host files of libnmr_hip.so (the C
// ABI in nmr_abi.hip and the learned corrector in nmr_learn.hip).  Not part
// of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "nmr_hip.h"
#include "nmr_kernels.h"

namespace nmri {
constexpr int kHomeRows = 1, kHomeBricks = 2;
// Record a status detail for nmr_hip_last_error() and return code.
int fail(int code, const char* fmt, ...);
inline int round_up4(int n) { return (n + 3) & ~3; }
// Bytes of the largest FP32 array of an nx*ny*nz grid in the row layout: the
// state (rows of round_up4(nx) floats) and the face arrays Vx (rows of
// round_up4(nx + 1)), Vy (ny + 1 rows per plane) and Vz (nz + 1 planes).
inline size_t biggest_array_bytes(int nx, int ny, int nz)
{
    const size_t p = (size_t)round_up4(nx), p1 = (size_t)round_up4(nx + 1);
    const size_t vx = p1 * ny * nz, vy = p * (ny + 1) * nz, vz = p * ny * (nz + 1);
    return 4 * std::max(vx, std::max(vy, vz));
}
} // namespace nmri

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return nmri::fail(e_ == hipErrorOutOfMemory ? NMR_ERR_NOMEM : NMR_ERR_DEVICE, "%s: %s (%s:%d)", \
                              #expr, hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

struct nmr_ctx {
    nmr_problem prob{};
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nc = 1;
    bool vel = false;
    bool exact = false;
    int pitch = 0, pitch1 = 0;
    size_t n_m = 0;  // floats per magnetization component (nz*ny*pitch)
    size_t n_vx = 0, n_vy = 0, n_vz = 0;
    float* m[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    float* pred[3] = {nullptr, nullptr, nullptr}; // MacCormack predictor
    bool general = false; // MacCormack, non-zero-derivative boundaries or per-voxel data: per-voxel kernel
    // per-voxel coefficients (diffusivity / d_ref, dt/T1, dt/T2) and phase weights
    float *dn = nullptr, *r1v = nullptr, *r2v = nullptr;
    float* bn = nullptr; // per-voxel field along the main-field axis bnax (nmr_hip_upload_field)
    int bnax = -1;
    float* w[2] = {nullptr, nullptr};
    double d_ref = 0.0;
    bool hetero = false, weighted = false;
    bool solids_zeroed = false; // MacCormack: both state buffers hold 0 at solids
    // the sweeps in 128-byte bricks (8 x-voxels by 4 rows): state copies and velocity
    bool brick_on = true;       // nmr_hip_set_layout
    float* mb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    float* pb[3] = {nullptr, nullptr, nullptr}; // MacCormack predictor in bricks
    // upwind velocity sums of the plain sweep (nmr::launch_adv), in rows
    // and, for brick sweeps, in bricks; computed once per velocity upload
    float* adv[3] = {nullptr, nullptr, nullptr};
    float* advb[3] = {nullptr, nullptr, nullptr};
    bool adv_valid = false, advb_valid = false;
    // the same for the plain face-velocity sums of the MacCormack stages
    float* csum[3] = {nullptr, nullptr, nullptr};
    float* csumb[3] = {nullptr, nullptr, nullptr};
    bool csum_valid = false, csumb_valid = false;
    // brick copies of the per-voxel arrays, in the order dn, r1v, r2v, bn,
    // w[0], w[1]; converted after each upload, before the next brick run
    float* pvb[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool brick_pv_valid = false;
    // Which copies hold the current state: the row buffer m[cur]
    // (nmri::kHomeRows), the brick buffer mb[bpar] (nmri::kHomeBricks), or
    // both.  A conversion writes both buffers of its target layout, so the
    // two buffers of a layout agree wherever a sweep does not write
    // (all-solid chunks); a sweep or a pulse leaves its own layout as the
    // only home of the state.
    int home = 1;
    int bpar = 0;
    // signal of the current state: host copy of the last row of the last run
    // (or of the last nmr_hip_signal); cleared when the state changes otherwise
    bool sig_valid = false;
    double sig_now[6] = {0, 0, 0, 0, 0, 0};
    int nsig() const { return weighted ? 2 * nc : nc; }
    uint8_t* mask = nullptr;
    float* vx = nullptr;
    float* vy = nullptr;
    float* vz = nullptr;
    double* partials = nullptr;
    size_t partials_cap = 0; // doubles
    double* dsignal = nullptr;
    size_t dsignal_cap = 0;  // doubles
    double* hsignal = nullptr; // pinned host staging of a run's signal rows (at most kPinnedSignal doubles)
    size_t hsignal_cap = 0;    // doubles
    int cur = 0;
    bool has_mask = false, has_mag = false;
    int64_t mask_version = 0; // bumped by every mask upload
    bool inert_dirty = true; // mask or velocity changed since the last check
    int inert = 0;           // no pore-solid face carries velocity
    int* dflag = nullptr;
    int variant = -1, lz = 0; // -1 / 0 = automatic
    double last_ms = 0.0;
    int64_t last_sweeps = 0;
    int64_t device_bytes = 0;
    // set while nmr_hip_run_static runs: every device buffer must exist
    // already and the partial-sum and signal buffers are the fixed scratch
    bool no_alloc = false;
    double* scratch = nullptr; // this device's fixed scratch (nmr_hip_run_static)
};

struct nmr_batch {
    nmr_problem prob{};
    int ninst = 0, nc = 1;
    bool vel = false, exact = false;
    std::vector<nmr_batch_instance> spec;
    struct Inst {
        int pitch = 0, pitch1 = 0;
        size_t n_m = 0, n_vx = 0, n_vy = 0, n_vz = 0;
        float* m[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
        uint8_t* mask = nullptr;
        float *vx = nullptr, *vy = nullptr, *vz = nullptr;
        bool uploaded = false;
        int inert = 0;
        int64_t npore = 0;
        // upwind velocity sums of the plain sweep (rows), computed at upload
        float* adv[3] = {nullptr, nullptr, nullptr};
        // brick copies (8 x-voxels by 4 rows) of the state and the velocity sums
        float* mb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
        float* advb[3] = {nullptr, nullptr, nullptr};
        bool adv_bricks = false; // the brick sums hold the uploaded velocity's
    };
    std::vector<Inst> inst;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int variant = -1, lz = 0, graph_steps = 64, persistent = 1;
    int parity = 0;
    // the sweeps on bricks (nmr_hip_batch_set_layout); which layout holds
    // every instance's current state (nmri::kHomeRows / kHomeBricks bits; a
    // conversion writes both buffers of its target layout)
    bool brick_on = true;
    int home = 1;
    // signal of every instance's current state ([ninst][nc], host copy of the
    // last rows of the last run) and, per instance, whether it is still the
    // signal of that state (an upload clears it)
    std::vector<double> sig_now;
    std::vector<uint8_t> sig_ok;
    int64_t steps_done = 0;
    // launch state for the current variant / lz / grid kind
    int cur_variant = -2, cur_lz = -1, cur_persistent = -1, grid = 0, ntiles = 0;
    bool cur_brick = false;
    // per-instance tile shapes (variant -3): one launch per shape, over the
    // tile range [first, first + count) of the tile list
    struct Group { int variant, first, count, grid; };
    std::vector<Group> groups;
    // time per padded voxel of each one-step tile shape (variant -3's
    // weights): single-domain figures until the autotuner measures the batch
    std::array<double, 7> tile_cost = {1.000, 1.088, 1.055, 1.154, 1.161, 1.257, 1.216};
    nmr::BatchInst* d_inst = nullptr;
    int2* d_tiles = nullptr;
    int *d_off = nullptr, *d_ids = nullptr, *d_step = nullptr;
    double* d_part = nullptr;
    size_t part_cap = 0;
    double* d_signal = nullptr;
    double* h_signal = nullptr; // pinned host staging of a run's signal block
    size_t h_signal_cap = 0;    // doubles
    size_t signal_cap = 0;
    double* d_fpart = nullptr; // field-sum partials (max nz)
    int* d_flag = nullptr;
    double last_ms = 0.0;
    int64_t last_voxel_steps = 0;
};

namespace nmri {
// Coefficients of one step of a context's problem under gradient g (T/m).
nmr::Coef make_coef(const nmr_ctx* c, const double g[3]);
// Grow the context's partial-sum buffer to hold `doubles` values.
int ensure_partials(nmr_ctx* c, size_t doubles);
// One z-sweep step with coefficients cf from buffer `in` into buffer `out`
// of the context (the automatic one-step tile), on its stream; the step's
// signal partials go to the context's partial buffer.
int sweep_step(nmr_ctx* c, const nmr::Coef& cf, bool exact, int in, int out);
// Make the row buffers hold the current state (both of them, converted from
// the bricks when only those hold it).
int state_rows(nmr_ctx* c);
// The same for every instance of a batch.
int batch_rows(nmr_batch* b);
} // namespace nmri
