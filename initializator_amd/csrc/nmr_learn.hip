// nmr_learn.hip -- the learned corrector: a small 3-D convolution net that
// turns one coarse step of factor*dt into the result of factor plain steps
// (host side of nmr_hip_corrector_* and nmr_hip_run_learned).
//
// Net: conv 6->16, ReLU, conv 16->16, ReLU, conv 16->2, 3x3x3 kernels with
// zero padding.  Input channels: (Mx, My) of the state before the coarse
// step (zero in solids), the pore mask, and the cell-centred velocity
// divided by a velocity scale.  Output: the correction, in units of an
// output scale, added at pores to the coarse result.  FP32 master weights
// with Adam; BF16 weights and activations in the kernels, FP32 accumulate.
//
// Training data come from the batched solver: every round advances all
// instances `stride` plain steps, keeps their state, advances `factor` more
// plain steps, and for each instance fits the net to (plain result - coarse
// step from the kept state).
//
// Inference: each coarse step is checked before it is accepted -- no
// non-finite value, pore energy sum(Mx^2 + My^2) not above (1 + tol) times
// the previous state's (the exact-rotation update only dissipates), and no
// correction larger than tol times the largest |M| of the previous state --
// and otherwise replaced by factor plain steps from the kept state.  The net
// computes only the rows its pore outputs depend on.
#include <random>

#include "nmr_internal.h"

using nmri::fail;

namespace {

constexpr int kParams = 10402;
// canonical layout: W1 [16][6][27], b1 [16], W2 [16][16][27], b2 [16], W3 [2][16][27], b3 [2]
constexpr int kW1 = 0, kB1 = 2592, kW2 = 2608, kB2 = 9520, kW3 = 9536, kB3 = 10400;
constexpr int kFragF1 = 7 * 512, kFragF = 14 * 512; // uint16 per fragment set (cin 8 / 16)
constexpr int kWgradGrid = 128;                     // workgroups of the weight-gradient kernel

struct Tensors {
    int nx = 0, ny = 0, nz = 0;
    size_t cap = 0; // padded voxels allocated
    uint16_t *x = nullptr, *a1 = nullptr, *a2 = nullptr, *g3 = nullptr, *g2 = nullptr, *g1 = nullptr;
    float* p = nullptr;
};

} // namespace

struct nmr_corrector {
    hipStream_t stream = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *f1 = nullptr, *f2 = nullptr, *f3 = nullptr, *d2 = nullptr, *d3 = nullptr; // fragments
    int64_t adam_t = 0;
    double vscale = 1e-3, out_scale = 0.0;
    Tensors t;
    float* wpart = nullptr; // wgrad partials
    double* lpart = nullptr;
    size_t lpart_cap = 0;
    double* apart = nullptr; // apply partials
    size_t apart_cap = 0;
    double* dstat = nullptr; // reduced statistics [6]
    bool train_ws = false;
    // rows of the three layers that the pores of one context need
    const nmr_ctx* rows_ctx = nullptr;
    int64_t rows_version = -1;
    int* rows[3] = {nullptr, nullptr, nullptr};
    int nrows[3] = {0, 0, 0};
    // convolution kernel (NMR_CONV_*) and planes per z-marching column
    int conv = 0;
    int zc = 16;
};

namespace {

// A cleared device buffer.  The clear runs on the corrector's stream and is
// waited for: that stream is non-blocking, so a clear on the null stream
// (hipMemset) would not be ordered before the corrector's own work on the
// buffer and could land after it.
template <typename T>
int dalloc(hipStream_t s, T** p, size_t count)
{
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_TRY(hipMalloc((void**)p, bytes));
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NMR_OK;
}

void tfree(Tensors& t)
{
    for (void* q : {(void*)t.x, (void*)t.a1, (void*)t.a2, (void*)t.g3, (void*)t.g2, (void*)t.g1, (void*)t.p})
        if (q) (void)hipFree(q);
    t = Tensors{};
}

// Tensors for an nx*ny*nz grid; halos zeroed.  Training needs the gradient
// tensors too.
int ensure_tensors(nmr_corrector* c, int nx, int ny, int nz, bool train)
{
    const size_t nv = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    Tensors& t = c->t;
    if (nv > t.cap || (train && !c->train_ws)) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        const size_t cap = std::max(nv, t.cap);
        tfree(t);
        int rc;
        if ((rc = dalloc(c->stream, &t.x, cap * 8)) || (rc = dalloc(c->stream, &t.a1, cap * 16)) ||
            (rc = dalloc(c->stream, &t.a2, cap * 16)) || (rc = dalloc(c->stream, &t.p, cap * 2)))
            return rc;
        if (train && ((rc = dalloc(c->stream, &t.g3, cap * 16)) || (rc = dalloc(c->stream, &t.g2, cap * 16)) ||
                      (rc = dalloc(c->stream, &t.g1, cap * 16))))
            return rc;
        c->train_ws = train;
        t.cap = cap;
    } else if (t.nx != nx || t.ny != ny || t.nz != nz) {
        // a different padded shape: old interiors would sit in the new halos
        HIP_TRY(hipMemsetAsync(t.x, 0, nv * 8 * 2, c->stream));
        for (uint16_t* q : {t.a1, t.a2, t.g3, t.g2, t.g1})
            if (q) HIP_TRY(hipMemsetAsync(q, 0, nv * 16 * 2, c->stream));
    }
    t.nx = nx;
    t.ny = ny;
    t.nz = nz;
    return NMR_OK;
}

int repack(nmr_corrector* c)
{
    HIP_TRY(nmr::launch_pack(c->w + kW1, 6, 16, 8, 0, c->f1, c->stream));
    HIP_TRY(nmr::launch_pack(c->w + kW2, 16, 16, 16, 0, c->f2, c->stream));
    HIP_TRY(nmr::launch_pack(c->w + kW3, 16, 2, 16, 0, c->f3, c->stream));
    HIP_TRY(nmr::launch_pack(c->w + kW2, 16, 16, 16, 1, c->d2, c->stream));
    HIP_TRY(nmr::launch_pack(c->w + kW3, 16, 2, 16, 1, c->d3, c->stream));
    return NMR_OK;
}

int conv_grid(int nrows) { return std::max(1, std::min(2048, (nrows + 3) / 4)); }
// persistent grid of the tiled convolution: 4 resident workgroups per CU
// (34.6 KB of LDS each at 16 channels) on 256 CUs
int tile_grid(const Tensors& t) { return std::max(1, std::min(1024, nmr::conv_tiles(t.nx, t.ny, t.nz))); }
// persistent grid of the z-marching convolution: 3 resident workgroups per
// CU at 16 input channels (41.5 KB of LDS each), 4 at 8
int zm_grid(const Tensors& t, int cin, int zc)
{
    return std::max(1, std::min(cin == 8 ? 1024 : 768, nmr::conv_columns(t.nx, t.ny, t.nz, zc)));
}

// One dense convolution layer with the context's kernel choice.
int conv_dense(nmr_corrector* c, int cin, int mode, nmr::ConvArgs a)
{
    const Tensors& t = c->t;
    if (c->conv == NMR_CONV_TILES) {
        HIP_TRY(nmr::launch_conv_tiled(cin, mode, a, tile_grid(t), c->stream));
    } else if (c->conv == NMR_CONV_ROWS) {
        a.rows = nullptr;
        a.nrows = a.nseg * t.ny * t.nz;
        HIP_TRY(nmr::launch_conv(cin, mode, a, conv_grid(a.nrows), c->stream));
    } else {
        a.zc = c->zc;
        HIP_TRY(nmr::launch_conv_zm(cin, mode, a, zm_grid(t, cin, c->zc), c->stream));
    }
    return NMR_OK;
}

// Forward pass on the tensors (x already filled): a1, a2, p.  Layers 1 and
// 2 run over every row; with row lists (sparse = true) layer 3 computes
// only the rows c->rows[2] (those holding a pore) unless the z-marching
// kernel runs it over every row.  With a state (smx, smy; z-marching kernels
// only) layer 1 takes input channels 0 and 1 from it instead of from x.
int forward(nmr_corrector* c, bool sparse = false, const float* smx = nullptr, const float* smy = nullptr,
            int spitch = 0)
{
    const Tensors& t = c->t;
    nmr::ConvArgs a{};
    a.nx = t.nx;
    a.ny = t.ny;
    a.nz = t.nz;
    a.nseg = (t.nx + 15) / 16;
    const int all = a.nseg * t.ny * t.nz;
    auto rows = [&](int l) {
        a.rows = sparse ? c->rows[l] : nullptr;
        a.nrows = sparse ? c->nrows[l] : all;
        return conv_grid(a.nrows);
    };
    a.in = t.x;
    a.wa = c->f1;
    a.bias = c->w + kB1;
    a.out = t.a1;
    int rc = NMR_OK;
    if (smx) {
        a.smx = smx;
        a.smy = smy;
        a.spitch = spitch;
        a.zc = c->zc;
        HIP_TRY(nmr::launch_conv_zm(8, 4, a, zm_grid(t, 8, c->zc), c->stream));
        a.smx = a.smy = nullptr;
    } else if ((rc = conv_dense(c, 8, 0, a))) {
        return rc;
    }
    a.in = t.a1;
    a.wa = c->f2;
    a.bias = c->w + kB2;
    a.out = t.a2;
    rc = conv_dense(c, 16, 0, a);
    if (rc) return rc;
    a.in = t.a2;
    a.wa = c->f3;
    a.bias = c->w + kB3;
    a.out = nullptr;
    a.outf = t.p;
    if (sparse && c->conv != NMR_CONV_ZMARCH) {
        const int grid = rows(2);
        if (a.nrows > 0) HIP_TRY(nmr::launch_conv(16, 1, a, grid, c->stream));
        return NMR_OK;
    }
    return conv_dense(c, 16, 1, a);
}

// Row lists of a context's mask: layer 3 computes the 16-voxel rows that
// hold a pore, layer 2 those rows dilated by one row in y, z and one
// segment in x, layer 1 that set dilated once more.
int build_rows(nmr_corrector* c, nmr_ctx* ctx)
{
    if (c->rows_ctx == ctx && c->rows_version == ctx->mask_version) return NMR_OK;
    const nmr_problem& p = ctx->prob;
    const int nseg = (p.nx + 15) / 16;
    const size_t n = (size_t)nseg * p.ny * p.nz;
    uint8_t* d = nullptr;
    HIP_TRY(hipMalloc((void**)&d, std::max<size_t>(n, 256)));
    std::vector<uint8_t> f(n);
    hipError_t e = nmr::launch_row_flags(ctx->mask, p.nx, p.ny, p.nz, ctx->pitch, d, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(f.data(), d, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(NMR_ERR_DEVICE, "row flags: %s", hipGetErrorString(e));
    auto dilate = [&](const std::vector<uint8_t>& in) {
        std::vector<uint8_t> a(in), b(n);
        const size_t st[3] = {1, (size_t)nseg, (size_t)nseg * p.ny};
        const int ext[3] = {nseg, p.ny, p.nz};
        for (int ax = 0; ax < 3; ax++) {
            for (size_t i = 0; i < n; i++) {
                const int coord = (int)((i / st[ax]) % ext[ax]);
                uint8_t v = a[i];
                if (coord > 0) v |= a[i - st[ax]];
                if (coord + 1 < ext[ax]) v |= a[i + st[ax]];
                b[i] = v;
            }
            a.swap(b);
        }
        return a;
    };
    std::vector<uint8_t> l3 = f, l2 = dilate(l3), l1 = dilate(l2);
    const std::vector<uint8_t>* sets[3] = {&l1, &l2, &l3};
    for (int l = 0; l < 3; l++) {
        std::vector<int> ids;
        for (size_t i = 0; i < n; i++)
            if ((*sets[l])[i]) ids.push_back((int)i);
        if (c->rows[l]) HIP_TRY(hipFree(c->rows[l]));
        c->rows[l] = nullptr;
        int rc = dalloc(c->stream, &c->rows[l], ids.size());
        if (rc) return rc;
        if (!ids.empty()) HIP_TRY(hipMemcpy(c->rows[l], ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
        c->nrows[l] = (int)ids.size();
    }
    c->rows_ctx = ctx;
    c->rows_version = ctx->mask_version;
    return NMR_OK;
}

// Backward pass from the output gradient in g3: parameter gradients into
// c->grad (overwritten).
int backward(nmr_corrector* c)
{
    const Tensors& t = c->t;
    nmr::ConvArgs a{};
    a.nx = t.nx;
    a.ny = t.ny;
    a.nz = t.nz;
    a.nseg = (t.nx + 15) / 16;
    a.nrows = a.nseg * t.ny * t.nz;
    a.in = t.g3;
    a.wa = c->d3;
    a.ref = t.a2;
    a.out = t.g2;
    int rc = conv_dense(c, 16, 2, a);
    if (rc) return rc;
    a.in = t.g2;
    a.wa = c->d2;
    a.ref = t.a1;
    a.out = t.g1;
    rc = conv_dense(c, 16, 2, a);
    if (rc) return rc;
    nmr::WgradArgs w{};
    w.nx = t.nx;
    w.ny = t.ny;
    w.nz = t.nz;
    w.part = c->wpart;
    const int nw = kWgradGrid * 4;
    w.in = t.a2;
    w.gout = t.g3;
    HIP_TRY(nmr::launch_wgrad(16, w, kWgradGrid, c->stream));
    HIP_TRY(nmr::launch_wreduce(c->wpart, nw, 16, 2, c->grad + kW3, c->stream));
    w.in = t.a1;
    w.gout = t.g2;
    HIP_TRY(nmr::launch_wgrad(16, w, kWgradGrid, c->stream));
    HIP_TRY(nmr::launch_wreduce(c->wpart, nw, 16, 16, c->grad + kW2, c->stream));
    w.in = t.x;
    w.gout = t.g1;
    HIP_TRY(nmr::launch_wgrad(8, w, kWgradGrid, c->stream));
    HIP_TRY(nmr::launch_wreduce(c->wpart, nw, 6, 16, c->grad + kW1, c->stream));
    return NMR_OK;
}

int adam_step(nmr_corrector* c, double lr)
{
    c->adam_t += 1;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, (double)c->adam_t)), c2 = (float)(1.0 - std::pow(b2, (double)c->adam_t));
    HIP_TRY(nmr::launch_adam(c->w, c->m, c->v, c->grad, kParams, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2,
                             c->stream));
    return repack(c);
}

int ensure_dev_doubles(nmr_corrector* c, double** p, size_t* cap, size_t n)
{
    if (n <= *cap) return NMR_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    int rc = dalloc(c->stream, p, n);
    if (rc) return rc;
    *cap = n;
    return NMR_OK;
}

// Interior of a padded FP32 2-channel tensor to host [nz][ny][nx][2].
int download_p(nmr_corrector* c, float* out)
{
    const Tensors& t = c->t;
    const size_t nv = (size_t)(t.nx + 2) * (t.ny + 2) * (t.nz + 2);
    std::vector<float> h(nv * 2);
    HIP_TRY(hipMemcpyAsync(h.data(), t.p, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int z = 0; z < t.nz; z++)
        for (int y = 0; y < t.ny; y++)
            for (int x = 0; x < t.nx; x++) {
                const size_t pv = ((size_t)(z + 1) * (t.ny + 2) + (y + 1)) * (t.nx + 2) + (x + 1);
                const size_t v = ((size_t)z * t.ny + y) * t.nx + x;
                out[2 * v] = h[2 * pv];
                out[2 * v + 1] = h[2 * pv + 1];
            }
    return NMR_OK;
}

int upload_features(nmr_corrector* c, int nx, int ny, int nz, const float* f6, float** dev)
{
    const size_t n = (size_t)nx * ny * nz * 6;
    HIP_TRY(hipMalloc((void**)dev, std::max<size_t>(n * 4, 256)));
    HIP_TRY(hipMemcpyAsync(*dev, f6, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(nmr::launch_feature_copy(*dev, nx, ny, nz, c->t.x, c->stream));
    return NMR_OK;
}

int check_dims(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1 || (double)(nx + 2) * (ny + 2) * (nz + 2) > 2.0e9)
        return fail(NMR_ERR_INVALID, "grid %d x %d x %d", nx, ny, nz);
    return NMR_OK;
}

// Coefficients of one coarse step of `factor` steps for problem q whose
// per-step z gradients are gz[0..factor-1]: dt*factor and the summed phase.
nmr::Coef coarse_coef(const nmr_problem& q, int factor, const double* gz_sum)
{
    nmr_ctx tmp;
    tmp.prob = q;
    tmp.prob.dt = q.dt * factor;
    const double g[3] = {0.0, 0.0, *gz_sum / factor};
    return nmri::make_coef(&tmp, g);
}

} // namespace

extern "C" {

int nmr_hip_corrector_num_params(void) { return kParams; }

int nmr_hip_corrector_create(uint64_t seed, nmr_corrector** out)
{
    if (!out) return fail(NMR_ERR_NULL, "nmr_hip_corrector_create: null argument");
    *out = nullptr;
    nmr_corrector* c = new (std::nothrow) nmr_corrector();
    if (!c) return fail(NMR_ERR_NOMEM, "host allocation");
    auto bail = [&](int rc) {
        nmr_hip_corrector_destroy(c);
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NMR_ERR_DEVICE, "stream creation failed"));
    int rc;
    hipStream_t s = c->stream;
    if ((rc = dalloc(s, &c->w, kParams)) || (rc = dalloc(s, &c->m, kParams)) || (rc = dalloc(s, &c->v, kParams)) ||
        (rc = dalloc(s, &c->grad, kParams)) || (rc = dalloc(s, &c->f1, kFragF1)) || (rc = dalloc(s, &c->f2, kFragF)) ||
        (rc = dalloc(s, &c->f3, kFragF)) || (rc = dalloc(s, &c->d2, kFragF)) || (rc = dalloc(s, &c->d3, kFragF)) ||
        (rc = dalloc(s, &c->wpart, (size_t)kWgradGrid * 4 * nmr::kWgradPart)) || (rc = dalloc(s, &c->dstat, 8)))
        return bail(rc);
    // He-normal weights (last layer scaled by 0.1), zero biases
    std::mt19937_64 gen(seed);
    std::vector<float> h(kParams, 0.f);
    auto fill = [&](int off, int n, double sd) {
        std::normal_distribution<double> nd(0.0, sd);
        for (int i = 0; i < n; i++) h[off + i] = (float)nd(gen);
    };
    fill(kW1, 2592, std::sqrt(2.0 / (6 * 27)));
    fill(kW2, 6912, std::sqrt(2.0 / (16 * 27)));
    fill(kW3, 864, 0.1 * std::sqrt(2.0 / (16 * 27)));
    if (hipMemcpy(c->w, h.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(NMR_ERR_DEVICE, "weight upload failed"));
    if ((rc = repack(c))) return bail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(NMR_ERR_DEVICE, "stream sync failed"));
    *out = c;
    return NMR_OK;
}

int nmr_hip_corrector_destroy(nmr_corrector* c)
{
    if (!c) return NMR_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    tfree(c->t);
    for (void* q : {(void*)c->w, (void*)c->m, (void*)c->v, (void*)c->grad, (void*)c->f1, (void*)c->f2, (void*)c->f3,
                    (void*)c->d2, (void*)c->d3, (void*)c->wpart, (void*)c->lpart, (void*)c->apart, (void*)c->dstat,
                    (void*)c->rows[0], (void*)c->rows[1], (void*)c->rows[2]})
        if (q) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NMR_OK;
}

int nmr_hip_corrector_get_params(nmr_corrector* c, float* params)
{
    if (!c || !params) return fail(NMR_ERR_NULL, "nmr_hip_corrector_get_params: null argument");
    HIP_TRY(hipMemcpyAsync(params, c->w, kParams * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NMR_OK;
}

int nmr_hip_corrector_set_params(nmr_corrector* c, const float* params)
{
    if (!c || !params) return fail(NMR_ERR_NULL, "nmr_hip_corrector_set_params: null argument");
    HIP_TRY(hipMemcpyAsync(c->w, params, kParams * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->m, 0, kParams * 4, c->stream));
    HIP_TRY(hipMemsetAsync(c->v, 0, kParams * 4, c->stream));
    c->adam_t = 0;
    int rc = repack(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return NMR_OK;
}

int nmr_hip_corrector_set_scales(nmr_corrector* c, double vscale, double out_scale)
{
    if (!c) return fail(NMR_ERR_NULL, "nmr_hip_corrector_set_scales: null argument");
    if (!(vscale > 0) || !(out_scale >= 0)) return fail(NMR_ERR_INVALID, "scales %g %g", vscale, out_scale);
    c->vscale = vscale;
    c->out_scale = out_scale;
    return NMR_OK;
}

int nmr_hip_corrector_set_conv(nmr_corrector* c, int32_t kind, int32_t zc)
{
    if (!c) return fail(NMR_ERR_NULL, "nmr_hip_corrector_set_conv: null argument");
    if (kind < NMR_CONV_ZMARCH || kind > NMR_CONV_ZMARCH_SPARSE)
        return fail(NMR_ERR_INVALID, "nmr_hip_corrector_set_conv: kind %d", kind);
    if (zc < 0 || zc > 4096) return fail(NMR_ERR_INVALID, "nmr_hip_corrector_set_conv: zc %d", zc);
    c->conv = kind;
    c->zc = zc > 0 ? zc : 16;
    return NMR_OK;
}

int nmr_hip_corrector_get_scales(nmr_corrector* c, double* vscale, double* out_scale)
{
    if (!c) return fail(NMR_ERR_NULL, "nmr_hip_corrector_get_scales: null argument");
    if (vscale) *vscale = c->vscale;
    if (out_scale) *out_scale = c->out_scale;
    return NMR_OK;
}

int nmr_hip_corrector_forward(nmr_corrector* c, int32_t nx, int32_t ny, int32_t nz, const float* features, float* out)
{
    if (!c || !features || !out) return fail(NMR_ERR_NULL, "nmr_hip_corrector_forward: null argument");
    int rc = check_dims(nx, ny, nz);
    if (rc || (rc = ensure_tensors(c, nx, ny, nz, false))) return rc;
    float* dev = nullptr;
    rc = upload_features(c, nx, ny, nz, features, &dev);
    if (!rc) rc = forward(c);
    if (!rc) rc = download_p(c, out);
    (void)hipStreamSynchronize(c->stream);
    if (dev) (void)hipFree(dev);
    return rc;
}

int nmr_hip_corrector_backward(nmr_corrector* c, int32_t nx, int32_t ny, int32_t nz, const float* features,
                               const float* grad_out, float* grad_params)
{
    if (!c || !features || !grad_out || !grad_params)
        return fail(NMR_ERR_NULL, "nmr_hip_corrector_backward: null argument");
    int rc = check_dims(nx, ny, nz);
    if (rc || (rc = ensure_tensors(c, nx, ny, nz, true))) return rc;
    float* dev = nullptr;
    float* dg = nullptr;
    const size_t n = (size_t)nx * ny * nz;
    rc = upload_features(c, nx, ny, nz, features, &dev);
    if (!rc) rc = forward(c);
    if (!rc) {
        if (hipMalloc((void**)&dg, std::max<size_t>(n * 8, 256)) != hipSuccess ||
            hipMemcpyAsync(dg, grad_out, n * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            nmr::launch_gout_copy(dg, nx, ny, nz, c->t.g3, c->stream) != hipSuccess)
            rc = fail(NMR_ERR_DEVICE, "output gradient upload failed");
    }
    if (!rc) rc = backward(c);
    if (!rc && (hipMemcpyAsync(grad_params, c->grad, kParams * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(NMR_ERR_DEVICE, "gradient download failed");
    (void)hipStreamSynchronize(c->stream);
    if (dev) (void)hipFree(dev);
    if (dg) (void)hipFree(dg);
    return rc;
}

int nmr_hip_corrector_train(nmr_corrector* c, nmr_batch* b, const nmr_train_opts* o, double* losses)
{
    if (!c || !b || !o) return fail(NMR_ERR_NULL, "nmr_hip_corrector_train: null argument");
    if (b->nc != 2 || !b->exact)
        return fail(NMR_ERR_UNSUPPORTED, "training needs a batch with ncomp 2 and the exact phase");
    if (o->factor < 1 || o->factor > 1024 || o->rounds < 1 || o->stride < 0 || !(o->lr > 0))
        return fail(NMR_ERR_INVALID, "training options");
    for (const auto& I : b->inst)
        if (!I.uploaded) return fail(NMR_ERR_STATE, "train before every instance is uploaded");
    // the samples are read from the instances' row buffers: the batch runs
    // on rows while training and leaves its state there
    if (int rc0 = nmri::batch_rows(b)) return rc0;
    struct RowsOnly {
        nmr_batch* b;
        bool keep;
        ~RowsOnly()
        {
            b->brick_on = keep;
            b->home = nmri::kHomeRows;
        }
    } guard{b, b->brick_on};
    b->brick_on = false;
    b->home = nmri::kHomeRows;
    const nmr_problem& P = b->prob;
    size_t nmax = 0;
    int mx = 1, my = 1, mz = 1;
    for (int i = 0; i < b->ninst; i++) {
        nmax = std::max(nmax, b->inst[i].n_m);
        mx = std::max(mx, b->spec[i].nx);
        my = std::max(my, b->spec[i].ny);
        mz = std::max(mz, b->spec[i].nz);
    }
    // allocate for the largest instance first so that later shapes only clear
    int rc = ensure_tensors(c, mx, my, mz, true);
    if (rc) return rc;
    float *coarse[2] = {nullptr, nullptr}, *tgt[2] = {nullptr, nullptr};
    double* spart = nullptr;
    std::vector<float*> kept((size_t)b->ninst * 2, nullptr); // state of every instance before the plain steps
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (float* q : {coarse[0], coarse[1], tgt[0], tgt[1]})
            if (q) (void)hipFree(q);
        for (float* q : kept)
            if (q) (void)hipFree(q);
        if (spart) (void)hipFree(spart);
    };
    for (int q = 0; q < 2 && !rc; q++)
        if ((rc = dalloc(c->stream, &coarse[q], nmax)) || (rc = dalloc(c->stream, &tgt[q], nmax))) break;
    for (int i = 0; i < b->ninst && !rc; i++)
        for (int q = 0; q < 2 && !rc; q++) rc = dalloc(c->stream, &kept[2 * i + q], b->inst[i].n_m);
    const int nlb = nmr::loss_blocks(mx, my, mz);
    if (!rc) rc = ensure_dev_doubles(c, &c->lpart, &c->lpart_cap, (size_t)nlb);
    // partials of the coarse sweep (tile 32 x 32, z-chunks of 8)
    const size_t nswp = (size_t)((mx + 31) / 32) * ((my + 31) / 32) * ((mz + 7) / 8);
    if (!rc) rc = dalloc(c->stream, &spart, nswp * 2);
    if (rc) {
        cleanup();
        return rc;
    }
    std::vector<double> hl(nlb);
    const float vinv = (float)(1.0 / c->vscale);
    for (int r = 0; r < o->rounds && !rc; r++) {
        if (o->stride > 0 && (rc = nmr_hip_batch_run(b, o->stride, nullptr))) break;
        const int p0 = b->parity;
        const int64_t step0 = b->steps_done;
        for (int i = 0; i < b->ninst && !rc; i++)
            for (int q = 0; q < 2 && !rc; q++)
                if (hipMemcpyAsync(kept[2 * i + q], b->inst[i].m[p0][q], b->inst[i].n_m * 4, hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess)
                    rc = fail(NMR_ERR_DEVICE, "state copy failed");
        if (rc) break;
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((rc = nmr_hip_batch_run(b, o->factor, nullptr))) break;
        const int p1 = b->parity;
        for (int i = 0; i < b->ninst && !rc; i++) {
            const nmr_batch_instance& s = b->spec[i];
            const auto& I = b->inst[i];
            // coarse step from the kept state
            nmr_problem q = P;
            q.nx = s.nx;
            q.ny = s.ny;
            q.nz = s.nz;
            q.rho1 = s.rho1;
            q.rho2 = s.rho2;
            const int nd = (int)std::lround(s.delta / P.dt), nD = (int)std::lround(s.big_delta / P.dt);
            double gsum = 0.0;
            for (int k = 0; k < o->factor; k++) {
                const int64_t st = step0 + k;
                gsum += (st >= nD && st < nD + nd) ? -s.grad_amp : (st < nd ? s.grad_amp : 0.0);
            }
            if (!(P.flags & NMR_F_PRECESSION)) gsum = 0.0;
            nmr::SweepArgs a{};
            a.min[0] = kept[2 * i];
            a.min[1] = kept[2 * i + 1];
            a.mout[0] = coarse[0];
            a.mout[1] = coarse[1];
            a.mask = I.mask;
            for (int qq = 0; qq < 3; qq++) a.adv[qq] = I.adv[qq];
            a.partials = spart;
            a.nx = s.nx;
            a.ny = s.ny;
            a.nz = s.nz;
            a.pitch = I.pitch;
            a.pitch1 = I.pitch1;
            // variant 6: tile 32 x 32, lz 8
            a.lz = std::min(8, s.nz);
            a.ntx = (s.nx + 31) / 32;
            a.nty = (s.ny + 31) / 32;
            a.ntz = (s.nz + a.lz - 1) / a.lz;
            a.nblocks = a.ntx * a.nty * a.ntz;
            if ((size_t)a.nblocks > nswp) {
                rc = fail(NMR_ERR_INVALID, "coarse partials too small");
                break;
            }
            a.bx0 = 0; a.bx1 = s.nx; a.by0 = 0; a.by1 = s.ny; a.bz0 = 0; a.bz1 = s.nz;
            a.inert = I.inert;
            a.c = coarse_coef(q, o->factor, &gsum);
            // coarse output: solids keep their value
            for (int qq = 0; qq < 2; qq++)
                HIP_TRY(hipMemcpyAsync(coarse[qq], kept[2 * i + qq], I.n_m * 4, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(nmr::launch_sweep(2, b->vel, true, 6, a, c->stream));
            HIP_TRY(nmr::launch_diff(I.m[p1][0], I.m[p1][1], coarse[0], coarse[1], tgt[0], tgt[1], (long)I.n_m,
                                     c->stream));
            if ((rc = ensure_tensors(c, s.nx, s.ny, s.nz, true))) break;
            nmr::FeatureArgs f{};
            f.mx = kept[2 * i];
            f.my = kept[2 * i + 1];
            f.mask = I.mask;
            f.vx = I.vx;
            f.vy = I.vy;
            f.vz = I.vz;
            f.out = c->t.x;
            f.nx = s.nx;
            f.ny = s.ny;
            f.nz = s.nz;
            f.pitch = I.pitch;
            f.pitch1 = I.pitch1;
            f.vinv = b->vel ? vinv : 0.f;
            HIP_TRY(nmr::launch_features(f, c->stream));
            if ((rc = forward(c))) break;
            nmr::LossArgs la{};
            la.pred = c->t.p;
            la.tx = tgt[0];
            la.ty = tgt[1];
            la.mask = I.mask;
            la.gout = c->t.g3;
            la.lpart = c->lpart;
            la.nx = s.nx;
            la.ny = s.ny;
            la.nz = s.nz;
            la.pitch = I.pitch;
            if (!(c->out_scale > 0)) {
                // output scale: RMS of the first sample's target over pores
                la.tinv = 1.f;
                la.winv = 0.f;
                HIP_TRY(hipMemsetAsync(c->t.p, 0, (size_t)(s.nx + 2) * (s.ny + 2) * (s.nz + 2) * 8, c->stream));
                HIP_TRY(nmr::launch_loss(la, c->stream));
                const int nb = nmr::loss_blocks(s.nx, s.ny, s.nz);
                HIP_TRY(hipMemcpyAsync(hl.data(), c->lpart, nb * 8, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                double ss = 0.0;
                for (int k = 0; k < nb; k++) ss += hl[k];
                c->out_scale = std::sqrt(ss / std::max<int64_t>(1, I.npore)) + 1e-30;
                if ((rc = forward(c))) break;
            }
            la.tinv = (float)(1.0 / c->out_scale);
            la.winv = (float)(1.0 / std::max<int64_t>(1, I.npore));
            HIP_TRY(nmr::launch_loss(la, c->stream));
            if ((rc = backward(c))) break;
            if ((rc = adam_step(c, o->lr))) break;
            if (losses) {
                const int nb = nmr::loss_blocks(s.nx, s.ny, s.nz);
                HIP_TRY(hipMemcpyAsync(hl.data(), c->lpart, nb * 8, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                double ss = 0.0;
                for (int k = 0; k < nb; k++) ss += hl[k];
                losses[(size_t)r * b->ninst + i] = ss / std::max<int64_t>(1, I.npore);
            }
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    cleanup();
    return rc;
}

int nmr_hip_run_learned(nmr_ctx* ctx, nmr_corrector* c, int64_t ncoarse, int32_t factor, const double* grad_sched,
                        double tol, double* signal, int64_t* fallbacks)
{
    if (!ctx || !c) return fail(NMR_ERR_NULL, "nmr_hip_run_learned: null argument");
    const nmr_problem& p = ctx->prob;
    if (ctx->nc != 2 || !ctx->exact || ctx->general)
        return fail(NMR_ERR_UNSUPPORTED, "the learned solve needs ncomp 2, the exact phase, the upwind scheme and "
                                         "zero-derivative boundaries");
    if (ncoarse < 0 || factor < 1 || factor > 1024 || !(tol > 0))
        return fail(NMR_ERR_INVALID, "ncoarse %lld, factor %d, tol %g", (long long)ncoarse, (int)factor, tol);
    if (!ctx->has_mask || !ctx->has_mag) return fail(NMR_ERR_STATE, "run before mask and magnetization upload");
    if (!(c->out_scale > 0)) return fail(NMR_ERR_STATE, "the corrector has no output scale (train or set it)");
    int rc = ensure_tensors(c, p.nx, p.ny, p.nz, false);
    if (rc) return rc;
    const int nab = nmr::apply_blocks(p.nx, p.ny, p.nz);
    if ((rc = ensure_dev_doubles(c, &c->apart, &c->apart_cap, (size_t)nab * 6))) return rc;
    if ((rc = build_rows(c, ctx))) return rc;
    // The loop works on the row buffers (the coarse step, the net, the
    // statistics and the fallback runs); it leaves the state there and the
    // context's layout switch as it found it.
    if ((rc = nmri::state_rows(ctx))) return rc;
    struct RowsOnly {
        nmr_ctx* c;
        bool keep;
        ~RowsOnly()
        {
            c->brick_on = keep;
            c->home = nmri::kHomeRows;
            c->sig_valid = false;
        }
    } guard{ctx, ctx->brick_on};
    ctx->brick_on = false;
    ctx->home = nmri::kHomeRows;
    ctx->sig_valid = false;
    nmr::ApplyArgs ap{};
    ap.mask = ctx->mask;
    ap.part = c->apart;
    ap.nx = p.nx;
    ap.ny = p.ny;
    ap.nz = p.nz;
    ap.pitch = ctx->pitch;
    ap.bx0 = p.buffer[0]; ap.bx1 = p.nx - p.buffer[0];
    ap.by0 = p.buffer[1]; ap.by1 = p.ny - p.buffer[1];
    ap.bz0 = p.buffer[2]; ap.bz1 = p.nz - p.buffer[2];
    ap.scale = (float)c->out_scale;
    // stats: [energy, max |correction|, max |M| before, non-finite, sum Mx, sum My]
    auto stats = [&](int buf, const float* pred, double st[6]) -> int {
        ap.mx = ctx->m[buf][0];
        ap.my = ctx->m[buf][1];
        ap.pred = pred;
        HIP_TRY(nmr::launch_apply(ap, ctx->stream));
        HIP_TRY(nmr::launch_stats_reduce(c->apart, nab, c->dstat, ctx->stream));
        HIP_TRY(hipMemcpyAsync(st, c->dstat, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return NMR_OK;
    };
    // the net's stream waits on the context's stream through explicit syncs
    HIP_TRY(hipStreamSynchronize(c->stream));
    double cur[6];
    if ((rc = stats(ctx->cur, nullptr, cur))) return rc;
    if (signal) {
        signal[0] = cur[4];
        signal[1] = cur[5];
    }
    int64_t nfall = 0, why[3] = {0, 0, 0};
    std::vector<double> sig8((size_t)(factor + 1) * 2);
    const float vinv = (float)(1.0 / c->vscale);
    // the z-marching kernels stage (Mx, My) from the state: the feature
    // tensor's other channels (pore, velocity) are written once per run
    const bool state_fed = c->conv == NMR_CONV_ZMARCH || c->conv == NMR_CONV_ZMARCH_SPARSE;
    auto features = [&](int buf) -> int {
        nmr::FeatureArgs f{};
        f.mx = ctx->m[buf][0];
        f.my = ctx->m[buf][1];
        f.mask = ctx->mask;
        f.vx = ctx->vx;
        f.vy = ctx->vy;
        f.vz = ctx->vz;
        f.out = c->t.x;
        f.nx = p.nx;
        f.ny = p.ny;
        f.nz = p.nz;
        f.pitch = ctx->pitch;
        f.pitch1 = ctx->pitch1;
        f.vinv = ctx->vel ? vinv : 0.f;
        HIP_TRY(nmr::launch_features(f, ctx->stream));
        return NMR_OK;
    };
    if (state_fed && ncoarse > 0 && (rc = features(ctx->cur))) return rc;
    for (int64_t k = 0; k < ncoarse; k++) {
        const int in = ctx->cur, out = in ^ 1;
        double gsum = 0.0;
        for (int s = 0; s < factor; s++) gsum += grad_sched ? grad_sched[3 * (k * factor + s) + 2] : p.grad[2];
        if (!(p.flags & NMR_F_PRECESSION)) gsum = 0.0;
        if (grad_sched)
            for (int s = 0; s < factor; s++) {
                const double* g = grad_sched + 3 * (k * factor + s);
                if (g[0] != 0.0 || g[1] != 0.0 || !std::isfinite(g[2]))
                    return fail(NMR_ERR_UNSUPPORTED, "ncomp 2 takes a finite z gradient only");
            }
        // coarse step into the other buffer
        if ((rc = nmri::sweep_step(ctx, coarse_coef(p, factor, &gsum), true, in, out))) return rc;
        if (!state_fed && (rc = features(in))) return rc;
        // the net runs on the context's stream
        hipStream_t keep_stream = c->stream;
        c->stream = ctx->stream;
        rc = state_fed ? forward(c, true, ctx->m[in][0], ctx->m[in][1], ctx->pitch) : forward(c, true);
        c->stream = keep_stream;
        if (rc) return rc;
        double nxt[6];
        if ((rc = stats(out, c->t.p, nxt))) return rc;
        const bool finite = nxt[3] == 0.0, energy = nxt[0] <= cur[0] * (1.0 + tol), size = nxt[1] <= tol * cur[2];
        if (finite && energy && size) {
            ctx->cur = out;
            std::copy(nxt, nxt + 6, cur);
        } else {
            nfall += 1;
            why[!finite ? 0 : (!size ? 2 : 1)] += 1;
            // plain steps from the kept state (buffer `in` is untouched)
            std::vector<double> gs;
            if (grad_sched) gs.assign(grad_sched + 3 * k * factor, grad_sched + 3 * (k + 1) * factor);
            ctx->sig_valid = false; // accepted coarse steps changed the state
            if ((rc = nmr_hip_run(ctx, factor, grad_sched ? gs.data() : nullptr, sig8.data()))) return rc;
            if ((rc = stats(ctx->cur, nullptr, cur))) return rc;
        }
        if (signal) {
            signal[2 * (k + 1)] = cur[4];
            signal[2 * (k + 1) + 1] = cur[5];
        }
    }
    if (fallbacks) {
        fallbacks[0] = nfall;
        fallbacks[1] = why[0];
        fallbacks[2] = why[1];
        fallbacks[3] = why[2];
    }
    return NMR_OK;
}

} // extern "C"
