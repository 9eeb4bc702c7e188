// fkgpu.cpp -- implementation of the C ABI declared in include/fkgpu.h.
//
// Host-side responsibilities: lattice hopping template (built once, kept in
// HBM), argument validation with the reference's error convention turned into
// status codes, buffer management, kernel dispatch by matrix size, HIP events
// for timing and HIP graph capture of move steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fkgpu.h"
#include "fk_internal.h"
#include "fk_rng.h"

struct fkgpu_ctx {
    int refs = 1;  // the owner plus one per live chain set
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cu = 0;
    int N = 0;
    int lat_kind = -1, lat_ndim = 0, lat_L = 0;
    int solver = -1;  // forced solver kind (fkgpu_set_solver / fkgpu_autotune), -1: by size
    struct mixed_bufs* mx = nullptr;  // buffers of fkgpu_eig_logz_mixed (kept between calls)
    double* d_hop = nullptr;
    double* d_hopv = nullptr;  // hopping value of each neighbour slot (N x zmax)
    int32_t* d_nbr = nullptr;
    int32_t* d_z = nullptr;
    int zmax = 0;
    // this device's static scratch (fkgpu_eig_logz_device_noalloc), resolved at creation
    double* noalloc_scratch = nullptr;
    size_t noalloc_doubles = 0;
    std::string noalloc_err;
    std::vector<double> hop;
    std::vector<int32_t> nbr, z;
    std::vector<double> hopv;
    std::vector<double> hopq;  // per-slot hopping value when it is the same at every site, else empty
    // grow-on-demand scratch for the host-pointer entry points
    struct buf { void* p = nullptr; size_t n = 0; };
    buf b_f, b_beta, b_U, b_mu, b_ev, b_lz, b_ce, b_cf, b_ws;
    buf b_vh, b_tau, b_d, b_e, b_Z, b_scr, b_m1, b_m2, b_m3, b_m4, b_m5, b_out, b_w, b_c1, b_c2, b_c3, b_c4;
};

// Measurement records of a chain set (fkgpu_chains_measure_enable): per-chain series of
// E, d2E and E_c, running means and optional histories, all on the device.
struct chain_meas_bufs {
    int cap = 0, flags = 0;
    long long* count = nullptr;
    double *e = nullptr, *d2e = nullptr, *ce = nullptr;  // B x cap
    double *esum = nullptr, *d2sum = nullptr;            // B
    double* spec_mean = nullptr;                         // B x N
    long long* focc_sum = nullptr;                       // B x N
    double* spec_hist = nullptr;                         // B x cap x N
    uint8_t* focc_hist = nullptr;                        // B x cap x N
    double *ipr = nullptr, *ipr_hist = nullptr;          // B x N, B x cap x N
    // stiffness / conductivity on nw frequencies with Lorentzian width `offset`
    int nw = 0;
    double offset = 0.05;
    double *wgrid = nullptr, *fermi = nullptr, *JV = nullptr, *mJ = nullptr, *tdiag = nullptr;
    double *stiff = nullptr, *cond = nullptr;                   // B, B x nw (this measurement)
    double *stiff_hist = nullptr, *cond_hist = nullptr;         // B x cap, B x cap x nw
    double *stiff_sum = nullptr, *cond_sum = nullptr;           // B, B x nw
    double* evec_hist = nullptr;                                // B x cap x N x N
    // decomposition of the current configurations, for chain sets that keep no
    // eigenvalues (stochastic trace) or no eigenvectors (IPR / stiffness outside the
    // rank-one mode)
    int32_t* f32 = nullptr;
    double *ev = nullptr, *lz = nullptr, *vh = nullptr, *tau = nullptr, *d = nullptr, *eo = nullptr;
    double *V = nullptr, *scr = nullptr;
};

static void meas_free(chain_meas_bufs* m) {
    if (!m) return;
    void* ps[] = {m->count, m->e, m->d2e, m->ce, m->esum, m->d2sum, m->spec_mean, m->focc_sum, m->spec_hist,
                  m->focc_hist, m->ipr, m->ipr_hist, m->f32, m->ev, m->lz, m->vh, m->tau, m->d, m->eo, m->V, m->scr,
                  m->wgrid, m->fermi, m->JV, m->mJ, m->tdiag, m->stiff, m->cond, m->stiff_hist, m->cond_hist,
                  m->stiff_sum, m->cond_sum, m->evec_hist};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete m;
}

struct fkgpu_chains {
    fkgpu_ctx* ctx = nullptr;
    int B = 0, N = 0, kind = -1, grid = 0;
    uint8_t* f = nullptr;
    int32_t* nf = nullptr;
    double* logz = nullptr;
    double* spectrum = nullptr;
    uint32_t* mt = nullptr;
    double *beta = nullptr, *U = nullptr, *mu_c = nullptr, *mu_f = nullptr;
    long long *naccept = nullptr, *nsolve = nullptr;
    double* ws = nullptr;
    size_t ws_stride = 0;
    sweep_args args{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    size_t trace_cap = 0;
    int32_t *t_kind = nullptr, *t_a = nullptr, *t_b = nullptr;
    double *t_u = nullptr, *t_logw = nullptr, *t_logz = nullptr;
    int8_t* t_acc = nullptr;
    // stochastic-trace mode
    int kpm = 0, kR = 0, kM = 0, kjack = 1;
    uint64_t kseed = 0;
    move_prop* k_props = nullptr;
    int32_t *k_fprop = nullptr, *k_skip = nullptr;
    double *k_ab = nullptr, *k_coef = nullptr, *k_mom = nullptr, *k_lz = nullptr;
    // how a proposal's dlogZ is estimated: local Chebyshev moments + Matsubara sum
    // (kl_M moments, kl_n0 frequencies; kl_s2, kl_s4 the tail sums of (2n+1)^-2, ^-4)
    // or the difference of the two configurations' stochastic traces
    int kdelta = FKGPU_KPM_DELTA_LOCAL, kl_M = 0, kl_n0 = 0;
    double kl_s2 = 0.0, kl_s4 = 0.0;
    double* k_lmom = nullptr;
    bool dense_ok = true;
    // mixed-precision mode (FP32 reduction + FP64 refinement)
    int mixed = 0;
    struct mixed_bufs* mx = nullptr;
    // rank-one update mode: eigenvectors kept on the device, full re-decomposition
    // every fu_refresh steps
    int fu = 0, fu_refresh = 0;
    double *fu_V = nullptr, *fu_Q1 = nullptr, *fu_Q2 = nullptr, *fu_T = nullptr;
    double *fu_vh = nullptr, *fu_tau = nullptr, *fu_d = nullptr, *fu_e = nullptr, *fu_scr = nullptr;
    int32_t *fu_upd = nullptr, *fu_f32 = nullptr;
    int32_t* fu_skip = nullptr;      // per chain: the periodic re-decomposition is not due this step
    long long* fu_count = nullptr;   // move steps since the rank-one mode was set (device counter)
    // chains on several lattices (fkgpu_chains_create_multi): the other contexts (held) and
    // the per-chain lattice index
    std::vector<fkgpu_ctx*> mctx;
    int32_t* lat = nullptr;
    int nlat = 0;
    // several dense FP64 move steps per launch (k_sweep_q work queue)
    int spl = 16;                    // move steps per launch
    unsigned* q_ctr = nullptr;
    int32_t *q_ring = nullptr, *q_err = nullptr;  // ring: spl x B slots
    long long q_ticks = 0;           // bound of one wait, in wall-clock ticks
    // on-device measurements and the graph of one measured sweep
    chain_meas_bufs* ms = nullptr;
    std::vector<double> m_wgrid{0.0};  // conductivity frequencies of the stiffness measure
    double m_offset = 0.05;            // their Lorentzian width (cond_offset)
    hipGraphExec_t mexec = nullptr;
    hipGraph_t mgraph = nullptr;
    int mgraph_len = 0;
};

// Captured graphs depend on the move set, the f-f terms and the update mode: drop them
// whenever one of these changes.
static void drop_graphs(fkgpu_chains* ch) {
    if (ch->gexec) { (void)hipGraphExecDestroy(ch->gexec); ch->gexec = nullptr; }
    if (ch->graph) { (void)hipGraphDestroy(ch->graph); ch->graph = nullptr; }
    ch->graph_steps = 0;
    if (ch->mexec) { (void)hipGraphExecDestroy(ch->mexec); ch->mexec = nullptr; }
    if (ch->mgraph) { (void)hipGraphDestroy(ch->mgraph); ch->mgraph = nullptr; }
    ch->mgraph_len = 0;
}

namespace {

int fail(fkgpu_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}
int hipfail(fkgpu_ctx* ctx, hipError_t e, const char* where) {
    return fail(ctx, FKGPU_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}
#define FK_HIP(ctx, call)                                   \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return hipfail(ctx, _e, #call); \
    } while (0)

template <class T>
int ensure(fkgpu_ctx* ctx, fkgpu_ctx::buf& b, size_t count, T** out) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    if (b.n < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) return fail(ctx, FKGPU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        b.n = bytes;
    }
    *out = (T*)b.p;
    return FKGPU_OK;
}

bool finite_all(const double* p, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// One workgroup per configuration / chain; the global-memory solver keeps one
// N x N workspace per workgroup, so its grid is one batch of resident groups
// and the kernel is launched repeatedly over slices of the batch.
// Solver kind for the context's lattice: the forced / autotuned one when it fits N.
// Solver variant for a batch of `batch` matrices (0: unknown): the forced or autotuned one,
// else by size; for N <= 64 a small batch (a single chain, the plain drop-in) takes the
// one-wave solver, whose solve has the lowest latency, and larger batches the 16 x 16
// block-cyclic one, which has the higher throughput.
int pick_kind(const fkgpu_ctx* ctx, int batch = 0) {
    if (ctx->solver >= 0 && ctx->N <= fk_solver_nmax(ctx->solver)) return ctx->solver;
    if (ctx->N <= 64 && batch > 0 && batch < 128) return 6;
    return fk_solver_kind(ctx->N);
}

int grid_for(fkgpu_ctx* ctx, int kind, int B) {
    (void)ctx;
    (void)kind;
    return B;
}

// Site index order of hypercubic_lattice::index_to_pos: last coordinate fastest.
// The dense N x N template (used by the dense eigensolvers) is kept up to N = 4096 only;
// the neighbour table and its per-slot values (hopv) describe the lattice at any size.
constexpr int kDenseTemplateMax = 4096;

void build_hypercubic(fkgpu_ctx* ctx, int kind, int D, int L, double t, double tp) {
    int n = 1;
    for (int i = 0; i < D; ++i) n *= L;
    ctx->N = n;
    const bool dense = n <= kDenseTemplateMax;
    ctx->hop.assign(dense ? (size_t)n * n : 0, 0.0);
    ctx->zmax = 2 * D + (kind == FKGPU_LATTICE_TRIANGULAR ? 2 : 0);
    ctx->nbr.assign((size_t)n * ctx->zmax, 0);
    ctx->hopv.assign((size_t)n * ctx->zmax, 0.0);
    ctx->z.assign(n, 0);
    int stride[3] = {1, 1, 1};
    for (int d = D - 2; d >= 0; --d) stride[d] = stride[d + 1] * L;
    for (int i = 0; i < n; ++i) {
        int p[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) p[d] = (i / stride[d]) % L;
        auto add = [&](int j, double v) {
            if (dense) ctx->hop[(size_t)i * n + j] += v;
            ctx->hopv[(size_t)i * ctx->zmax + ctx->z[i]] = v;  // L >= 3: no two slots share a site
            ctx->nbr[(size_t)i * ctx->zmax + ctx->z[i]++] = j;
        };
        for (int d = 0; d < D; ++d) {
            int pl = p[d] > 0 ? p[d] - 1 : L - 1;
            int pr = p[d] < L - 1 ? p[d] + 1 : 0;
            add(i + (pl - p[d]) * stride[d], -t);
            add(i + (pr - p[d]) * stride[d], -t);
        }
        if (kind == FKGPU_LATTICE_TRIANGULAR) {
            int xl = p[0] > 0 ? p[0] - 1 : L - 1, yl = p[1] > 0 ? p[1] - 1 : L - 1;
            int xr = p[0] < L - 1 ? p[0] + 1 : 0, yr = p[1] < L - 1 ? p[1] + 1 : 0;
            add(xl * stride[0] + yl, -tp);
            add(xr * stride[0] + yr, -tp);
        }
    }
}

static const double* hopq_ptr(const fkgpu_ctx* ctx) { return ctx->hopq.empty() ? nullptr : ctx->hopq.data(); }

int upload_lattice(fkgpu_ctx* ctx) {
    if (ctx->d_hop) (void)hipFree(ctx->d_hop);
    if (ctx->d_hopv) (void)hipFree(ctx->d_hopv);
    if (ctx->d_nbr) (void)hipFree(ctx->d_nbr);
    if (ctx->d_z) (void)hipFree(ctx->d_z);
    ctx->d_hop = ctx->d_hopv = nullptr;
    ctx->d_nbr = ctx->d_z = nullptr;
    size_t n = ctx->N;
    const bool dense = !ctx->hop.empty();
    // per-slot values from the dense template (a CSR lattice may list a site in several slots);
    // stencil lattices above the dense range filled hopv themselves
    if (dense) ctx->hopv.assign(ctx->nbr.size(), 0.0);
    for (size_t i = 0; i < n && dense; ++i)
        for (int q = 0; q < ctx->z[i]; ++q) {
            const int j = ctx->nbr[i * ctx->zmax + q];
            // the dense template sums all slots pointing at j; split the value evenly so that
            // the slots of one row add up to the template entry
            int mult = 0;
            for (int q2 = 0; q2 < ctx->z[i]; ++q2) mult += ctx->nbr[i * ctx->zmax + q2] == j;
            ctx->hopv[i * ctx->zmax + q] = ctx->hop[i * n + j] / mult;
        }
    ctx->hopq.assign(ctx->hopv.begin(), ctx->hopv.begin() + std::min<size_t>(ctx->zmax, ctx->hopv.size()));
    for (size_t i = 0; i < n && !ctx->hopq.empty(); ++i) {
        bool same = ctx->z[i] == ctx->zmax;
        for (int q = 0; same && q < ctx->zmax; ++q) same = ctx->hopv[i * ctx->zmax + q] == ctx->hopq[q];
        if (!same) ctx->hopq.clear();
    }
    FK_HIP(ctx, hipMalloc(&ctx->d_hopv, sizeof(double) * ctx->hopv.size()));
    FK_HIP(ctx, hipMemcpy(ctx->d_hopv, ctx->hopv.data(), sizeof(double) * ctx->hopv.size(), hipMemcpyHostToDevice));
    if (dense) {
        FK_HIP(ctx, hipMalloc(&ctx->d_hop, sizeof(double) * n * n));
        FK_HIP(ctx, hipMemcpy(ctx->d_hop, ctx->hop.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
    }
    FK_HIP(ctx, hipMalloc(&ctx->d_nbr, sizeof(int32_t) * ctx->nbr.size()));
    FK_HIP(ctx, hipMalloc(&ctx->d_z, sizeof(int32_t) * n));
    FK_HIP(ctx, hipMemcpy(ctx->d_nbr, ctx->nbr.data(), sizeof(int32_t) * ctx->nbr.size(), hipMemcpyHostToDevice));
    FK_HIP(ctx, hipMemcpy(ctx->d_z, ctx->z.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    return FKGPU_OK;
}

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

}  // namespace

extern "C" {

int fkgpu_create(int device, fkgpu_ctx** out) {
    if (!out) return FKGPU_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FKGPU_EHIP;
    if (device < 0 || device >= ndev) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = new fkgpu_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return FKGPU_EHIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
    if (ctx->num_cu <= 0) ctx->num_cu = 256;
    // the allocation-free entry point's scratch: its device address, once per device
    const hipError_t se = fk_noalloc_scratch(device, &ctx->noalloc_scratch, &ctx->noalloc_doubles);
    if (se != hipSuccess) ctx->noalloc_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    *out = ctx;
    return FKGPU_OK;
}

static void mixed_free(mixed_bufs* m);

static void ctx_release(fkgpu_ctx* ctx) {
    if (!ctx || --ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    mixed_free(ctx->mx);
    for (auto* b : {&ctx->b_f, &ctx->b_beta, &ctx->b_U, &ctx->b_mu, &ctx->b_ev, &ctx->b_lz, &ctx->b_ce, &ctx->b_cf,
                    &ctx->b_ws, &ctx->b_vh, &ctx->b_tau, &ctx->b_d, &ctx->b_e, &ctx->b_Z, &ctx->b_scr, &ctx->b_m1,
                    &ctx->b_m2, &ctx->b_m3, &ctx->b_m4, &ctx->b_m5, &ctx->b_out, &ctx->b_w, &ctx->b_c1, &ctx->b_c2,
                    &ctx->b_c3, &ctx->b_c4})
        if (b->p) (void)hipFree(b->p);
    if (ctx->d_hop) (void)hipFree(ctx->d_hop);
    if (ctx->d_hopv) (void)hipFree(ctx->d_hopv);
    if (ctx->d_nbr) (void)hipFree(ctx->d_nbr);
    if (ctx->d_z) (void)hipFree(ctx->d_z);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// Destroying a context while chain sets built on it are alive defers the
// release until the last of them is destroyed.
void fkgpu_destroy(fkgpu_ctx* ctx) { ctx_release(ctx); }

const char* fkgpu_last_error(const fkgpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
void* fkgpu_stream(fkgpu_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
int fkgpu_msize(const fkgpu_ctx* ctx) { return ctx ? ctx->N : 0; }

int fkgpu_set_lattice(fkgpu_ctx* ctx, int kind, int ndim, int L, double t, double tp) {
    if (!ctx) return FKGPU_EINVAL;
    if (kind != FKGPU_LATTICE_HYPERCUBIC && kind != FKGPU_LATTICE_TRIANGULAR)
        return fail(ctx, FKGPU_EINVAL, "unknown lattice kind");
    if (ndim < 1 || ndim > 3) return fail(ctx, FKGPU_EINVAL, "ndim must be 1..3");
    if (kind == FKGPU_LATTICE_TRIANGULAR && ndim != 2) return fail(ctx, FKGPU_EINVAL, "triangular lattice is 2-D");
    if (L < 3) return fail(ctx, FKGPU_EINVAL, "lattice size L must be >= 3");
    if (!std::isfinite(t) || !std::isfinite(tp)) return fail(ctx, FKGPU_EINVAL, "non-finite hopping");
    long n = 1;
    for (int i = 0; i < ndim; ++i) n *= L;
    if (n > 16384) return fail(ctx, FKGPU_EUNSUPPORTED, "lattice volume above 16384 sites");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    build_hypercubic(ctx, kind, ndim, L, t, tp);
    ctx->lat_kind = kind;
    ctx->lat_ndim = ndim;
    ctx->lat_L = L;
    return upload_lattice(ctx);
}

int fkgpu_set_lattice_csr(fkgpu_ctx* ctx, int n, const int32_t* row_ptr, const int32_t* col, const double* val) {
    if (!ctx) return FKGPU_EINVAL;
    if (n <= 0 || !row_ptr || !col || !val) return fail(ctx, FKGPU_EINVAL, "invalid CSR arguments");
    if (n > 4096) return fail(ctx, FKGPU_EUNSUPPORTED, "matrix size above 4096");
    std::vector<double> h((size_t)n * n, 0.0);
    std::vector<int32_t> z(n, 0);
    int zmax = 0;
    for (int i = 0; i < n; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) return fail(ctx, FKGPU_EINVAL, "row_ptr not monotone");
        for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
            int j = col[p];
            if (j < 0 || j >= n || !std::isfinite(val[p])) return fail(ctx, FKGPU_EINVAL, "invalid CSR entry");
            h[(size_t)i * n + j] += val[p];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (h[(size_t)i * n + j] != h[(size_t)j * n + i]) return fail(ctx, FKGPU_EINVAL, "hopping matrix not symmetric");
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j)
            if (i != j && h[(size_t)i * n + j] != 0.0) ++z[i];
        zmax = std::max(zmax, (int)z[i]);
    }
    if (zmax == 0) zmax = 1;
    ctx->N = n;
    ctx->hop.swap(h);
    ctx->zmax = zmax;
    ctx->z.assign(n, 0);
    ctx->nbr.assign((size_t)n * zmax, 0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j && ctx->hop[(size_t)i * n + j] != 0.0) ctx->nbr[(size_t)i * zmax + ctx->z[i]++] = j;
    ctx->lat_kind = 2;
    ctx->lat_ndim = 0;
    ctx->lat_L = 0;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    return upload_lattice(ctx);
}

static int check_params(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                        const double* mu_c) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (batch <= 0 || !f || !beta || !U || !mu_c) return fail(ctx, FKGPU_EINVAL, "null argument or empty batch");
    if (!finite_all(beta, batch) || !finite_all(U, batch) || !finite_all(mu_c, batch))
        return fail(ctx, FKGPU_EINVAL, "non-finite beta, U or mu_c");
    const size_t n = (size_t)batch * ctx->N;
    for (size_t i = 0; i < n; ++i)
        if (f[i] != 0 && f[i] != 1) return fail(ctx, FKGPU_EINVAL, "occupations must be 0 or 1");
    return FKGPU_OK;
}

// Device-resident eigenpairs of `batch` configurations given on the host.
struct eig_dev {
    int32_t* f;
    double *beta, *U, *mu, *ev, *lz, *ce, *cf, *V;
};
static int eig_pairs_dev(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                         const double* mu_c, bool vectors, eig_dev& o) {
    const int N = ctx->N;
    if (fk_solver_kind(N) < 0) return fail(ctx, FKGPU_EUNSUPPORTED, "dense eigensolver implemented for N <= 2048");
    const size_t BN = (size_t)batch * N, BNN = BN * N;
    int r;
    if ((r = ensure(ctx, ctx->b_f, BN, &o.f)) || (r = ensure(ctx, ctx->b_beta, batch, &o.beta)) ||
        (r = ensure(ctx, ctx->b_U, batch, &o.U)) || (r = ensure(ctx, ctx->b_mu, batch, &o.mu)) ||
        (r = ensure(ctx, ctx->b_ev, BN, &o.ev)) || (r = ensure(ctx, ctx->b_lz, batch, &o.lz)) ||
        (r = ensure(ctx, ctx->b_ce, BN, &o.ce)) || (r = ensure(ctx, ctx->b_cf, BN, &o.cf)))
        return r;
    o.V = nullptr;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(o.f, f, sizeof(int32_t) * BN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(o.beta, beta, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(o.U, U, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(o.mu, mu_c, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    const int kind = pick_kind(ctx, batch);
    const int grid = grid_for(ctx, kind, batch);
    eig_args a{};
    a.B = batch;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = o.f;
    a.beta = o.beta;
    a.U = o.U;
    a.mu_c = o.mu;
    a.evals = o.ev;
    a.logz = o.lz;
    a.cexp = o.ce;
    a.cfermi = o.cf;
    a.ws_stride = fk_solver_ws_doubles(kind, N);
    if (fk_solver_needs_ws(kind) && (r = ensure(ctx, ctx->b_ws, (size_t)grid * a.ws_stride, &a.ws))) return r;
    double *vh = nullptr, *tau = nullptr, *dd = nullptr, *ee = nullptr, *Z = nullptr, *scr = nullptr;
    if (vectors) {
        if (N > 1024) return fail(ctx, FKGPU_EUNSUPPORTED, "eigenvectors are implemented for N <= 1024");
        if ((r = ensure(ctx, ctx->b_vh, BNN, &vh)) || (r = ensure(ctx, ctx->b_tau, BN, &tau)) ||
            (r = ensure(ctx, ctx->b_d, BN, &dd)) || (r = ensure(ctx, ctx->b_e, BN, &ee)) ||
            (r = ensure(ctx, ctx->b_Z, BNN, &Z)) || (r = ensure(ctx, ctx->b_scr, fk_evecs_scratch_doubles(batch, N), &scr)))
            return r;
        a.vout = vh;
        a.tauout = tau;
        a.d_out = dd;
        a.e_out = ee;
    }
    FK_HIP(ctx, fk_launch_eig(kind, a, grid, st));
    if (vectors) {
        FK_HIP(ctx, fk_launch_evecs(batch, N, dd, ee, o.ev, vh, tau, Z, scr, st));
        o.V = Z;
    }
    return FKGPU_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 blocked reduction (solver kind 5), FP64 tridiagonal
// eigenvalues and eigenvectors, FP64 Rayleigh-Ritz refinement (fk_mixed.hip),
// FP64 solve of the configurations the refinement flags.
// ---------------------------------------------------------------------------
struct mixed_bufs {
    int B = 0, N = 0;
    double *ws = nullptr, *vh = nullptr, *tau = nullptr, *d = nullptr, *e = nullptr, *evh = nullptr, *lzh = nullptr;
    double *X = nullptr, *scr = nullptr, *Q = nullptr, *ev = nullptr, *lz = nullptr, *ws64 = nullptr;
    int32_t* fb = nullptr;
    unsigned long long* nflag = nullptr;
    size_t ws_stride = 0, ws64_stride = 0;
    int kind64 = -1;
};

static void mixed_free(mixed_bufs* m) {
    if (!m) return;
    void* ptrs[] = {m->ws, m->vh, m->tau, m->d, m->e, m->evh, m->lzh, m->X, m->scr, m->Q, m->ev, m->lz, m->ws64,
                    m->fb, m->nflag};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
}

static hipError_t mixed_alloc(fkgpu_ctx* ctx, int B, int N, mixed_bufs** out) {
    mixed_bufs* m = new mixed_bufs();
    m->B = B;
    m->N = N;
    m->ws_stride = fk_solver_ws_doubles(FK_SOLVER_MIXED32, N);
    m->kind64 = pick_kind(ctx);
    m->ws64_stride = fk_solver_needs_ws(m->kind64) ? fk_solver_ws_doubles(m->kind64, N) : 0;
    const size_t BN = (size_t)B * N, BNN = BN * N;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(dalloc(&m->ws, (size_t)B * m->ws_stride));
    chk(dalloc(&m->vh, BNN));
    chk(dalloc(&m->tau, BN));
    chk(dalloc(&m->d, BN));
    chk(dalloc(&m->e, BN));
    chk(dalloc(&m->evh, BN));
    chk(dalloc(&m->lzh, B));
    chk(dalloc(&m->X, BNN));
    chk(dalloc(&m->scr, fk_evecs_scratch_doubles(B, N)));
    chk(dalloc(&m->Q, BN * fk_mixed_cluster_max()));
    chk(dalloc(&m->ev, BN));
    chk(dalloc(&m->lz, B));
    if (m->ws64_stride) chk(dalloc(&m->ws64, (size_t)B * m->ws64_stride));
    chk(dalloc(&m->fb, B));
    chk(dalloc(&m->nflag, 1));
    if (e == hipSuccess) e = hipMemset(m->nflag, 0, sizeof(unsigned long long));
    // eigenvector-stage inputs of skipped items must hold finite numbers
    if (e == hipSuccess) e = hipMemset(m->d, 0, sizeof(double) * BN);
    if (e == hipSuccess) e = hipMemset(m->e, 0, sizeof(double) * BN);
    if (e == hipSuccess) e = hipMemset(m->evh, 0, sizeof(double) * BN);
    if (e != hipSuccess) { mixed_free(m); return e; }
    *out = m;
    return hipSuccess;
}

// Mixed-precision eigenvalues + logZ of B configurations (device f, int32) into m->ev, m->lz.
static hipError_t mixed_eig_launch(fkgpu_ctx* ctx, mixed_bufs* m, int B, const int32_t* df, const double* beta,
                                   const double* U, const double* mu, const int32_t* skip, hipStream_t st) {
    const int N = m->N;
    const size_t BNN = (size_t)B * N * N;
    eig_args a{};
    a.B = B;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = df;
    a.beta = beta;
    a.U = U;
    a.mu_c = mu;
    a.evals = m->evh;
    a.logz = m->lzh;
    a.ws = m->ws;
    a.ws_stride = m->ws_stride;
    a.skip = skip;
    a.vout = m->vh;
    a.tauout = m->tau;
    a.d_out = m->d;
    a.e_out = m->e;
    hipError_t e = fk_launch_eig(FK_SOLVER_MIXED32, a, B, st);
    if (e != hipSuccess) return e;
    e = fk_launch_evecs(B, N, m->d, m->e, m->evh, m->vh, m->tau, m->X, m->scr, st, skip);
    if (e != hipSuccess) return e;
    e = fk_launch_mixed_refine(B, N, m->X, df, U, mu, ctx->d_nbr, ctx->d_z, ctx->zmax, ctx->d_hopv, m->scr,
                               m->scr + BNN, m->scr + 2 * BNN, m->Q, beta, m->ev, m->lz, m->fb, skip, m->nflag, st);
    if (e != hipSuccess) return e;
    // FP64 solve where the refinement could not certify its result
    eig_args f{};
    f.B = B;
    f.N = N;
    f.hop = ctx->d_hop;
    f.f = df;
    f.beta = beta;
    f.U = U;
    f.mu_c = mu;
    f.evals = m->ev;
    f.logz = m->lz;
    f.ws = m->ws64;
    f.ws_stride = m->ws64_stride;
    f.skip = m->fb;
    return fk_launch_eig(m->kind64, f, B, st);
}

int fkgpu_eig_logz_mixed(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                         const double* mu_c, double* evals, double* logZ, int64_t* nfallback) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!evals || !logZ) return fail(ctx, FKGPU_EINVAL, "null output");
    const int N = ctx->N;
    if (N > fk_solver_nmax(FK_SOLVER_MIXED32)) return fail(ctx, FKGPU_EUNSUPPORTED, "mixed-precision path implemented for N <= 1024");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t BN = (size_t)batch * N;
    int32_t* df;
    double *dbeta, *dU, *dmu;
    if ((r = ensure(ctx, ctx->b_f, BN, &df)) || (r = ensure(ctx, ctx->b_beta, batch, &dbeta)) ||
        (r = ensure(ctx, ctx->b_U, batch, &dU)) || (r = ensure(ctx, ctx->b_mu, batch, &dmu)))
        return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(df, f, sizeof(int32_t) * BN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dbeta, beta, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dU, U, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dmu, mu_c, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    if (ctx->mx && (ctx->mx->N != N || ctx->mx->B < batch || ctx->mx->kind64 != pick_kind(ctx))) {
        mixed_free(ctx->mx);
        ctx->mx = nullptr;
    }
    if (!ctx->mx) FK_HIP(ctx, mixed_alloc(ctx, batch, N, &ctx->mx));
    mixed_bufs* m = ctx->mx;
    FK_HIP(ctx, hipMemsetAsync(m->nflag, 0, sizeof(unsigned long long), st));
    hipError_t e = mixed_eig_launch(ctx, m, batch, df, dbeta, dU, dmu, nullptr, st);
    unsigned long long nf = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(evals, m->ev, sizeof(double) * BN, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(logZ, m->lz, sizeof(double) * batch, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&nf, m->nflag, sizeof(nf), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hipfail(ctx, e, "mixed-precision eigensolve");
    if (nfallback) *nfallback = (int64_t)nf;
    return FKGPU_OK;
}

int fkgpu_set_solver(fkgpu_ctx* ctx, int variant) {
    if (!ctx) return FKGPU_EINVAL;
    if (variant < FKGPU_SOLVER_AUTO || variant >= FKGPU_SOLVER_VARIANTS)
        return fail(ctx, FKGPU_EINVAL, "unknown solver variant");
    if (variant >= 0 && ctx->N > 0 && ctx->N > fk_solver_nmax(variant))
        return fail(ctx, FKGPU_EUNSUPPORTED, "solver variant does not fit this lattice size");
    ctx->solver = variant;
    return FKGPU_OK;
}

int fkgpu_get_solver(const fkgpu_ctx* ctx) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return FKGPU_ENOLATTICE;
    return pick_kind(ctx);
}

int fkgpu_autotune(fkgpu_ctx* ctx, int batch, double* ms) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (batch <= 0) return fail(ctx, FKGPU_EINVAL, "empty batch");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const int N = ctx->N;
    const size_t BN = (size_t)batch * N;
    // synthetic half-filled configurations (a fixed LCG), beta = 1, U = 2, mu_c = 1
    std::vector<int32_t> f(BN);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < BN; ++i) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        f[i] = (int32_t)(x >> 63);
    }
    std::vector<double> one(batch, 1.0), two(batch, 2.0);
    int32_t* df;
    double *dbeta, *dU, *dmu, *dev, *dlz;
    int r;
    if ((r = ensure(ctx, ctx->b_f, BN, &df)) || (r = ensure(ctx, ctx->b_beta, batch, &dbeta)) ||
        (r = ensure(ctx, ctx->b_U, batch, &dU)) || (r = ensure(ctx, ctx->b_mu, batch, &dmu)) ||
        (r = ensure(ctx, ctx->b_ev, BN, &dev)) || (r = ensure(ctx, ctx->b_lz, batch, &dlz)))
        return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(df, f.data(), sizeof(int32_t) * BN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dbeta, one.data(), sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dU, two.data(), sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dmu, one.data(), sizeof(double) * batch, hipMemcpyHostToDevice, st));
    hipEvent_t e0, e1;
    FK_HIP(ctx, hipEventCreate(&e0));
    FK_HIP(ctx, hipEventCreate(&e1));
    int best = -1;
    double best_ms = 0.0;
    for (int v = 0; v < FKGPU_SOLVER_VARIANTS; ++v) {
        if (ms) ms[v] = -1.0;
        if (N > fk_solver_nmax(v)) continue;
        const int grid = grid_for(ctx, v, batch);
        eig_args a{};
        a.B = batch;
        a.N = N;
        a.hop = ctx->d_hop;
        a.f = df;
        a.beta = dbeta;
        a.U = dU;
        a.mu_c = dmu;
        a.evals = dev;
        a.logz = dlz;
        a.ws_stride = fk_solver_ws_doubles(v, N);
        if (fk_solver_needs_ws(v) && (r = ensure(ctx, ctx->b_ws, (size_t)grid * a.ws_stride, &a.ws))) break;
        hipError_t he = fk_launch_eig(v, a, grid, st);
        if (he == hipSuccess) he = hipEventRecord(e0, st);
        for (int it = 0; it < 3 && he == hipSuccess; ++it) he = fk_launch_eig(v, a, grid, st);
        if (he == hipSuccess) he = hipEventRecord(e1, st);
        if (he == hipSuccess) he = hipEventSynchronize(e1);
        float t = 0.f;
        if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
        if (he != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return hipfail(ctx, he, "autotune launch");
        }
        const double tv = t / 3.0;
        if (ms) ms[v] = tv;
        if (best < 0 || tv < best_ms) { best = v; best_ms = tv; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    if (best < 0) return fail(ctx, FKGPU_EUNSUPPORTED, "no solver variant fits this lattice size");
    ctx->solver = best;
    return best;
}

int fkgpu_eig_logz_device(fkgpu_ctx* ctx, int batch, const int32_t* d_f, const double* d_beta, const double* d_U,
                          const double* d_mu_c, double* d_evals, double* d_logZ) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (batch <= 0 || !d_f || !d_beta || !d_U || !d_mu_c || !d_evals || !d_logZ)
        return fail(ctx, FKGPU_EINVAL, "null argument or empty batch");
    if (fk_solver_kind(ctx->N) < 0) return fail(ctx, FKGPU_EUNSUPPORTED, "dense eigensolver implemented for N <= 2048");
    const int N = ctx->N;
    const int kind = pick_kind(ctx, batch);
    const int grid = grid_for(ctx, kind, batch);
    eig_args a{};
    a.B = batch;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = d_f;
    a.beta = d_beta;
    a.U = d_U;
    a.mu_c = d_mu_c;
    a.evals = d_evals;
    a.logz = d_logZ;
    a.ws_stride = fk_solver_ws_doubles(kind, N);
    if (fk_solver_needs_ws(kind)) {
        int r = ensure(ctx, ctx->b_ws, (size_t)grid * a.ws_stride, &a.ws);
        if (r) return r;
    }
    FK_HIP(ctx, hipSetDevice(ctx->device));
    FK_HIP(ctx, fk_launch_eig(kind, a, grid, ctx->stream));
    return FKGPU_OK;
}

int fkgpu_eig_logz_device_noalloc(fkgpu_ctx* ctx, int batch, const int32_t* d_f, const double* d_beta,
                                  const double* d_U, const double* d_mu_c, double* d_evals, double* d_logZ) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (batch <= 0 || !d_f || !d_beta || !d_U || !d_mu_c || !d_evals || !d_logZ)
        return fail(ctx, FKGPU_EINVAL, "null argument or empty batch");
    if (fk_solver_kind(ctx->N) < 0) return fail(ctx, FKGPU_EUNSUPPORTED, "dense eigensolver implemented for N <= 2048");
    const int N = ctx->N;
    const int kind = pick_kind(ctx, batch);
    eig_args a{};
    a.B = batch;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = d_f;
    a.beta = d_beta;
    a.U = d_U;
    a.mu_c = d_mu_c;
    a.evals = d_evals;
    a.logz = d_logZ;
    a.ws_stride = fk_solver_ws_doubles(kind, N);
    if (fk_solver_needs_ws(kind)) {
        // one workgroup per configuration, each with its own slice of the static scratch
        if (!ctx->noalloc_scratch) return fail(ctx, FKGPU_EHIP, "static scratch unavailable: " + ctx->noalloc_err);
        const size_t need = (size_t)batch * a.ws_stride;
        if (need > ctx->noalloc_doubles) {
            const size_t most = ctx->noalloc_doubles / a.ws_stride;
            return fail(ctx, FKGPU_EUNSUPPORTED,
                        "batch of " + std::to_string(batch) + " configurations at N=" + std::to_string(N) + " needs " +
                            std::to_string(need * sizeof(double) >> 20) + " MiB of scratch, more than the " +
                            std::to_string(FK_NOALLOC_SCRATCH_BYTES >> 20) + " MiB static scratch (at most " +
                            std::to_string(most) + " configurations per call at this N)");
        }
        a.ws = ctx->noalloc_scratch;
    }
    FK_HIP(ctx, hipSetDevice(ctx->device));
    FK_HIP(ctx, fk_launch_eig(kind, a, batch, ctx->stream));
    return FKGPU_OK;
}

int fkgpu_eig_logz(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                   const double* mu_c, int want_vectors, double* evals, double* evecs, double* logZ,
                   double* cached_exp, double* cached_fermi) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!evals || !logZ) return fail(ctx, FKGPU_EINVAL, "null output");
    if (fk_solver_kind(ctx->N) < 0) return fail(ctx, FKGPU_EUNSUPPORTED, "dense eigensolver implemented for N <= 2048");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    if (want_vectors) {
        if (!evecs) return fail(ctx, FKGPU_EINVAL, "null evecs with want_vectors");
        eig_dev o;
        if ((r = eig_pairs_dev(ctx, batch, f, beta, U, mu_c, true, o))) return r;
        const size_t BN = (size_t)batch * ctx->N;
        hipStream_t st = ctx->stream;
        FK_HIP(ctx, hipMemcpyAsync(evals, o.ev, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
        FK_HIP(ctx, hipMemcpyAsync(logZ, o.lz, sizeof(double) * batch, hipMemcpyDeviceToHost, st));
        FK_HIP(ctx, hipMemcpyAsync(evecs, o.V, sizeof(double) * BN * ctx->N, hipMemcpyDeviceToHost, st));
        if (cached_exp) FK_HIP(ctx, hipMemcpyAsync(cached_exp, o.ce, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
        if (cached_fermi) FK_HIP(ctx, hipMemcpyAsync(cached_fermi, o.cf, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
        FK_HIP(ctx, hipStreamSynchronize(st));
        return FKGPU_OK;
    }
    const int N = ctx->N;
    const size_t BN = (size_t)batch * N;
    int32_t* df;
    double *dbeta, *dU, *dmu, *dev, *dlz, *dce = nullptr, *dcf = nullptr;
    if ((r = ensure(ctx, ctx->b_f, BN, &df)) || (r = ensure(ctx, ctx->b_beta, batch, &dbeta)) ||
        (r = ensure(ctx, ctx->b_U, batch, &dU)) || (r = ensure(ctx, ctx->b_mu, batch, &dmu)) ||
        (r = ensure(ctx, ctx->b_ev, BN, &dev)) || (r = ensure(ctx, ctx->b_lz, batch, &dlz)))
        return r;
    if (cached_exp || cached_fermi) {
        if ((r = ensure(ctx, ctx->b_ce, BN, &dce)) || (r = ensure(ctx, ctx->b_cf, BN, &dcf))) return r;
    }
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(df, f, sizeof(int32_t) * BN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dbeta, beta, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dU, U, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dmu, mu_c, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    const int kind = pick_kind(ctx, batch);
    const int grid = grid_for(ctx, kind, batch);
    eig_args a{};
    a.B = batch;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = df;
    a.beta = dbeta;
    a.U = dU;
    a.mu_c = dmu;
    a.evals = dev;
    a.logz = dlz;
    a.cexp = dce;
    a.cfermi = dcf;
    a.ws_stride = fk_solver_ws_doubles(kind, N);
    if (fk_solver_needs_ws(kind)) {
        if ((r = ensure(ctx, ctx->b_ws, (size_t)grid * a.ws_stride, &a.ws))) return r;
    }
    FK_HIP(ctx, fk_launch_eig(kind, a, grid, st));
    FK_HIP(ctx, hipMemcpyAsync(evals, dev, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipMemcpyAsync(logZ, dlz, sizeof(double) * batch, hipMemcpyDeviceToHost, st));
    if (cached_exp) FK_HIP(ctx, hipMemcpyAsync(cached_exp, dce, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    if (cached_fermi) FK_HIP(ctx, hipMemcpyAsync(cached_fermi, dcf, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

// ---------------------------------------------------------------------------
// Measures and adjoint
// ---------------------------------------------------------------------------
int fkgpu_measure_ipr(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                      const double* mu_c, double* evals, double* ipr) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!ipr) return fail(ctx, FKGPU_EINVAL, "null output");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    eig_dev o;
    if ((r = eig_pairs_dev(ctx, batch, f, beta, U, mu_c, true, o))) return r;
    const int N = ctx->N;
    const size_t BN = (size_t)batch * N;
    double* dout;
    if ((r = ensure(ctx, ctx->b_m1, BN, &dout))) return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, fk_launch_ipr(batch, N, o.V, dout, st));
    FK_HIP(ctx, hipMemcpyAsync(ipr, dout, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    if (evals) FK_HIP(ctx, hipMemcpyAsync(evals, o.ev, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_measure_stiffness(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                            const double* mu_c, int nw, const double* wgrid, double offset, double* stiffness,
                            double* cond) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!stiffness || nw < 0 || (nw > 0 && (!wgrid || !cond)) || !std::isfinite(offset))
        return fail(ctx, FKGPU_EINVAL, "invalid stiffness arguments");
    if (ctx->lat_kind != FKGPU_LATTICE_HYPERCUBIC || ctx->lat_ndim < 2)
        return fail(ctx, FKGPU_EUNSUPPORTED, "stiffness needs a hypercubic lattice of dimension >= 2");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    eig_dev o;
    if ((r = eig_pairs_dev(ctx, batch, f, beta, U, mu_c, true, o))) return r;
    const int N = ctx->N;
    const size_t BN = (size_t)batch * N, BNN = BN * N;
    double *JV, *mJ, *td, *st_d, *cd, *wg;
    if ((r = ensure(ctx, ctx->b_m1, BNN, &JV)) || (r = ensure(ctx, ctx->b_m2, BNN, &mJ)) ||
        (r = ensure(ctx, ctx->b_m3, BN, &td)) || (r = ensure(ctx, ctx->b_m4, batch, &st_d)) ||
        (r = ensure(ctx, ctx->b_m5, (size_t)batch * (nw > 0 ? nw : 1), &cd)) ||
        (r = ensure(ctx, ctx->b_w, (size_t)(nw > 0 ? nw : 1), &wg)))
        return r;
    hipStream_t st = ctx->stream;
    if (nw > 0) FK_HIP(ctx, hipMemcpyAsync(wg, wgrid, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    int stride0 = 1;
    for (int d = 1; d < ctx->lat_ndim; ++d) stride0 *= ctx->lat_L;
    // T_x uses hopping 1 whatever t is (stiffness.hpp:59, t_ = 1)
    FK_HIP(ctx, fk_launch_stiffness(batch, N, stride0, ctx->lat_L, 1.0, o.V, o.ev, o.cf, nw, wg, offset, JV, mJ, td,
                                    st_d, cd, st));
    FK_HIP(ctx, hipMemcpyAsync(stiffness, st_d, sizeof(double) * batch, hipMemcpyDeviceToHost, st));
    if (nw > 0) FK_HIP(ctx, hipMemcpyAsync(cond, cd, sizeof(double) * batch * nw, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_dos(fkgpu_ctx* ctx, int nspectra, const double* spectra, int nw, const double* wgrid, double eta,
              double* dos) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (nspectra <= 0 || nw <= 0 || !spectra || !wgrid || !dos || !(eta > 0.0))
        return fail(ctx, FKGPU_EINVAL, "invalid DOS arguments");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const int N = ctx->N;
    double *sp, *wg, *out;
    int r;
    if ((r = ensure(ctx, ctx->b_m1, (size_t)nspectra * N, &sp)) || (r = ensure(ctx, ctx->b_w, nw, &wg)) ||
        (r = ensure(ctx, ctx->b_m2, (size_t)nspectra * nw, &out)))
        return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(sp, spectra, sizeof(double) * nspectra * N, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(wg, wgrid, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, fk_launch_dos(nspectra, N, sp, nw, wg, eta, out, st));
    FK_HIP(ctx, hipMemcpyAsync(dos, out, sizeof(double) * nspectra * nw, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_ipr_w(fkgpu_ctx* ctx, int nspectra, const double* spectra, const double* ipr, int nw, const double* wgrid,
                double eta, double* out) {
    if (!ctx) return FKGPU_EINVAL;
    if (ctx->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "no lattice set");
    if (nspectra <= 0 || nw <= 0 || !spectra || !ipr || !wgrid || !out || !(eta > 0.0))
        return fail(ctx, FKGPU_EINVAL, "invalid IPR(w) arguments");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const int N = ctx->N;
    const size_t SN = (size_t)nspectra * N;
    double *sp, *wg, *o, *ip;
    int r;
    if ((r = ensure(ctx, ctx->b_m1, SN, &sp)) || (r = ensure(ctx, ctx->b_m3, SN, &ip)) ||
        (r = ensure(ctx, ctx->b_w, nw, &wg)) || (r = ensure(ctx, ctx->b_m2, (size_t)nspectra * nw, &o)))
        return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(sp, spectra, sizeof(double) * SN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(ip, ipr, sizeof(double) * SN, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(wg, wgrid, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, fk_launch_ipr_w(nspectra, N, sp, ip, nw, wg, eta, o, st));
    FK_HIP(ctx, hipMemcpyAsync(out, o, sizeof(double) * nspectra * nw, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_adjoint(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                  const double* mu_c, double* logZ, double* dens, double* grads, double* rho) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!logZ || !dens || !grads) return fail(ctx, FKGPU_EINVAL, "null output");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    eig_dev o;
    if ((r = eig_pairs_dev(ctx, batch, f, beta, U, mu_c, true, o))) return r;
    const int N = ctx->N;
    const size_t BN = (size_t)batch * N;
    double *dd, *bt, *btp, *out, *rf = nullptr;
    if ((r = ensure(ctx, ctx->b_m1, BN, &dd)) || (r = ensure(ctx, ctx->b_m3, BN, &bt)) ||
        (r = ensure(ctx, ctx->b_m4, BN, &btp)) || (r = ensure(ctx, ctx->b_out, (size_t)batch * 8, &out)))
        return r;
    if (rho && (r = ensure(ctx, ctx->b_m2, BN * N, &rf))) return r;
    const int nt_bonds = ctx->lat_kind == 2 ? ctx->zmax : 2 * ctx->lat_ndim;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, fk_launch_adjoint(batch, N, o.V, o.cf, o.f, o.beta, ctx->d_nbr, ctx->d_z, ctx->zmax, nt_bonds, dd, bt,
                                  btp, out, rf, st));
    FK_HIP(ctx, hipMemcpyAsync(logZ, o.lz, sizeof(double) * batch, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipMemcpyAsync(dens, dd, sizeof(double) * BN, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipMemcpyAsync(grads, out, sizeof(double) * batch * 8, hipMemcpyDeviceToHost, st));
    if (rho) FK_HIP(ctx, hipMemcpyAsync(rho, rf, sizeof(double) * BN * N, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

// Newton iteration on mu_c for a target conduction filling N_c / N, with N_c and
// dN_c/dmu_c = beta sum_k f_k (1 - f_k) from the adjoint (dOmega/dmu_c = -N_c),
// safeguarded by bisection on a bracket from the spectrum bounds of H.
int fkgpu_tune_mu(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U, double* mu_c,
                  double filling, double tol, int maxit, int32_t* iters) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    if (!(filling > 0.0 && filling < 1.0) || !(tol > 0.0) || maxit < 1)
        return fail(ctx, FKGPU_EINVAL, "filling must lie in (0, 1), tol > 0, maxit >= 1");
    const int N = ctx->N;
    double rad = 0.0;  // largest Gershgorin radius of the hopping matrix
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int q = 0; q < ctx->z[i]; ++q) s += std::fabs(ctx->hopv[(size_t)i * ctx->zmax + q]);
        rad = std::max(rad, s);
    }
    std::vector<double> lo(batch), hi(batch), mu(mu_c, mu_c + batch), lz(batch), dens((size_t)batch * N),
        gr((size_t)batch * 8);
    std::vector<int32_t> done(batch, 0), it(batch, 0);
    for (int b = 0; b < batch; ++b) {
        const double span = rad + std::fabs(U[b]) + 60.0 / beta[b];
        lo[b] = std::min(0.0, U[b]) - span;
        hi[b] = std::max(0.0, U[b]) + span;
        if (!(mu[b] > lo[b] && mu[b] < hi[b])) mu[b] = 0.5 * (lo[b] + hi[b]);
    }
    const double target = filling * N;
    for (int iter = 0; iter < maxit; ++iter) {
        if ((r = fkgpu_adjoint(ctx, batch, f, beta, U, mu.data(), lz.data(), dens.data(), gr.data(), nullptr)))
            return r;
        int left = 0;
        for (int b = 0; b < batch; ++b) {
            if (done[b]) continue;
            const double nc = gr[(size_t)b * 8 + 4], dn = gr[(size_t)b * 8 + 5];
            it[b] = iter + 1;
            if (std::fabs(nc - target) <= tol * N) { done[b] = 1; continue; }
            if (nc < target) lo[b] = mu[b]; else hi[b] = mu[b];
            double next = dn > 0.0 ? mu[b] + (target - nc) / dn : 0.5 * (lo[b] + hi[b]);
            if (!(next > lo[b] && next < hi[b])) next = 0.5 * (lo[b] + hi[b]);
            mu[b] = next;
            ++left;
        }
        if (!left) break;
    }
    for (int b = 0; b < batch; ++b) {
        mu_c[b] = mu[b];
        if (iters) iters[b] = done[b] ? it[b] : -it[b];
    }
    return FKGPU_OK;
}

// ---------------------------------------------------------------------------
// Chebyshev paths
// ---------------------------------------------------------------------------
namespace {
// chebyshev_eval (include/fk_mc/chebyshev.hpp:15-63): angle grid linspace(0,1,G),
// x_i = -cos(pi theta_i), T_k(x_i) = cos(k acos x_i), trapezoid in theta.
struct cheb_quad {
    int M, G;
    std::vector<double> th, x;
    cheb_quad(int max_moment, int grid) : G(grid) {
        M = max_moment + max_moment % 2;
        th.resize(G);
        x.resize(G);
        for (int i = 0; i < G; ++i) {
            th[i] = G > 1 ? (double)i / (double)(G - 1) : 0.0;
            x[i] = -std::cos(M_PI * th[i]);
        }
    }
    double moment(const std::vector<double>& v, int k) const {
        double s = 0.0;
        for (int i = 0; i + 1 < G; ++i)
            s += (v[i + 1] * std::cos(k * std::acos(x[i + 1])) + v[i] * std::cos(k * std::acos(x[i]))) * (th[i + 1] - th[i]);
        return s * 0.5;
    }
};
}  // namespace

// Extreme eigenvalues of a symmetric tridiagonal (alpha, beta: n and n-1 entries) by
// bisection on Sturm counts.
static double tri_sturm_count(const std::vector<double>& al, const std::vector<double>& be, int n, double x) {
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
        const double b2 = i > 0 ? be[i - 1] * be[i - 1] : 0.0;
        q = (al[i] - x) - (i > 0 ? b2 / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0.0) ++c;
    }
    return c;
}
static void tri_extremes(const std::vector<double>& al, const std::vector<double>& be, int n, double* lo_out,
                         double* hi_out) {
    double glo = 1e300, ghi = -1e300;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(be[i]) : 0.0);
        glo = std::min(glo, al[i] - r);
        ghi = std::max(ghi, al[i] + r);
    }
    for (int which = 0; which < 2; ++which) {
        double lo = glo, hi = ghi;  // count(lo) <= k < count(hi), k = 0 or n - 1
        const int k = which == 0 ? 0 : n - 1;
        for (int it = 0; it < 200 && hi - lo > 0.0; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (mid <= lo || mid >= hi) break;
            if (tri_sturm_count(al, be, n, mid) > k) hi = mid;
            else lo = mid;
        }
        (which == 0 ? *lo_out : *hi_out) = 0.5 * (lo + hi);
    }
}

// lambda_min, lambda_max of H = hop + diag(d) by Lanczos on the device (the ARPACK calls of
// src/configuration.cpp:99-100, which ask for full accuracy): steps in chunks of 64 until both
// extreme Ritz values change by at most 4 eps max(|lambda|, 1) over a chunk, the Krylov space
// is exhausted (beta = 0) or N steps were taken.
static int lanczos_bounds(fkgpu_ctx* ctx, const double* d_diag, double* e_min, double* e_max, int* steps) {
    const int N = ctx->N;
    const int cap = std::min(N, 6000), chunk = 64;
    double *work, *al_d;
    int r;
    if ((r = ensure(ctx, ctx->b_m4, (size_t)3 * N + 2, &work)) || (r = ensure(ctx, ctx->b_m5, (size_t)2 * cap + 2, &al_d)))
        return r;
    double* be_d = al_d + cap + 1;
    hipStream_t st = ctx->stream;
    std::vector<double> al(cap + 1), be(cap + 1);
    double plo = 0.0, phi = 0.0;
    int done = 0;
    for (;;) {
        const int it = std::min(chunk, cap - done);
        FK_HIP(ctx, fk_launch_lanczos(N, ctx->d_nbr, ctx->d_z, ctx->zmax, ctx->d_hopv, d_diag, done == 0, it, work,
                                      al_d, be_d, st));
        FK_HIP(ctx, hipMemcpyAsync(al.data(), al_d, sizeof(double) * (done + it), hipMemcpyDeviceToHost, st));
        FK_HIP(ctx, hipMemcpyAsync(be.data(), be_d, sizeof(double) * (done + it), hipMemcpyDeviceToHost, st));
        FK_HIP(ctx, hipStreamSynchronize(st));
        int n = done + it;
        bool exhausted = false;
        for (int j = done; j < n; ++j)
            if (!(be[j] > 1e-14 * (std::fabs(al[j]) + 1.0))) { n = j + 1; exhausted = true; break; }
        double lo, hi;
        tri_extremes(al, be, n, &lo, &hi);
        const double tol = 4.0 * 2.220446049250313e-16;
        const bool conv = done > 0 && std::fabs(lo - plo) <= tol * std::max(std::fabs(lo), 1.0) &&
                          std::fabs(hi - phi) <= tol * std::max(std::fabs(hi), 1.0);
        plo = lo;
        phi = hi;
        done = n;
        if (conv || exhausted || done >= cap) break;
    }
    *e_min = plo;
    *e_max = phi;
    if (steps) *steps = done;
    return FKGPU_OK;
}

int fkgpu_cheb_logz(fkgpu_ctx* ctx, const int32_t* f, double beta, double U, double mu_c, int cheb_size,
                    int grid_size, double* moments, double* logZ, double* ab) {
    int r = check_params(ctx, 1, f, &beta, &U, &mu_c);
    if (r) return r;
    if (cheb_size < 2 || grid_size < 2 || !moments || !logZ || !ab) return fail(ctx, FKGPU_EINVAL, "invalid Chebyshev arguments");
    const int N = ctx->N;
    if (N > 16384) return fail(ctx, FKGPU_EUNSUPPORTED, "exact-trace Chebyshev keeps two N x N matrices: N <= 16384");
    cheb_quad cq(cheb_size, grid_size);
    const int M = cq.M;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t NN = (size_t)N * N;
    double *T0, *T1, *parts, *traces, *diag;
    if ((r = ensure(ctx, ctx->b_c1, NN, &T0)) || (r = ensure(ctx, ctx->b_c2, NN, &T1)) ||
        (r = ensure(ctx, ctx->b_c3, 512, &parts)) || (r = ensure(ctx, ctx->b_c4, (size_t)3 * M + N, &traces)))
        return r;
    diag = traces + 3 * M;
    std::vector<double> dg(N);
    for (int i = 0; i < N; ++i) dg[i] = -mu_c + U * f[i];
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(diag, dg.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    // extreme eigenvalues by Lanczos (ARPACK in the reference)
    double e_min = 0.0, e_max = 0.0;
    int lsteps = 0;
    if ((r = lanczos_bounds(ctx, diag, &e_min, &e_max, &lsteps))) return r;
    const double a = (e_max - e_min) / 2.0, b = (e_max + e_min) / 2.0;
    if (!(a > 0.0)) return fail(ctx, FKGPU_EINVAL, "degenerate spectrum: e_max == e_min");
    double trX = 0.0;
    for (int i = 0; i < N; ++i) trX += (dg[i] - b) / a;
    FK_HIP(ctx, hipMemsetAsync(traces, 0, sizeof(double) * 3 * M, st));
    FK_HIP(ctx, fk_cheb_exact_moments(N, ctx->d_nbr, ctx->d_z, ctx->zmax, ctx->d_hopv, diag, a, b, M, T0, T1, parts,
                                      nullptr, traces, st));
    std::vector<double> tr(3 * M);
    FK_HIP(ctx, hipMemcpyAsync(tr.data(), traces, sizeof(double) * 3 * M, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    // moment bookkeeping of src/configuration.cpp:217-288
    std::vector<double> mom(M, 0.0);
    std::vector<char> is_set(M, 0);
    mom[0] = 1.0; is_set[0] = 1;
    mom[1] = tr[1] / N; is_set[1] = 1;
    for (int m = 2; m <= M / 2; ++m) {
        if (!is_set[m]) { mom[m] = tr[m] / N; is_set[m] = 1; }
        int k = 2 * m - 1;
        if (k < M && k >= M / 2) {
            mom[k] = (tr[M + m] * 2.0 - trX) / N;
            is_set[k] = 1;
            if (k != M - 1) {
                ++k;
                mom[k] = tr[2 * M + m] / N * 2.0 - 1.0;
                is_set[k] = 1;
            }
        }
    }
    std::vector<double> fv(cq.G);
    for (int i = 0; i < cq.G; ++i) fv[i] = N * std::log(1.0 + std::exp(-beta * (a * cq.x[i] + b)));
    double s = cq.moment(fv, 0);
    for (int m = 1; m < M; ++m) s += 2.0 * cq.moment(fv, m) * mom[m];
    for (int m = 0; m < M; ++m) moments[m] = mom[m];
    *logZ = s;
    ab[0] = a; ab[1] = b; ab[2] = e_min; ab[3] = e_max;
    return FKGPU_OK;
}

int fkgpu_kpm_logz(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                   const double* mu_c, int R, int M, uint64_t seed, int jackson, double* logZ, double* moments,
                   double* ab) {
    int r = check_params(ctx, batch, f, beta, U, mu_c);
    if (r) return r;
    const int N = ctx->N;
    if (!logZ || R < 1 || M < 2 || (M & 1) || M > 512) return fail(ctx, FKGPU_EINVAL, "invalid KPM arguments");
    if (N > 4096 || ctx->zmax > 8) return fail(ctx, FKGPU_EUNSUPPORTED, "KPM path needs N <= 4096 and <= 8 neighbours");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    // Gershgorin bounds and Chebyshev-Gauss coefficients of log(1+e^{-beta(a x + b)})
    std::vector<double> abv(2 * (size_t)batch), coef((size_t)batch * M), abo(2 * (size_t)batch);
    const int K = std::max(2 * M, 64);
    for (int c = 0; c < batch; ++c) {
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < N; ++i) {
            double rad = 0.0;
            for (int q = 0; q < ctx->z[i]; ++q) rad += std::fabs(ctx->hopv[(size_t)i * ctx->zmax + q]);
            const double dgi = -mu_c[c] + U[c] * f[(size_t)c * N + i];
            lo = std::min(lo, dgi - rad);
            hi = std::max(hi, dgi + rad);
        }
        double a = 0.5 * (hi - lo) * 1.01, b = 0.5 * (hi + lo);
        if (!(a > 0.0)) a = 1.0;
        abv[2 * c] = 1.0 / a;
        abv[2 * c + 1] = b;
        abo[2 * c] = a;
        abo[2 * c + 1] = b;
        for (int m = 0; m < M; ++m) {
            double sacc = 0.0;
            for (int k = 0; k < K; ++k) {
                const double th = M_PI * (k + 0.5) / K;
                const double x = std::cos(th);
                const double y = beta[c] * (a * x + b);
                const double F = y > 0 ? std::log1p(std::exp(-y)) : -y + std::log1p(std::exp(y));
                sacc += F * std::cos(m * th);
            }
            coef[(size_t)c * M + m] = (m == 0 ? 1.0 : 2.0) * sacc / K;
        }
    }
    int32_t* df;
    double *dU, *dmu, *dab, *dcoef, *dmom, *dmuo, *dlz;
    if ((r = ensure(ctx, ctx->b_f, (size_t)batch * N, &df)) || (r = ensure(ctx, ctx->b_U, batch, &dU)) ||
        (r = ensure(ctx, ctx->b_mu, batch, &dmu)) || (r = ensure(ctx, ctx->b_c1, 2 * (size_t)batch, &dab)) ||
        (r = ensure(ctx, ctx->b_c2, (size_t)batch * M, &dcoef)) ||
        (r = ensure(ctx, ctx->b_c3, (size_t)batch * R * M, &dmom)) ||
        (r = ensure(ctx, ctx->b_c4, (size_t)batch * M, &dmuo)) || (r = ensure(ctx, ctx->b_lz, batch, &dlz)))
        return r;
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipMemcpyAsync(df, f, sizeof(int32_t) * batch * N, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dU, U, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dmu, mu_c, sizeof(double) * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dab, abv.data(), sizeof(double) * 2 * batch, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(dcoef, coef.data(), sizeof(double) * batch * M, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, fk_launch_kpm(batch, N, R, M, df, dU, dmu, dab, ctx->d_nbr, ctx->d_z, ctx->zmax, ctx->d_hopv, hopq_ptr(ctx), seed, 0,
                              nullptr, dcoef, jackson, dmom, dmuo, dlz, st));
    FK_HIP(ctx, hipMemcpyAsync(logZ, dlz, sizeof(double) * batch, hipMemcpyDeviceToHost, st));
    if (moments) FK_HIP(ctx, hipMemcpyAsync(moments, dmuo, sizeof(double) * batch * M, hipMemcpyDeviceToHost, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    if (ab) for (int i = 0; i < 2 * batch; ++i) ab[i] = abo[i];
    return FKGPU_OK;
}

// ---------------------------------------------------------------------------
// Chains
// ---------------------------------------------------------------------------
void fkgpu_chains_destroy(fkgpu_chains* ch) {
    if (!ch) return;
    if (ch->ctx) {
        (void)hipSetDevice(ch->ctx->device);
        (void)hipStreamSynchronize(ch->ctx->stream);
    }
    drop_graphs(ch);
    meas_free(ch->ms);
    void* ptrs[] = {ch->f, ch->nf, ch->logz, ch->spectrum, ch->mt, ch->beta, ch->U, ch->mu_c, ch->mu_f,
                    ch->naccept, ch->nsolve, ch->ws, ch->t_kind, ch->t_a, ch->t_b, ch->t_u, ch->t_logw,
                    ch->t_logz, ch->t_acc, ch->k_props, ch->k_fprop, ch->k_skip, ch->k_ab, ch->k_coef,
                    ch->k_mom, ch->k_lz, ch->k_lmom, ch->q_ctr, ch->q_ring, ch->q_err, ch->lat};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (ch->ev0) (void)hipEventDestroy(ch->ev0);
    if (ch->ev1) (void)hipEventDestroy(ch->ev1);
    mixed_free(ch->mx);
    void* fptrs[] = {ch->fu_V, ch->fu_Q1, ch->fu_Q2, ch->fu_T, ch->fu_vh, ch->fu_tau, ch->fu_d, ch->fu_e, ch->fu_scr,
                     ch->fu_upd, ch->fu_f32, ch->fu_skip, ch->fu_count};
    for (void* p : fptrs)
        if (p) (void)hipFree(p);
    fkgpu_ctx* ctx = ch->ctx;
    std::vector<fkgpu_ctx*> others = ch->mctx;
    delete ch;
    ctx_release(ctx);
    for (fkgpu_ctx* c : others) ctx_release(c);
}

}  // extern "C"

// A chain set of B chains whose f / spectrum rows have N entries (the lattice size, or the
// largest one of a multi-lattice set), stepped by the dense solver `kind`.  With
// initial_eig the starting logZ and spectra come from ctx's lattice; otherwise they are
// left at zero for the caller.
static int chains_create_impl(fkgpu_ctx* ctx, int N, int kind, int nchains, const int32_t* f_init,
                              const double* beta, const double* U, const double* mu_c, const double* mu_f,
                              const uint32_t* seeds, bool initial_eig, fkgpu_chains** out) {
    int r = 0;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    const int B = nchains;
    fkgpu_chains* ch = new fkgpu_chains();
    ch->ctx = ctx;
    ctx->refs += 1;
    ch->B = B;
    ch->N = N;
    ch->kind = kind;
    ch->dense_ok = ch->kind >= 0;
    ch->grid = grid_for(ctx, ch->kind, B);
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(dalloc(&ch->f, (size_t)B * N));
    chk(dalloc(&ch->nf, B));
    chk(dalloc(&ch->logz, B));
    chk(dalloc(&ch->spectrum, (size_t)B * N));
    chk(dalloc(&ch->mt, (size_t)B * fkrng::MT_WORDS));
    chk(dalloc(&ch->beta, B));
    chk(dalloc(&ch->U, B));
    chk(dalloc(&ch->mu_c, B));
    chk(dalloc(&ch->mu_f, B));
    chk(dalloc(&ch->naccept, B));
    chk(dalloc(&ch->nsolve, B));
    ch->ws_stride = ch->dense_ok ? fk_solver_ws_doubles(ch->kind, N) : 0;
    if (ch->dense_ok && fk_solver_needs_ws(ch->kind)) chk(dalloc(&ch->ws, (size_t)ch->grid * ch->ws_stride));
    chk(hipEventCreate(&ch->ev0));
    chk(hipEventCreate(&ch->ev1));
    chk(dalloc(&ch->q_ctr, 2));
    chk(dalloc(&ch->q_ring, (size_t)ch->spl * B));
    chk(dalloc(&ch->q_err, 1));
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess || khz <= 0)
            khz = 100000;
        ch->q_ticks = 30LL * 1000LL * khz;  // 30 s: far beyond any single move step
    }
    if (e != hipSuccess) {
        fkgpu_chains_destroy(ch);
        return fail(ctx, FKGPU_ENOMEM, std::string("chain allocation: ") + hipGetErrorString(e));
    }
    std::vector<uint8_t> f8((size_t)B * N);
    std::vector<int32_t> nf(B, 0);
    for (int c = 0; c < B; ++c)
        for (int i = 0; i < N; ++i) {
            f8[(size_t)c * N + i] = (uint8_t)f_init[(size_t)c * N + i];
            nf[c] += f_init[(size_t)c * N + i];
        }
    std::vector<uint32_t> mts((size_t)B * fkrng::MT_WORDS);
    for (int c = 0; c < B; ++c) fkrng::mt_seed(&mts[(size_t)c * fkrng::MT_WORDS], seeds[c]);
    std::vector<long long> zeros(B, 0);
    hipStream_t st = ctx->stream;
    chk(hipMemcpyAsync(ch->f, f8.data(), f8.size(), hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->nf, nf.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->mt, mts.data(), sizeof(uint32_t) * mts.size(), hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->beta, beta, sizeof(double) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->U, U, sizeof(double) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->mu_c, mu_c, sizeof(double) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->mu_f, mu_f, sizeof(double) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->naccept, zeros.data(), sizeof(long long) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->nsolve, zeros.data(), sizeof(long long) * B, hipMemcpyHostToDevice, st));
    chk(hipMemsetAsync(ch->q_err, 0, sizeof(int32_t), st));
    if (e != hipSuccess) {
        fkgpu_chains_destroy(ch);
        return hipfail(ctx, e, "chain upload");
    }
    // initial logZ and spectrum of every chain
    int32_t* df;
    if (!initial_eig) {
        chk(hipMemsetAsync(ch->logz, 0, sizeof(double) * B, st));
        chk(hipMemsetAsync(ch->spectrum, 0, sizeof(double) * B * N, st));
    }
    if (!initial_eig) {
    } else if ((r = ensure(ctx, ctx->b_f, (size_t)B * N, &df))) {
        fkgpu_chains_destroy(ch);
        return r;
    } else {
    chk(hipMemcpyAsync(df, f_init, sizeof(int32_t) * B * N, hipMemcpyHostToDevice, st));
    eig_args a{};
    a.B = B;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = df;
    a.beta = ch->beta;
    a.U = ch->U;
    a.mu_c = ch->mu_c;
    a.evals = ch->spectrum;
    a.logz = ch->logz;
    a.ws = ch->ws;
    a.ws_stride = ch->ws_stride;
    if (ch->dense_ok) {
        chk(fk_launch_eig(ch->kind, a, ch->grid, st));
    } else {  // logZ comes from fkgpu_chains_set_kpm
        chk(hipMemsetAsync(ch->logz, 0, sizeof(double) * B, st));
        chk(hipMemsetAsync(ch->spectrum, 0, sizeof(double) * B * N, st));
    }
    }
    chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        fkgpu_chains_destroy(ch);
        return hipfail(ctx, e, "initial eigensolve");
    }
    sweep_args& s = ch->args;
    s.B = B;
    s.N = N;
    s.hop = ctx->d_hop;
    s.nbr = ctx->d_nbr;
    s.z = ctx->d_z;
    s.zmax = ctx->zmax;
    s.f = ch->f;
    s.nf = ch->nf;
    s.logz = ch->logz;
    s.spectrum = ch->spectrum;
    s.mt = ch->mt;
    s.beta = ch->beta;
    s.U = ch->U;
    s.mu_c = ch->mu_c;
    s.mu_f = ch->mu_f;
    s.naccept = ch->naccept;
    s.nsolve = ch->nsolve;
    s.ws = ch->ws;
    s.ws_stride = ch->ws_stride;
    s.ncp = 1;
    s.cp[0] = 1.0;
    s.kinds[0] = FKGPU_MOVE_ADDREMOVE;
    *out = ch;
    return FKGPU_OK;
}

extern "C" {

int fkgpu_chains_create(fkgpu_ctx* ctx, int nchains, const int32_t* f_init, const double* beta, const double* U,
                        const double* mu_c, const double* mu_f, const uint32_t* seeds, fkgpu_chains** out) {
    if (!out) return FKGPU_EINVAL;
    *out = nullptr;
    int r = check_params(ctx, nchains, f_init, beta, U, mu_c);
    if (r) return r;
    if (!mu_f || !seeds || !finite_all(mu_f, nchains)) return fail(ctx, FKGPU_EINVAL, "invalid mu_f or seeds");
    return chains_create_impl(ctx, ctx->N, pick_kind(ctx, nchains), nchains, f_init, beta, U, mu_c, mu_f, seeds, true,
                              out);
}

int fkgpu_chains_create_multi(int ncontexts, fkgpu_ctx* const* contexts, int nchains, const int32_t* lattice,
                              const int32_t* f_init, const double* beta, const double* U, const double* mu_c,
                              const double* mu_f, const uint32_t* seeds, fkgpu_chains** out) {
    if (!out) return FKGPU_EINVAL;
    *out = nullptr;
    if (ncontexts < 1 || !contexts || !contexts[0]) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = contexts[0];
    if (ncontexts > FK_MAX_LATTICES) return fail(ctx, FKGPU_EUNSUPPORTED, "at most 8 lattices in one chain set");
    if (nchains < 1 || !lattice || !f_init || !beta || !U || !mu_c || !mu_f || !seeds)
        return fail(ctx, FKGPU_EINVAL, "null argument or empty chain set");
    int nmax = 0;
    for (int l = 0; l < ncontexts; ++l) {
        const fkgpu_ctx* c = contexts[l];
        if (!c || c->N == 0) return fail(ctx, FKGPU_ENOLATTICE, "a context without a lattice");
        if (c->device != ctx->device) return fail(ctx, FKGPU_EINVAL, "contexts on different devices");
        if (fk_solver_kind(c->N) < 0 || c->N > 1024)
            return fail(ctx, FKGPU_EUNSUPPORTED, "multi-lattice chain sets need N <= 1024 on every lattice");
        nmax = std::max(nmax, c->N);
    }
    // one solver for all chains: the register-resident wave-row solver when every N <= 256,
    // else the blocked HBM-workspace solver (both take N at run time)
    const int kind = nmax <= 256 ? 10 : 8;
    const int B = nchains;
    std::vector<int32_t> fpad((size_t)B * nmax, 0);
    for (int c = 0; c < B; ++c) {
        if (lattice[c] < 0 || lattice[c] >= ncontexts) return fail(ctx, FKGPU_EINVAL, "lattice index out of range");
        if (!std::isfinite(beta[c]) || !std::isfinite(U[c]) || !std::isfinite(mu_c[c]) || !std::isfinite(mu_f[c]))
            return fail(ctx, FKGPU_EINVAL, "non-finite beta, U, mu_c or mu_f");
        const int n = contexts[lattice[c]]->N;
        for (int i = 0; i < n; ++i) {
            const int32_t v = f_init[(size_t)c * nmax + i];
            if (v != 0 && v != 1) return fail(ctx, FKGPU_EINVAL, "occupations must be 0 or 1");
            fpad[(size_t)c * nmax + i] = v;
        }
    }
    fkgpu_chains* ch = nullptr;
    int r = chains_create_impl(ctx, nmax, kind, B, fpad.data(), beta, U, mu_c, mu_f, seeds, false, &ch);
    if (r) return r;
    for (int l = 1; l < ncontexts; ++l) {
        contexts[l]->refs += 1;
        ch->mctx.push_back(contexts[l]);
    }
    // starting logZ and spectra, one batched eigensolve per lattice with the chain set's solver
    // (so that a chain starts from the bits its own steps would produce)
    hipStream_t st = ctx->stream;
    std::vector<double> lz(B, 0.0), spec((size_t)B * nmax, 0.0);
    int32_t* dfl;
    double *dbl, *dul, *dml, *devl, *dlzl;
    if ((r = ensure(ctx, ctx->b_f, (size_t)B * nmax, &dfl)) || (r = ensure(ctx, ctx->b_beta, B, &dbl)) ||
        (r = ensure(ctx, ctx->b_U, B, &dul)) || (r = ensure(ctx, ctx->b_mu, B, &dml)) ||
        (r = ensure(ctx, ctx->b_ev, (size_t)B * nmax, &devl)) || (r = ensure(ctx, ctx->b_lz, B, &dlzl))) {
        fkgpu_chains_destroy(ch);
        return r;
    }
    for (int l = 0; l < ncontexts; ++l) {
        fkgpu_ctx* c = contexts[l];
        std::vector<int> idx;
        for (int q = 0; q < B; ++q)
            if (lattice[q] == l) idx.push_back(q);
        if (idx.empty()) continue;
        const int n = c->N, nb = (int)idx.size();
        std::vector<int32_t> fl((size_t)nb * n);
        std::vector<double> bl(nb), ul(nb), ml(nb), ev((size_t)nb * n), lzl(nb);
        for (int k = 0; k < nb; ++k) {
            std::copy(fpad.begin() + (size_t)idx[k] * nmax, fpad.begin() + (size_t)idx[k] * nmax + n,
                      fl.begin() + (size_t)k * n);
            bl[k] = beta[idx[k]];
            ul[k] = U[idx[k]];
            ml[k] = mu_c[idx[k]];
        }
        eig_args a{};
        a.B = nb;
        a.N = n;
        a.hop = c->d_hop;
        a.f = dfl;
        a.beta = dbl;
        a.U = dul;
        a.mu_c = dml;
        a.evals = devl;
        a.logz = dlzl;
        a.ws = ch->ws;  // B workspaces of the largest N: enough for nb <= B of this one
        a.ws_stride = ch->ws_stride;
        hipError_t he = hipMemcpyAsync(dfl, fl.data(), sizeof(int32_t) * fl.size(), hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipMemcpyAsync(dbl, bl.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipMemcpyAsync(dul, ul.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipMemcpyAsync(dml, ml.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = fk_launch_eig(kind, a, nb, st);
        if (he == hipSuccess) he = hipMemcpyAsync(ev.data(), devl, sizeof(double) * ev.size(), hipMemcpyDeviceToHost, st);
        if (he == hipSuccess) he = hipMemcpyAsync(lzl.data(), dlzl, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
        if (he == hipSuccess) he = hipStreamSynchronize(st);
        if (he != hipSuccess) {
            fkgpu_chains_destroy(ch);
            return hipfail(ctx, he, "multi-lattice initial eigensolve");
        }
        for (int k = 0; k < nb; ++k) {
            lz[idx[k]] = lzl[k];
            std::copy(ev.begin() + (size_t)k * n, ev.begin() + (size_t)(k + 1) * n, spec.begin() + (size_t)idx[k] * nmax);
        }
    }
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(dalloc(&ch->lat, B));
    chk(hipMemcpyAsync(ch->lat, lattice, sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->logz, lz.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
    chk(hipMemcpyAsync(ch->spectrum, spec.data(), sizeof(double) * spec.size(), hipMemcpyHostToDevice, st));
    chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        fkgpu_chains_destroy(ch);
        return hipfail(ctx, e, "multi-lattice chain set upload");
    }
    sweep_args& s = ch->args;
    s.lat = ch->lat;
    s.ld = nmax;
    for (int l = 0; l < ncontexts; ++l) {
        s.lN[l] = contexts[l]->N;
        s.lhop[l] = contexts[l]->d_hop;
        s.lnbr[l] = contexts[l]->d_nbr;
        s.lz[l] = contexts[l]->d_z;
        s.lzmax[l] = contexts[l]->zmax;
    }
    ch->nlat = ncontexts;
    *out = ch;
    return FKGPU_OK;
}

int fkgpu_chains_set_moves(fkgpu_chains* ch, int nmoves, const int32_t* kinds, const double* weights) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (nmoves < 1 || nmoves > 4 || !kinds || !weights) return fail(ctx, FKGPU_EINVAL, "1..4 moves required");
    double sum = 0.0;
    for (int m = 0; m < nmoves; ++m) {
        if (kinds[m] < 0 || kinds[m] > 3) return fail(ctx, FKGPU_EINVAL, "unknown move kind");
        if (!(weights[m] >= 0.0) || !std::isfinite(weights[m])) return fail(ctx, FKGPU_EINVAL, "invalid move weight");
        sum += weights[m];
    }
    if (!(sum > 0.0)) return fail(ctx, FKGPU_EINVAL, "move weights sum to zero");
    // The exchange draws a neighbour slot uniformly from the site's z neighbours, so
    // every site needs one (a CSR lattice may hold isolated sites).
    for (int m = 0; m < nmoves; ++m)
        if (kinds[m] == FKGPU_MOVE_HOP && weights[m] > 0.0) {
            std::vector<const fkgpu_ctx*> lats{ctx};
            for (const fkgpu_ctx* c : ch->mctx) lats.push_back(c);
            for (const fkgpu_ctx* c : lats)
                for (int i = 0; i < c->N; ++i)
                    if (c->z[i] == 0)
                        return fail(ctx, FKGPU_EINVAL, "nearest-neighbour exchange on a lattice with an isolated site");
        }
    if (ch->fu)
        for (int m = 0; m < nmoves; ++m)
            if (kinds[m] == FKGPU_MOVE_RANDOMIZE)
                return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one updates cover moves of one or two sites (not randomize)");
    // libstdc++ discrete_distribution::param_type::_M_initialize: normalise by
    // division, partial sums, last entry forced to 1; a single weight draws nothing.
    sweep_args& s = ch->args;
    s.ncp = nmoves;
    double acc = 0.0;
    for (int m = 0; m < nmoves; ++m) {
        acc += weights[m] / sum;
        s.cp[m] = acc;
        s.kinds[m] = kinds[m];
    }
    s.cp[nmoves - 1] = 1.0;
    if (nmoves < 2) s.ncp = 1;
    drop_graphs(ch);
    return FKGPU_OK;
}

int fkgpu_chains_set_ff(fkgpu_chains* ch, int nW, const double* W) {
    if (!ch) return FKGPU_EINVAL;
    if (nW < 0 || nW > 8 || (nW > 0 && !W) || (nW > 0 && !finite_all(W, nW)))
        return fail(ch->ctx, FKGPU_EINVAL, "W must hold 0..8 finite values");
    if (nW > 0 && ch->ctx->lat_ndim != 1) nW = 0;  // calc_ff_energy returns 0 unless ndim == 1
    ch->args.nW = nW;
    for (int i = 0; i < nW; ++i) ch->args.W[i] = W[i];
    drop_graphs(ch);
    return FKGPU_OK;
}

// One move step of every chain: the fused dense kernel, or propose / stochastic
// trace / accept for the Chebyshev (KPM) mode.
// Full eigendecomposition of every chain's current configuration into spectrum, logZ
// and the eigenvector matrices of the rank-one update mode.
// Full re-decomposition of the chains' current configurations; with skip, only of the
// chains whose skip flag is clear (all or none: the flags come from the step counter).
static hipError_t fu_refresh_launch(fkgpu_chains* ch, hipStream_t st, const int32_t* skip = nullptr) {
    fkgpu_ctx* ctx = ch->ctx;
    const int B = ch->B, N = ch->N;
    hipError_t e = fk_launch_f8to32(B * N, ch->f, ch->fu_f32, st, skip);
    if (e != hipSuccess) return e;
    eig_args a{};
    a.B = B;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = ch->fu_f32;
    a.beta = ch->beta;
    a.U = ch->U;
    a.mu_c = ch->mu_c;
    a.evals = ch->spectrum;
    a.logz = ch->logz;
    a.ws = ch->ws;
    a.ws_stride = ch->ws_stride;
    a.vout = ch->fu_vh;
    a.tauout = ch->fu_tau;
    a.d_out = ch->fu_d;
    a.e_out = ch->fu_e;
    a.skip = skip;
    e = fk_launch_eig(ch->kind, a, ch->grid, st);
    if (e != hipSuccess) return e;
    return fk_launch_evecs(B, N, ch->fu_d, ch->fu_e, ch->spectrum, ch->fu_vh, ch->fu_tau, ch->fu_V, ch->fu_scr, st,
                           skip);
}

static hipError_t launch_step(fkgpu_chains* ch, const sweep_args& a, hipStream_t st) {
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->fu) {
        hipError_t e = fk_launch_fu_step(a, ch->fu_V, ch->fu_Q1, ch->fu_Q2, ch->fu_upd, ch->fu_T, ch->fu_V, st);
        if (e != hipSuccess) return e;
        // the step counter lives on the device, so the period runs across calls and graph
        // replays; the re-decomposition kernels are launched every step and exit at once
        // unless it is due
        if (ch->fu_refresh > 0) {
            e = fk_launch_fu_tick(ch->B, ch->fu_refresh, ch->fu_count, ch->fu_skip, st);
            if (e == hipSuccess) e = fu_refresh_launch(ch, st, ch->fu_skip);
        }
        return e;
    }
    if (ch->mixed) {
        hipError_t e = fk_launch_kpm_propose(a, ch->k_props, ch->k_fprop, ch->k_skip, st);
        if (e != hipSuccess) return e;
        e = mixed_eig_launch(ctx, ch->mx, ch->B, ch->k_fprop, ch->beta, ch->U, ch->mu_c, ch->k_skip, st);
        if (e != hipSuccess) return e;
        return fk_launch_kpm_accept(a, ch->k_props, ch->k_fprop, ch->mx->lz, st, ch->mx->ev);
    }
    if (!ch->kpm) return fk_launch_sweep(ch->kind, a, ch->grid, st);
    // the moves of src/moves_chebyshev.cpp: their global exchange includes -beta dE_ff
    sweep_args ak = a;
    ak.flip_ff = 1;
    hipError_t e = fk_launch_kpm_propose(ak, ch->k_props, ch->k_fprop, ch->k_skip, st);
    if (e != hipSuccess) return e;
    if (ch->kdelta == FKGPU_KPM_DELTA_LOCAL) {
        e = fk_launch_kpm_local(ch->B, ch->N, ch->kl_M, ch->kl_n0, ch->f, ch->k_fprop, ch->k_props, ch->k_skip, ch->U,
                                ch->mu_c, ch->beta, ch->k_ab, ctx->d_nbr, ctx->d_z, ctx->zmax, ctx->d_hopv,
                                hopq_ptr(ctx), ch->kl_s2, ch->kl_s4, ch->k_lmom, ch->logz, ch->k_lz, st);
        if (e != hipSuccess) return e;
        return fk_launch_kpm_accept(a, ch->k_props, ch->k_fprop, ch->k_lz, st);
    }
    e = fk_launch_kpm(ch->B, ch->N, ch->kR, ch->kM, ch->k_fprop, ch->U, ch->mu_c, ch->k_ab, ctx->d_nbr, ctx->d_z,
                      ctx->zmax, ctx->d_hopv, hopq_ptr(ctx), ch->kseed, 1, ch->k_skip, ch->k_coef, ch->kjack, ch->k_mom, nullptr,
                      ch->k_lz, st);
    if (e != hipSuccess) return e;
    return fk_launch_kpm_accept(a, ch->k_props, ch->k_fprop, ch->k_lz, st);
}

// K move steps of every chain from trace row a.step on: one work-queue launch for the
// dense FP64 mode, one launch_step per step otherwise.
static hipError_t launch_steps(fkgpu_chains* ch, sweep_args a, int K, hipStream_t st) {
    if (K > 1 && !ch->fu && !ch->mixed && !ch->kpm) {
        sweep_queue q{ch->q_ctr, ch->q_ring, ch->spl, ch->q_err, ch->q_ticks};
        return fk_launch_sweep_q(ch->kind, a, K, q, st);
    }
    const int step0 = a.step;
    for (int s = 0; s < K; ++s) {
        a.step = step0 + s;
        hipError_t e = launch_step(ch, a, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// nsteps move steps in launches of at most spl steps each.
static hipError_t launch_sweep_chunks(fkgpu_chains* ch, const sweep_args& a0, int nsteps, hipStream_t st) {
    sweep_args a = a0;
    for (int s = 0; s < nsteps; s += ch->spl) {
        a.step = s;
        hipError_t e = launch_steps(ch, a, std::min(ch->spl, nsteps - s), st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The work queue's wait bound expired in a launch since the chain set was created.
static int queue_check(fkgpu_chains* ch) {
    int32_t err = 0;
    FK_HIP(ch->ctx, hipMemcpy(&err, ch->q_err, sizeof(err), hipMemcpyDeviceToHost));
    if (err) return fail(ch->ctx, FKGPU_EHIP, "move-step work queue: a wait exceeded its bound; chain state undefined");
    return FKGPU_OK;
}

int fkgpu_chains_randomize_f(fkgpu_chains* ch, const int32_t* nf_start) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->nlat) return fail(ctx, FKGPU_EUNSUPPORTED, "randomize_f: not for chain sets on several lattices");
    const int B = ch->B, N = ch->N;
    if (!nf_start) return fail(ctx, FKGPU_EINVAL, "null nf_start");
    for (int c = 0; c < B; ++c)
        if (nf_start[c] < 0 || nf_start[c] > N)
            return fail(ctx, FKGPU_EINVAL, "nf_start must lie in [0, msize]");
    if (ch->fu || ch->mixed || ch->kpm || !ch->dense_ok)
        return fail(ctx, FKGPU_EUNSUPPORTED, "randomize_f: call it on a new chain set (FP64 dense mode)");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    int32_t *dnf, *df;
    int r;
    if ((r = ensure(ctx, ctx->b_f, (size_t)B * N, &df)) || (r = ensure(ctx, ctx->b_m4, (size_t)B, &dnf))) return r;
    drop_graphs(ch);
    FK_HIP(ctx, hipMemcpyAsync(dnf, nf_start, sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, fk_launch_randomize_f(B, N, ch->mt, ch->f, ch->nf, dnf, df, st));
    eig_args a{};
    a.B = B;
    a.N = N;
    a.hop = ctx->d_hop;
    a.f = df;
    a.beta = ch->beta;
    a.U = ch->U;
    a.mu_c = ch->mu_c;
    a.evals = ch->spectrum;
    a.logz = ch->logz;
    a.ws = ch->ws;
    a.ws_stride = ch->ws_stride;
    FK_HIP(ctx, fk_launch_eig(ch->kind, a, ch->grid, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_chains_set_queue_timeout(fkgpu_chains* ch, double seconds) {
    if (!ch) return FKGPU_EINVAL;
    if (!(seconds >= 0.0) || seconds > 3600.0) return fail(ch->ctx, FKGPU_EINVAL, "timeout must lie in [0, 3600] s");
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ch->ctx->device) != hipSuccess || khz <= 0)
        khz = 100000;
    ch->q_ticks = (long long)(seconds * 1000.0 * khz);
    drop_graphs(ch);  // the bound is a kernel argument
    return FKGPU_OK;
}

int fkgpu_chains_set_steps_per_launch(fkgpu_chains* ch, int steps) {
    if (!ch) return FKGPU_EINVAL;
    if (steps < 1 || (long long)steps * ch->B > 0x7fffffffLL)
        return fail(ch->ctx, FKGPU_EINVAL, "steps per launch must be >= 1 (and steps x chains < 2^31)");
    FK_HIP(ch->ctx, hipSetDevice(ch->ctx->device));
    FK_HIP(ch->ctx, hipStreamSynchronize(ch->ctx->stream));
    if (steps > ch->spl) {
        int32_t* ring = nullptr;
        if (dalloc(&ring, (size_t)steps * ch->B) != hipSuccess) return fail(ch->ctx, FKGPU_ENOMEM, "chain queue");
        (void)hipFree(ch->q_ring);
        ch->q_ring = ring;
    }
    ch->spl = steps;
    drop_graphs(ch);
    return FKGPU_OK;
}

int fkgpu_chains_set_update(fkgpu_chains* ch, int mode, int refresh_steps) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->nlat) return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one updates: not for chain sets on several lattices");
    const int B = ch->B, N = ch->N;
    if (mode != FKGPU_UPDATE_DENSE && mode != FKGPU_UPDATE_RANK1) return fail(ctx, FKGPU_EINVAL, "unknown update mode");
    if (refresh_steps < 0) return fail(ctx, FKGPU_EINVAL, "negative refresh interval");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipStreamSynchronize(st));
    drop_graphs(ch);
    if (mode == FKGPU_UPDATE_DENSE) {
        ch->fu = 0;
        return FKGPU_OK;
    }
    if (ch->kpm || ch->mixed) return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one updates need the FP64 dense eigensolver");
    if (!ch->dense_ok || N > fk_fu_nmax() || N > 1024)
        return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one updates implemented for N <= 1024");
    for (int m = 0; m < ch->args.ncp; ++m)
        if (ch->args.kinds[m] == FKGPU_MOVE_RANDOMIZE)
            return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one updates cover moves of one or two sites (not randomize)");
    if (!ch->fu_V) {
        const size_t BN = (size_t)B * N, BNN = BN * N;
        hipError_t e = hipSuccess;
        auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
        chk(dalloc(&ch->fu_V, BNN));
        chk(dalloc(&ch->fu_Q1, BNN));
        chk(dalloc(&ch->fu_Q2, BNN));
        chk(dalloc(&ch->fu_T, BNN));
        chk(dalloc(&ch->fu_vh, BNN));
        chk(dalloc(&ch->fu_tau, BN));
        chk(dalloc(&ch->fu_d, BN));
        chk(dalloc(&ch->fu_e, BN));
        chk(dalloc(&ch->fu_scr, fk_evecs_scratch_doubles(B, N)));
        chk(dalloc(&ch->fu_upd, B));
        chk(dalloc(&ch->fu_f32, BN));
        chk(dalloc(&ch->fu_skip, B));
        chk(dalloc(&ch->fu_count, 1));
        if (e != hipSuccess) {
            // release what was allocated so that a later call starts from scratch
            void** bufs[] = {(void**)&ch->fu_V, (void**)&ch->fu_Q1, (void**)&ch->fu_Q2, (void**)&ch->fu_T,
                             (void**)&ch->fu_vh, (void**)&ch->fu_tau, (void**)&ch->fu_d, (void**)&ch->fu_e,
                             (void**)&ch->fu_scr, (void**)&ch->fu_upd, (void**)&ch->fu_f32, (void**)&ch->fu_skip,
                             (void**)&ch->fu_count};
            for (void** q : bufs) {
                if (*q) (void)hipFree(*q);
                *q = nullptr;
            }
            return fail(ctx, FKGPU_ENOMEM, "rank-one update buffers");
        }
    }
    FK_HIP(ctx, hipMemsetAsync(ch->fu_count, 0, sizeof(long long), st));
    FK_HIP(ctx, fu_refresh_launch(ch, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    ch->fu = 1;
    ch->fu_refresh = refresh_steps;
    return FKGPU_OK;
}

int fkgpu_chains_set_precision(fkgpu_chains* ch, int precision) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->nlat) return fail(ctx, FKGPU_EUNSUPPORTED, "mixed precision: not for chain sets on several lattices");
    const int B = ch->B, N = ch->N;
    if (precision != FKGPU_PRECISION_FP64 && precision != FKGPU_PRECISION_MIXED)
        return fail(ctx, FKGPU_EINVAL, "unknown precision");
    if (ch->kpm) return fail(ctx, FKGPU_EUNSUPPORTED, "precision applies to the dense eigensolver, not the stochastic trace");
    if (ch->fu && precision == FKGPU_PRECISION_MIXED)
        return fail(ctx, FKGPU_EUNSUPPORTED, "rank-one update mode is FP64; switch it off first");
    if (precision == FKGPU_PRECISION_MIXED && N > fk_solver_nmax(FK_SOLVER_MIXED32))
        return fail(ctx, FKGPU_EUNSUPPORTED, "mixed-precision path implemented for N <= 1024");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipStreamSynchronize(st));
    drop_graphs(ch);
    if (!ch->k_props) {
        hipError_t e = hipSuccess;
        auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
        chk(dalloc(&ch->k_props, B));
        chk(dalloc(&ch->k_fprop, (size_t)B * N));
        chk(dalloc(&ch->k_skip, B));
        if (e != hipSuccess) return fail(ctx, FKGPU_ENOMEM, "proposal buffers");
    }
    if (precision == FKGPU_PRECISION_MIXED && !ch->mx) FK_HIP(ctx, mixed_alloc(ctx, B, N, &ch->mx));
    // logZ and spectrum of the current configurations at the requested precision
    std::vector<uint8_t> f8((size_t)B * N);
    FK_HIP(ctx, hipMemcpy(f8.data(), ch->f, f8.size(), hipMemcpyDeviceToHost));
    std::vector<int32_t> f32(f8.begin(), f8.end());
    FK_HIP(ctx, hipMemcpyAsync(ch->k_fprop, f32.data(), sizeof(int32_t) * f32.size(), hipMemcpyHostToDevice, st));
    if (precision == FKGPU_PRECISION_MIXED) {
        FK_HIP(ctx, mixed_eig_launch(ctx, ch->mx, B, ch->k_fprop, ch->beta, ch->U, ch->mu_c, nullptr, st));
        FK_HIP(ctx, hipMemcpyAsync(ch->logz, ch->mx->lz, sizeof(double) * B, hipMemcpyDeviceToDevice, st));
        FK_HIP(ctx, hipMemcpyAsync(ch->spectrum, ch->mx->ev, sizeof(double) * B * N, hipMemcpyDeviceToDevice, st));
    } else {
        eig_args a{};
        a.B = B;
        a.N = N;
        a.hop = ctx->d_hop;
        a.f = ch->k_fprop;
        a.beta = ch->beta;
        a.U = ch->U;
        a.mu_c = ch->mu_c;
        a.evals = ch->spectrum;
        a.logz = ch->logz;
        a.ws = ch->ws;
        a.ws_stride = ch->ws_stride;
        FK_HIP(ctx, fk_launch_eig(ch->kind, a, ch->grid, st));
    }
    FK_HIP(ctx, hipStreamSynchronize(st));
    ch->mixed = precision == FKGPU_PRECISION_MIXED;
    return FKGPU_OK;
}

int fkgpu_chains_fallbacks(const fkgpu_chains* ch, int64_t* n) {
    if (!ch || !n) return FKGPU_EINVAL;
    unsigned long long v = 0;
    if (ch->mx) {
        FK_HIP(ch->ctx, hipSetDevice(ch->ctx->device));
        if (hipStreamSynchronize(ch->ctx->stream) != hipSuccess ||
            hipMemcpy(&v, ch->mx->nflag, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
            return FKGPU_EHIP;
    }
    *n = (int64_t)v;
    return FKGPU_OK;
}

int fkgpu_chains_set_kpm(fkgpu_chains* ch, int R, int M, uint64_t seed, int jackson) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->nlat) return fail(ctx, FKGPU_EUNSUPPORTED, "the Chebyshev path: not for chain sets on several lattices");
    const int B = ch->B, N = ch->N;
    if (R < 1 || M < 2 || (M & 1) || M > 512) return fail(ctx, FKGPU_EINVAL, "invalid KPM arguments");
    if (N > 4096 || ctx->zmax > 8) return fail(ctx, FKGPU_EUNSUPPORTED, "KPM path needs N <= 4096 and <= 8 neighbours");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::vector<double> beta(B), U(B), mu(B);
    FK_HIP(ctx, hipMemcpy(beta.data(), ch->beta, sizeof(double) * B, hipMemcpyDeviceToHost));
    FK_HIP(ctx, hipMemcpy(U.data(), ch->U, sizeof(double) * B, hipMemcpyDeviceToHost));
    FK_HIP(ctx, hipMemcpy(mu.data(), ch->mu_c, sizeof(double) * B, hipMemcpyDeviceToHost));
    // bounds valid for every configuration of a chain: diag in {-mu_c, -mu_c + U}
    double rmax = 0.0;
    for (int i = 0; i < N; ++i) {
        double rad = 0.0;
        for (int q = 0; q < ctx->z[i]; ++q) rad += std::fabs(ctx->hopv[(size_t)i * ctx->zmax + q]);
        rmax = std::max(rmax, rad);
    }
    std::vector<double> ab(2 * (size_t)B), coef((size_t)B * M);
    const int K = std::max(2 * M, 64);
    for (int c = 0; c < B; ++c) {
        const double lo = std::min(-mu[c], -mu[c] + U[c]) - rmax, hi = std::max(-mu[c], -mu[c] + U[c]) + rmax;
        double a = 0.5 * (hi - lo) * 1.01;
        if (!(a > 0.0)) a = 1.0;
        const double b = 0.5 * (hi + lo);
        ab[2 * c] = 1.0 / a;
        ab[2 * c + 1] = b;
        for (int m = 0; m < M; ++m) {
            double sacc = 0.0;
            for (int k = 0; k < K; ++k) {
                const double th = M_PI * (k + 0.5) / K;
                const double y = beta[c] * (a * std::cos(th) + b);
                const double F = y > 0 ? std::log1p(std::exp(-y)) : -y + std::log1p(std::exp(y));
                sacc += F * std::cos(m * th);
            }
            coef[(size_t)c * M + m] = (m == 0 ? 1.0 : 2.0) * sacc / K;
        }
    }
    void* old[] = {ch->k_props, ch->k_fprop, ch->k_skip, ch->k_ab, ch->k_coef, ch->k_mom, ch->k_lz};
    for (void* p : old)
        if (p) (void)hipFree(p);
    ch->k_props = nullptr; ch->k_fprop = ch->k_skip = nullptr;
    ch->k_ab = ch->k_coef = ch->k_mom = ch->k_lz = nullptr;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(dalloc(&ch->k_props, B));
    chk(dalloc(&ch->k_fprop, (size_t)B * N));
    chk(dalloc(&ch->k_skip, B));
    chk(dalloc(&ch->k_ab, 2 * (size_t)B));
    chk(dalloc(&ch->k_coef, (size_t)B * M));
    chk(dalloc(&ch->k_mom, (size_t)B * R * M));
    chk(dalloc(&ch->k_lz, B));
    if (e != hipSuccess) return fail(ctx, FKGPU_ENOMEM, "KPM buffers");
    FK_HIP(ctx, hipMemcpyAsync(ch->k_ab, ab.data(), sizeof(double) * 2 * B, hipMemcpyHostToDevice, st));
    FK_HIP(ctx, hipMemcpyAsync(ch->k_coef, coef.data(), sizeof(double) * B * M, hipMemcpyHostToDevice, st));
    // logZ of the current configurations with the chain's own probe vectors
    std::vector<uint8_t> f8((size_t)B * N);
    FK_HIP(ctx, hipMemcpy(f8.data(), ch->f, f8.size(), hipMemcpyDeviceToHost));
    std::vector<int32_t> f32(f8.begin(), f8.end());
    FK_HIP(ctx, hipMemcpyAsync(ch->k_fprop, f32.data(), sizeof(int32_t) * f32.size(), hipMemcpyHostToDevice, st));
    FK_HIP(ctx, fk_launch_kpm(B, N, R, M, ch->k_fprop, ch->U, ch->mu_c, ch->k_ab, ctx->d_nbr, ctx->d_z, ctx->zmax,
                              ctx->d_hopv, hopq_ptr(ctx), seed, 1, nullptr, ch->k_coef, jackson, ch->k_mom, nullptr, ch->logz, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    ch->kpm = 1;
    ch->mixed = 0;
    ch->fu = 0;
    ch->kR = R;
    ch->kM = M;
    ch->kseed = seed;
    ch->kjack = jackson;
    drop_graphs(ch);
    // proposals' dlogZ from local moments by default
    return fkgpu_chains_set_kpm_delta(ch, FKGPU_KPM_DELTA_LOCAL, 0, 0);
}

// Sums of (2n + 1)^-k over n >= n0 (k = 2, 4): direct sum of 2^20 terms in long double,
// then the Euler-Maclaurin remainder (integral + half the first omitted term).
static double odd_power_tail(int n0, int k) {
    const long nt = 1L << 20;
    long double sum = 0.0L;
    for (long n = n0 + nt - 1; n >= n0; --n) sum += 1.0L / powl(2.0L * n + 1.0L, k);
    const long double x = 2.0L * (n0 + nt) + 1.0L;
    sum += 1.0L / ((k - 1) * 2.0L * powl(x, k - 1)) + 0.5L / powl(x, k);
    return (double)sum;
}

int fkgpu_chains_set_kpm_delta(fkgpu_chains* ch, int mode, int moments, int frequencies) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (!ch->kpm) return fail(ctx, FKGPU_EINVAL, "the chain set is not on the stochastic-trace path (fkgpu_chains_set_kpm)");
    if (mode == FKGPU_KPM_DELTA_STOCHASTIC) {
        ch->kdelta = mode;
        drop_graphs(ch);
        return FKGPU_OK;
    }
    if (mode != FKGPU_KPM_DELTA_LOCAL) return fail(ctx, FKGPU_EINVAL, "unknown dlogZ estimator");
    if (ch->N > FK_KPM_LOCAL_NMAX) return fail(ctx, FKGPU_EUNSUPPORTED, "local dlogZ estimator needs N <= 4096");
    if (moments < 0 || moments == 1 || moments > FK_KPM_LOCAL_MMAX || frequencies < 0 || frequencies == 1)
        return fail(ctx, FKGPU_EINVAL, "moments must be 2..2048 and frequencies >= 2 (0: automatic)");
    const int B = ch->B;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> beta(B), ab(2 * (size_t)B);
    FK_HIP(ctx, hipMemcpy(beta.data(), ch->beta, sizeof(double) * B, hipMemcpyDeviceToHost));
    FK_HIP(ctx, hipMemcpy(ab.data(), ch->k_ab, sizeof(double) * 2 * B, hipMemcpyDeviceToHost));
    if (moments == 0) {
        // the series of 1/(z' - x) at the lowest frequency decays like |r|^m with
        // |r| <= 1 - pi / (beta a) (z' on the imaginary axis, the slowest case): enough
        // terms for |r|^M < 1e-12, rounded up to a multiple of 32
        double worst = 0.0;
        for (int c = 0; c < B; ++c) worst = std::max(worst, beta[c] / ab[2 * c]);
        const double need = std::log(1e12) * worst / M_PI;
        moments = std::max(64, std::min(FK_KPM_LOCAL_MMAX, 32 * (int)std::ceil(need / 32.0)));
    }
    if (frequencies == 0) frequencies = 512;
    if (ch->k_lmom) (void)hipFree(ch->k_lmom);
    ch->k_lmom = nullptr;
    if (dalloc(&ch->k_lmom, (size_t)B * 3 * moments) != hipSuccess) return fail(ctx, FKGPU_ENOMEM, "local moment buffers");
    ch->kl_M = moments;
    ch->kl_n0 = frequencies;
    ch->kl_s2 = odd_power_tail(frequencies, 2);
    ch->kl_s4 = odd_power_tail(frequencies, 4);
    ch->kdelta = FKGPU_KPM_DELTA_LOCAL;
    drop_graphs(ch);
    return FKGPU_OK;
}

int fkgpu_chains_kpm_delta(const fkgpu_chains* ch, int* mode, int* moments, int* frequencies) {
    if (!ch) return FKGPU_EINVAL;
    if (mode) *mode = ch->kpm ? ch->kdelta : -1;
    if (moments) *moments = ch->kl_M;
    if (frequencies) *frequencies = ch->kl_n0;
    return FKGPU_OK;
}

// The local dlogZ estimator covers moves that change one or two sites.
static int kpm_moves_ok(fkgpu_chains* ch) {
    if (!ch->kpm || ch->kdelta != FKGPU_KPM_DELTA_LOCAL) return FKGPU_OK;
    for (int m = 0; m < ch->args.ncp; ++m)
        if (ch->args.kinds[m] == FKGPU_MOVE_RANDOMIZE)
            return fail(ch->ctx, FKGPU_EUNSUPPORTED,
                        "the local dlogZ estimator covers moves of one or two sites; randomize needs "
                        "FKGPU_KPM_DELTA_STOCHASTIC");
    return FKGPU_OK;
}

int fkgpu_sweep(fkgpu_chains* ch, int nsteps, const fkgpu_trace* trace) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (nsteps < 0) return fail(ctx, FKGPU_EINVAL, "negative step count");
    if (int r = kpm_moves_ok(ch)) return r;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    sweep_args a = ch->args;
    const size_t need = (size_t)nsteps * ch->B;
    if (trace) {
        if (ch->trace_cap < need) {
            void* ps[] = {ch->t_kind, ch->t_a, ch->t_b, ch->t_u, ch->t_logw, ch->t_logz, ch->t_acc};
            for (void* p : ps)
                if (p) (void)hipFree(p);
            ch->t_kind = ch->t_a = ch->t_b = nullptr;
            ch->t_u = ch->t_logw = ch->t_logz = nullptr;
            ch->t_acc = nullptr;
            ch->trace_cap = 0;
            hipError_t e = hipSuccess;
            auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
            chk(dalloc(&ch->t_kind, need));
            chk(dalloc(&ch->t_a, need));
            chk(dalloc(&ch->t_b, need));
            chk(dalloc(&ch->t_u, need));
            chk(dalloc(&ch->t_logw, need));
            chk(dalloc(&ch->t_logz, need));
            chk(dalloc(&ch->t_acc, need));
            if (e != hipSuccess) return fail(ctx, FKGPU_ENOMEM, "trace allocation failed");
            ch->trace_cap = need;
        }
        a.trace = 1;
        a.tr_kind = ch->t_kind;
        a.tr_a = ch->t_a;
        a.tr_b = ch->t_b;
        a.tr_u = ch->t_u;
        a.tr_logw = ch->t_logw;
        a.tr_logz = ch->t_logz;
        a.tr_acc = ch->t_acc;
    }
    if (!ch->kpm && !ch->dense_ok)
        return fail(ctx, FKGPU_EUNSUPPORTED, "no dense solver for this size: enable the stochastic trace first");
    FK_HIP(ctx, hipEventRecord(ch->ev0, st));
    FK_HIP(ctx, launch_sweep_chunks(ch, a, nsteps, st));
    FK_HIP(ctx, hipEventRecord(ch->ev1, st));
    FK_HIP(ctx, hipEventSynchronize(ch->ev1));
    float ms = 0.f;
    FK_HIP(ctx, hipEventElapsedTime(&ms, ch->ev0, ch->ev1));
    ch->last_ms = ms;
    if (int r = queue_check(ch)) return r;
    if (trace && nsteps > 0) {
        auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
            return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
        };
        FK_HIP(ctx, cp(trace->kind, ch->t_kind, sizeof(int32_t) * need));
        FK_HIP(ctx, cp(trace->site_a, ch->t_a, sizeof(int32_t) * need));
        FK_HIP(ctx, cp(trace->site_b, ch->t_b, sizeof(int32_t) * need));
        FK_HIP(ctx, cp(trace->u, ch->t_u, sizeof(double) * need));
        FK_HIP(ctx, cp(trace->logw, ch->t_logw, sizeof(double) * need));
        FK_HIP(ctx, cp(trace->logz_new, ch->t_logz, sizeof(double) * need));
        FK_HIP(ctx, cp(trace->accepted, ch->t_acc, sizeof(int8_t) * need));
    }
    return FKGPU_OK;
}

int fkgpu_sweep_graph(fkgpu_chains* ch, int nsteps) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (nsteps <= 0) return fail(ctx, FKGPU_EINVAL, "step count must be positive");
    if (int r = kpm_moves_ok(ch)) return r;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (!ch->gexec || ch->graph_steps != nsteps) {
        if (ch->gexec) { (void)hipGraphExecDestroy(ch->gexec); ch->gexec = nullptr; }
        if (ch->graph) { (void)hipGraphDestroy(ch->graph); ch->graph = nullptr; }
        sweep_args a = ch->args;
        a.trace = 0;
        if (!ch->kpm && !ch->dense_ok)
            return fail(ctx, FKGPU_EUNSUPPORTED, "no dense solver for this size: enable the stochastic trace first");
        FK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_sweep_chunks(ch, a, nsteps, st);
        if (e != hipSuccess) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            return hipfail(ctx, e, "graph capture");
        }
        FK_HIP(ctx, hipStreamEndCapture(st, &ch->graph));
        FK_HIP(ctx, hipGraphInstantiate(&ch->gexec, ch->graph, nullptr, nullptr, 0));
        ch->graph_steps = nsteps;
    }
    FK_HIP(ctx, hipEventRecord(ch->ev0, st));
    FK_HIP(ctx, hipGraphLaunch(ch->gexec, st));
    FK_HIP(ctx, hipEventRecord(ch->ev1, st));
    FK_HIP(ctx, hipEventSynchronize(ch->ev1));
    float ms = 0.f;
    FK_HIP(ctx, hipEventElapsedTime(&ms, ch->ev0, ch->ev1));
    ch->last_ms = ms;
    return queue_check(ch);
}

int fkgpu_chains_last_ms(const fkgpu_chains* ch, double* ms) {
    if (!ch || !ms) return FKGPU_EINVAL;
    *ms = ch->last_ms;
    return FKGPU_OK;
}

int fkgpu_chains_state(fkgpu_chains* ch, int32_t* f, int32_t* nf, double* logZ, double* spectrum,
                       int64_t* naccept, int64_t* nsolve) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    FK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t BN = (size_t)ch->B * ch->N;
    if (f) {
        std::vector<uint8_t> f8(BN);
        FK_HIP(ctx, hipMemcpy(f8.data(), ch->f, BN, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < BN; ++i) f[i] = f8[i];
    }
    if (nf) FK_HIP(ctx, hipMemcpy(nf, ch->nf, sizeof(int32_t) * ch->B, hipMemcpyDeviceToHost));
    if (logZ) FK_HIP(ctx, hipMemcpy(logZ, ch->logz, sizeof(double) * ch->B, hipMemcpyDeviceToHost));
    if (spectrum) FK_HIP(ctx, hipMemcpy(spectrum, ch->spectrum, sizeof(double) * BN, hipMemcpyDeviceToHost));
    if (naccept) FK_HIP(ctx, hipMemcpy(naccept, ch->naccept, sizeof(int64_t) * ch->B, hipMemcpyDeviceToHost));
    if (nsolve) FK_HIP(ctx, hipMemcpy(nsolve, ch->nsolve, sizeof(int64_t) * ch->B, hipMemcpyDeviceToHost));
    return FKGPU_OK;
}

// ---- on-device measurements ----------------------------------------------------

// The decomposition buffers a measurement needs beyond the chain set's own state, allocated
// before any graph capture: the spectrum of stochastic-trace chains and, for the IPR, the
// eigenvectors of chains outside the rank-one mode.
static int meas_prepare(fkgpu_chains* ch) {
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if (!m) return fail(ctx, FKGPU_EINVAL, "measurements are not enabled on this chain set");
    if (m->flags & FKGPU_MEASURE_NO_SPECTRUM) return FKGPU_OK;  // occupations only: no eigensolve
    if (!ch->kpm && !ch->dense_ok)
        return fail(ctx, FKGPU_EUNSUPPORTED, "no dense solver for this size: enable the stochastic trace first");
    const bool ipr = m->flags & (FKGPU_MEASURE_IPR | FKGPU_MEASURE_STIFFNESS | FKGPU_MEASURE_EIGENFUNCTIONS);
    if (!ch->kpm && !(ipr && !ch->fu)) return FKGPU_OK;
    if (!ch->dense_ok) return fail(ctx, FKGPU_EUNSUPPORTED, "measurements need the dense eigensolver (N <= 2048)");
    const int B = ch->B, N = ch->N;
    const size_t BN = (size_t)B * N, BNN = BN * N;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    if (!m->f32) chk(dalloc(&m->f32, BN));
    if (!m->ev) chk(dalloc(&m->ev, BN));
    if (!m->lz) chk(dalloc(&m->lz, B));
    if (ipr && !m->V) {
        chk(dalloc(&m->vh, BNN));
        chk(dalloc(&m->tau, BN));
        chk(dalloc(&m->d, BN));
        chk(dalloc(&m->eo, BN));
        chk(dalloc(&m->V, BNN));
        chk(dalloc(&m->scr, fk_evecs_scratch_doubles(B, N)));
    }
    if (e != hipSuccess) return fail(ctx, FKGPU_ENOMEM, "measurement decomposition buffers");
    return FKGPU_OK;
}

// One measurement of every chain, stream-ordered (capturable: no allocation, no memset).
static hipError_t meas_launch(fkgpu_chains* ch, hipStream_t st) {
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    const int B = ch->B, N = ch->N;
    const bool vec = m->flags & (FKGPU_MEASURE_IPR | FKGPU_MEASURE_STIFFNESS | FKGPU_MEASURE_EIGENFUNCTIONS);
    const bool ipr = vec;  // eigenvectors needed
    const double* spec = ch->spectrum;
    const double* V = (ipr && ch->fu) ? ch->fu_V : nullptr;
    const double* specV = ch->spectrum;  // the eigenvalues that belong to V
    const bool nospec = m->flags & FKGPU_MEASURE_NO_SPECTRUM;
    hipError_t e;
    if (!nospec && (ch->kpm || (ipr && !ch->fu))) {
        if ((e = fk_launch_f8to32(B * N, ch->f, m->f32, st)) != hipSuccess) return e;
        eig_args a{};
        a.B = B;
        a.N = N;
        a.hop = ctx->d_hop;
        a.f = m->f32;
        a.beta = ch->beta;
        a.U = ch->U;
        a.mu_c = ch->mu_c;
        a.evals = m->ev;
        a.logz = m->lz;
        a.ws = ch->ws;
        a.ws_stride = ch->ws_stride;
        if (ipr) {
            a.vout = m->vh;
            a.tauout = m->tau;
            a.d_out = m->d;
            a.e_out = m->eo;
        }
        if ((e = fk_launch_eig(ch->kind, a, ch->grid, st)) != hipSuccess) return e;
        if (ipr) {
            e = fk_launch_evecs(B, N, m->d, m->eo, m->ev, m->vh, m->tau, m->V, m->scr, st);
            if (e != hipSuccess) return e;
            V = m->V;
            specV = m->ev;
        }
        if (ch->kpm) spec = m->ev;
    }
    if (V && (m->flags & FKGPU_MEASURE_IPR) && (e = fk_launch_ipr(B, N, V, m->ipr, st)) != hipSuccess) return e;
    if (V && (m->flags & FKGPU_MEASURE_STIFFNESS)) {
        // measure_stiffness::accumulate on the chains' eigenpairs, direction 0, T_x hopping 1
        if ((e = fk_launch_fermi(B, N, specV, ch->beta, m->fermi, st)) != hipSuccess) return e;
        int stride0 = 1;
        for (int d = 1; d < ctx->lat_ndim; ++d) stride0 *= ctx->lat_L;
        e = fk_launch_stiffness(B, N, stride0, ctx->lat_L, 1.0, V, specV, m->fermi, m->nw, m->wgrid, m->offset, m->JV,
                                m->mJ, m->tdiag, m->stiff, m->cond, st);
        if (e != hipSuccess) return e;
    }
    chain_meas_args a{};
    a.B = B;
    a.N = N;
    a.cap = m->cap;
    a.nW = ch->args.nW;
    for (int i = 0; i < 8; ++i) a.W[i] = ch->args.W[i];
    a.f = ch->f;
    a.nf = ch->nf;
    a.spectrum = nospec ? nullptr : spec;
    a.beta = ch->beta;
    a.mu_f = ch->mu_f;
    a.ipr = (V && (m->flags & FKGPU_MEASURE_IPR)) ? m->ipr : nullptr;
    const bool stf = V && (m->flags & FKGPU_MEASURE_STIFFNESS);
    a.nw = m->nw;
    a.stiff = stf ? m->stiff : nullptr;
    a.cond = m->cond;
    a.stiff_hist = m->stiff_hist;
    a.stiff_sum = m->stiff_sum;
    a.cond_hist = m->cond_hist;
    a.cond_sum = m->cond_sum;
    a.V = (V && (m->flags & FKGPU_MEASURE_EIGENFUNCTIONS)) ? V : nullptr;
    a.evec_hist = m->evec_hist;
    a.count = m->count;
    a.e = m->e;
    a.d2e = m->d2e;
    a.ce = m->ce;
    a.esum = m->esum;
    a.d2sum = m->d2sum;
    a.spec_mean = m->spec_mean;
    a.focc_sum = m->focc_sum;
    a.spec_hist = m->spec_hist;
    a.focc_hist = m->focc_hist;
    a.ipr_hist = m->ipr_hist;
    return fk_launch_chain_measure(a, m->count, st);
}

static int meas_count(fkgpu_chains* ch, long long* count) {
    fkgpu_ctx* ctx = ch->ctx;
    if (!ch->ms) return fail(ctx, FKGPU_EINVAL, "measurements are not enabled on this chain set");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    FK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FK_HIP(ctx, hipMemcpy(count, ch->ms->count, sizeof(long long), hipMemcpyDeviceToHost));
    return FKGPU_OK;
}

int fkgpu_chains_measure_enable(fkgpu_chains* ch, int capacity, int flags) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (ch->nlat) return fail(ctx, FKGPU_EUNSUPPORTED, "on-device measurements: not for chain sets on several lattices");
    const int vecs = FKGPU_MEASURE_IPR | FKGPU_MEASURE_STIFFNESS | FKGPU_MEASURE_EIGENFUNCTIONS;
    const int all = FKGPU_MEASURE_SPECTRUM_HISTORY | FKGPU_MEASURE_FOCC_HISTORY | vecs | FKGPU_MEASURE_NO_SPECTRUM;
    if (capacity < 0 || (flags & ~all)) return fail(ctx, FKGPU_EINVAL, "invalid capacity or measurement flags");
    if ((flags & FKGPU_MEASURE_NO_SPECTRUM) && (flags & (vecs | FKGPU_MEASURE_SPECTRUM_HISTORY)))
        return fail(ctx, FKGPU_EINVAL, "FKGPU_MEASURE_NO_SPECTRUM records occupations only");
    if ((flags & vecs) && (!ch->dense_ok || ch->N > 1024))
        return fail(ctx, FKGPU_EUNSUPPORTED, "IPR, stiffness and eigenfunctions need eigenvectors, implemented for N <= 1024");
    if ((flags & FKGPU_MEASURE_STIFFNESS) && (ctx->lat_kind != FKGPU_LATTICE_HYPERCUBIC || ctx->lat_ndim < 2))
        return fail(ctx, FKGPU_EUNSUPPORTED, "stiffness needs a hypercubic lattice of dimension >= 2");
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, hipStreamSynchronize(st));
    drop_graphs(ch);
    meas_free(ch->ms);
    ch->ms = nullptr;
    const int B = ch->B, N = ch->N;
    const size_t BN = (size_t)B * N, Bc = (size_t)B * capacity, BcN = Bc * N;
    chain_meas_bufs* m = new chain_meas_bufs();
    m->cap = capacity;
    m->flags = flags;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (x != hipSuccess && e == hipSuccess) e = x; };
    chk(dalloc(&m->count, 1));
    chk(dalloc(&m->e, Bc));
    chk(dalloc(&m->d2e, Bc));
    chk(dalloc(&m->ce, Bc));
    chk(dalloc(&m->esum, B));
    chk(dalloc(&m->d2sum, B));
    chk(dalloc(&m->spec_mean, BN));
    chk(dalloc(&m->focc_sum, BN));
    if (flags & FKGPU_MEASURE_SPECTRUM_HISTORY) chk(dalloc(&m->spec_hist, BcN));
    if (flags & FKGPU_MEASURE_FOCC_HISTORY) chk(dalloc(&m->focc_hist, BcN));
    if (flags & FKGPU_MEASURE_IPR) {
        chk(dalloc(&m->ipr, BN));
        chk(dalloc(&m->ipr_hist, BcN));
    }
    if (flags & FKGPU_MEASURE_EIGENFUNCTIONS) chk(dalloc(&m->evec_hist, BcN * N));
    const int nw = (int)ch->m_wgrid.size();
    m->nw = nw;
    m->offset = ch->m_offset;
    if (flags & FKGPU_MEASURE_STIFFNESS) {
        const size_t BNN = BN * N;
        chk(dalloc(&m->wgrid, nw));
        chk(dalloc(&m->fermi, BN));
        chk(dalloc(&m->JV, BNN));
        chk(dalloc(&m->mJ, BNN));
        chk(dalloc(&m->tdiag, BN));
        chk(dalloc(&m->stiff, B));
        chk(dalloc(&m->cond, (size_t)B * nw));
        chk(dalloc(&m->stiff_hist, Bc));
        chk(dalloc(&m->cond_hist, Bc * nw));
        chk(dalloc(&m->stiff_sum, B));
        chk(dalloc(&m->cond_sum, (size_t)B * nw));
    }
    if (e != hipSuccess) {
        meas_free(m);
        return fail(ctx, FKGPU_ENOMEM, std::string("measurement records: ") + hipGetErrorString(e));
    }
    chk(hipMemsetAsync(m->count, 0, sizeof(long long), st));
    chk(hipMemsetAsync(m->esum, 0, sizeof(double) * B, st));
    chk(hipMemsetAsync(m->d2sum, 0, sizeof(double) * B, st));
    chk(hipMemsetAsync(m->spec_mean, 0, sizeof(double) * BN, st));
    chk(hipMemsetAsync(m->focc_sum, 0, sizeof(long long) * BN, st));
    if (flags & FKGPU_MEASURE_STIFFNESS) {
        chk(hipMemcpyAsync(m->wgrid, ch->m_wgrid.data(), sizeof(double) * nw, hipMemcpyHostToDevice, st));
        chk(hipMemsetAsync(m->stiff_sum, 0, sizeof(double) * B, st));
        chk(hipMemsetAsync(m->cond_sum, 0, sizeof(double) * B * nw, st));
    }
    chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        meas_free(m);
        return hipfail(ctx, e, "measurement records");
    }
    ch->ms = m;
    return FKGPU_OK;
}

int fkgpu_chains_measure_conductivity(fkgpu_chains* ch, int nw, const double* wgrid, double offset) {
    if (!ch) return FKGPU_EINVAL;
    if (nw < 1 || !wgrid || !finite_all(wgrid, nw) || !(offset > 0.0) || !std::isfinite(offset))
        return fail(ch->ctx, FKGPU_EINVAL, "conductivity grid: nw >= 1 finite frequencies and a positive width");
    ch->m_wgrid.assign(wgrid, wgrid + nw);
    ch->m_offset = offset;
    return FKGPU_OK;
}

int fkgpu_chains_stiffness(fkgpu_chains* ch, int64_t* count, double* stiffness, double* cond, double* stiffness_mean,
                           double* cond_mean) {
    if (!ch) return FKGPU_EINVAL;
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if (!(m->flags & FKGPU_MEASURE_STIFFNESS)) return fail(ctx, FKGPU_EINVAL, "stiffness was not enabled");
    if (count) *count = c;
    const int B = ch->B, nw = m->nw;
    const size_t n = (size_t)std::min<long long>(c, m->cap);
    if (n > 0) {
        if (stiffness)
            FK_HIP(ctx, hipMemcpy2D(stiffness, sizeof(double) * n, m->stiff_hist, sizeof(double) * m->cap,
                                    sizeof(double) * n, B, hipMemcpyDeviceToHost));
        if (cond)
            FK_HIP(ctx, hipMemcpy2D(cond, sizeof(double) * n * nw, m->cond_hist, sizeof(double) * m->cap * nw,
                                    sizeof(double) * n * nw, B, hipMemcpyDeviceToHost));
    }
    const double z = c > 0 ? (double)c : 1.0;
    if (stiffness_mean) {
        FK_HIP(ctx, hipMemcpy(stiffness_mean, m->stiff_sum, sizeof(double) * B, hipMemcpyDeviceToHost));
        for (int b = 0; b < B; ++b) stiffness_mean[b] /= z;
    }
    if (cond_mean) {
        FK_HIP(ctx, hipMemcpy(cond_mean, m->cond_sum, sizeof(double) * B * nw, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < (size_t)B * nw; ++i) cond_mean[i] /= z;
    }
    return FKGPU_OK;
}

int fkgpu_chains_eigenfunctions(fkgpu_chains* ch, int64_t* count, double* evecs) {
    if (!ch) return FKGPU_EINVAL;
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if (!(m->flags & FKGPU_MEASURE_EIGENFUNCTIONS)) return fail(ctx, FKGPU_EINVAL, "eigenfunctions were not enabled");
    if (count) *count = c;
    const size_t n = (size_t)std::min<long long>(c, m->cap), NN = (size_t)ch->N * ch->N;
    if (n > 0 && evecs)
        FK_HIP(ctx, hipMemcpy2D(evecs, sizeof(double) * n * NN, m->evec_hist, sizeof(double) * m->cap * NN,
                                sizeof(double) * n * NN, ch->B, hipMemcpyDeviceToHost));
    return FKGPU_OK;
}

// The records averaged by the Green's-function readers: every recorded measurement of
// every chain (eigenvectors and spectrum history both enabled).
static int green_records(fkgpu_chains* ch, int* n) {
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    chain_meas_bufs* m = ch->ms;
    const int need = FKGPU_MEASURE_EIGENFUNCTIONS | FKGPU_MEASURE_SPECTRUM_HISTORY;
    if ((m->flags & need) != need)
        return fail(ch->ctx, FKGPU_EINVAL, "the Green's function needs the eigenfunction and spectrum histories");
    *n = (int)std::min<long long>(c, m->cap);
    if (*n < 1) return fail(ch->ctx, FKGPU_EINVAL, "no recorded measurements");
    return FKGPU_OK;
}

int fkgpu_chains_green(fkgpu_chains* ch, double w, double offset, double* g_re, double* g_im) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (!(offset > 0.0) || !std::isfinite(w) || !g_re || !g_im)
        return fail(ctx, FKGPU_EINVAL, "invalid Green's function arguments");
    int n = 0, r = green_records(ch, &n);
    if (r) return r;
    chain_meas_bufs* m = ch->ms;
    const int N = ch->N, tb = (N + 63) / 64;
    const int splits = fk_green_splits(n * ch->B, tb * tb);
    const size_t NN = (size_t)N * N;
    double* buf = nullptr;
    if (dalloc(&buf, (2 * (size_t)splits + 2) * NN) != hipSuccess)
        return fail(ctx, FKGPU_ENOMEM, "Green's function buffers");
    hipStream_t st = ctx->stream;
    hipError_t e = hipEventRecord(ch->ev0, st);
    if (e == hipSuccess)
        e = fk_launch_green(n, m->cap, ch->B, N, m->evec_hist, m->spec_hist, w, offset, splits, buf, buf + splits * NN,
                            buf + 2 * splits * NN, buf + (2 * splits + 1) * NN, st);
    if (e == hipSuccess) e = hipEventRecord(ch->ev1, st);
    if (e == hipSuccess) e = hipMemcpyAsync(g_re, buf + 2 * splits * NN, sizeof(double) * NN, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g_im, buf + (2 * splits + 1) * NN, sizeof(double) * NN, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ch->ev0, ch->ev1);
    (void)hipFree(buf);
    FK_HIP(ctx, e);
    ch->last_ms = ms;
    return FKGPU_OK;
}

int fkgpu_chains_local_dos(fkgpu_chains* ch, int nw, const double* wgrid, double offset, double* ldos) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (nw < 1 || !wgrid || !ldos || !(offset > 0.0)) return fail(ctx, FKGPU_EINVAL, "invalid local DOS arguments");
    for (int i = 0; i < nw; ++i)
        if (!std::isfinite(wgrid[i])) return fail(ctx, FKGPU_EINVAL, "invalid local DOS arguments");
    int n = 0, r = green_records(ch, &n);
    if (r) return r;
    chain_meas_bufs* m = ch->ms;
    const int N = ch->N;
    const int splits = fk_green_splits(n * ch->B, ((N + 63) / 64) * ((nw + 63) / 64));
    const size_t len = (size_t)nw * N;
    double* buf = nullptr;
    if (dalloc(&buf, ((size_t)splits + 1) * len + nw) != hipSuccess)
        return fail(ctx, FKGPU_ENOMEM, "local DOS buffers");
    double* wg = buf + (splits + 1) * len;
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(wg, wgrid, sizeof(double) * nw, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ch->ev0, st);
    if (e == hipSuccess)
        e = fk_launch_ldos(n, m->cap, ch->B, N, m->evec_hist, m->spec_hist, nw, wg, offset, splits, buf,
                           buf + splits * len, st);
    if (e == hipSuccess) e = hipEventRecord(ch->ev1, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ldos, buf + splits * len, sizeof(double) * len, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ch->ev0, ch->ev1);
    (void)hipFree(buf);
    FK_HIP(ctx, e);
    ch->last_ms = ms;
    return FKGPU_OK;
}

int fkgpu_chains_measure(fkgpu_chains* ch) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    int r = meas_prepare(ch);
    if (r) return r;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    FK_HIP(ctx, meas_launch(ch, st));
    FK_HIP(ctx, hipStreamSynchronize(st));
    return FKGPU_OK;
}

int fkgpu_sweep_measure(fkgpu_chains* ch, int nsweeps, int sweep_len) {
    if (!ch) return FKGPU_EINVAL;
    fkgpu_ctx* ctx = ch->ctx;
    if (nsweeps < 0 || sweep_len <= 0) return fail(ctx, FKGPU_EINVAL, "invalid sweep count or length");
    if (!ch->kpm && !ch->dense_ok)
        return fail(ctx, FKGPU_EUNSUPPORTED, "no dense solver for this size: enable the stochastic trace first");
    int r = kpm_moves_ok(ch);
    if (r) return r;
    r = meas_prepare(ch);
    if (r) return r;
    FK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (!ch->mexec || ch->mgraph_len != sweep_len) {
        if (ch->mexec) { (void)hipGraphExecDestroy(ch->mexec); ch->mexec = nullptr; }
        if (ch->mgraph) { (void)hipGraphDestroy(ch->mgraph); ch->mgraph = nullptr; }
        sweep_args a = ch->args;
        a.trace = 0;
        FK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_sweep_chunks(ch, a, sweep_len, st);
        if (e == hipSuccess) e = meas_launch(ch, st);
        if (e != hipSuccess) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            return hipfail(ctx, e, "graph capture");
        }
        FK_HIP(ctx, hipStreamEndCapture(st, &ch->mgraph));
        FK_HIP(ctx, hipGraphInstantiate(&ch->mexec, ch->mgraph, nullptr, nullptr, 0));
        ch->mgraph_len = sweep_len;
    }
    FK_HIP(ctx, hipEventRecord(ch->ev0, st));
    for (int i = 0; i < nsweeps; ++i) FK_HIP(ctx, hipGraphLaunch(ch->mexec, st));
    FK_HIP(ctx, hipEventRecord(ch->ev1, st));
    FK_HIP(ctx, hipEventSynchronize(ch->ev1));
    float ms = 0.f;
    FK_HIP(ctx, hipEventElapsedTime(&ms, ch->ev0, ch->ev1));
    ch->last_ms = ms;
    return queue_check(ch);
}

int fkgpu_chains_energies(fkgpu_chains* ch, int64_t* count, double* energies, double* d2energies,
                          double* c_energies) {
    if (!ch) return FKGPU_EINVAL;
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if (count) *count = c;
    const size_t n = (size_t)std::min<long long>(c, m->cap);
    if (n == 0) return FKGPU_OK;
    const double* src[3] = {m->e, m->d2e, m->ce};
    double* dst[3] = {energies, d2energies, c_energies};
    for (int i = 0; i < 3; ++i)
        if (dst[i])
            FK_HIP(ctx, hipMemcpy2D(dst[i], sizeof(double) * n, src[i], sizeof(double) * m->cap, sizeof(double) * n,
                                    ch->B, hipMemcpyDeviceToHost));
    return FKGPU_OK;
}

int fkgpu_chains_means(fkgpu_chains* ch, int64_t* count, double* energy, double* d2energy, double* spectrum,
                       double* focc) {
    if (!ch) return FKGPU_EINVAL;
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if (count) *count = c;
    const int B = ch->B;
    const size_t BN = (size_t)B * ch->N;
    const double z = c > 0 ? (double)c : 1.0;
    std::vector<double> tmp(B);
    if (energy) {
        FK_HIP(ctx, hipMemcpy(tmp.data(), m->esum, sizeof(double) * B, hipMemcpyDeviceToHost));
        for (int b = 0; b < B; ++b) energy[b] = tmp[b] / z;
    }
    if (d2energy) {
        FK_HIP(ctx, hipMemcpy(tmp.data(), m->d2sum, sizeof(double) * B, hipMemcpyDeviceToHost));
        for (int b = 0; b < B; ++b) d2energy[b] = tmp[b] / z;
    }
    if (spectrum) FK_HIP(ctx, hipMemcpy(spectrum, m->spec_mean, sizeof(double) * BN, hipMemcpyDeviceToHost));
    if (focc) {
        std::vector<long long> cnt(BN);
        FK_HIP(ctx, hipMemcpy(cnt.data(), m->focc_sum, sizeof(long long) * BN, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < BN; ++i) focc[i] = (double)cnt[i] / z;
    }
    return FKGPU_OK;
}

int fkgpu_chains_histories(fkgpu_chains* ch, int64_t* count, double* spectrum, int32_t* focc, double* ipr) {
    if (!ch) return FKGPU_EINVAL;
    long long c = 0;
    int r = meas_count(ch, &c);
    if (r) return r;
    fkgpu_ctx* ctx = ch->ctx;
    chain_meas_bufs* m = ch->ms;
    if ((spectrum && !m->spec_hist) || (focc && !m->focc_hist) || (ipr && !m->ipr_hist))
        return fail(ctx, FKGPU_EINVAL, "history requested that was not enabled");
    if (count) *count = c;
    const size_t n = (size_t)std::min<long long>(c, m->cap), N = ch->N;
    if (n == 0) return FKGPU_OK;
    const size_t row = n * N, srow = (size_t)m->cap * N;
    if (spectrum)
        FK_HIP(ctx, hipMemcpy2D(spectrum, sizeof(double) * row, m->spec_hist, sizeof(double) * srow,
                                sizeof(double) * row, ch->B, hipMemcpyDeviceToHost));
    if (ipr)
        FK_HIP(ctx, hipMemcpy2D(ipr, sizeof(double) * row, m->ipr_hist, sizeof(double) * srow, sizeof(double) * row,
                                ch->B, hipMemcpyDeviceToHost));
    if (focc) {
        std::vector<uint8_t> f8(row * ch->B);
        FK_HIP(ctx, hipMemcpy2D(f8.data(), row, m->focc_hist, srow, row, ch->B, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f8.size(); ++i) focc[i] = f8[i];
    }
    return FKGPU_OK;
}

int fkgpu_rng_draws(uint32_t seed, int mode, int hi, int nprobs, const double* probs, int count, double* out) {
    if (!out || count < 0 || mode < 0 || mode > 3 || (mode == 3 && (nprobs < 1 || nprobs > 64 || !probs)))
        return FKGPU_EINVAL;
    std::vector<uint32_t> st(fkrng::MT_WORDS);
    fkrng::mt_seed(st.data(), seed);
    std::vector<double> cp;
    if (mode == 3) {
        double sum = 0.0;
        for (int i = 0; i < nprobs; ++i) sum += probs[i];
        double acc = 0.0;
        for (int i = 0; i < nprobs; ++i) { acc += probs[i] / sum; cp.push_back(acc); }
        cp.back() = 1.0;
    }
    for (int i = 0; i < count; ++i) {
        switch (mode) {
            case 0: out[i] = (double)fkrng::mt_next(st.data()); break;
            case 1: out[i] = fkrng::uniform_int(st.data(), 0, hi); break;
            case 2: out[i] = fkrng::canonical(st.data()); break;
            case 3: out[i] = fkrng::discrete(st.data(), cp.data(), nprobs < 2 ? 1 : nprobs); break;
        }
    }
    return FKGPU_OK;
}

}  // extern "C"
