// C-ABI layer of libspecfab_gpu (declared in include/specfab_gpu.h).
//
// A context owns: the device copies of the per-L real-basis operators and of
// the coupling-table CSR (uploaded lazily, once per L), a HIP stream, and a
// grow-only device workspace for per-parcel parameters and RK scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/specfab_gpu.h"
#include "sfg_internal.h"
#include "sfg_kernels.h"
#include "sfg_fast.h"
#include "sfg_homog.h"
#include "sfg_init.h"
#include "sfg_reduced.h"
#include "sfg_rotate.h"
#include "sfg_discrete.h"
#include "sfg_noalloc.h"
#include "sfg_wave.h"
#include "sfg_mlp_net.h"
#include "sfg_surrogate.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(SFG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t upload(T** dptr, const std::vector<T>& v) {
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    hipError_t e = hipMalloc((void**)dptr, bytes);
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

struct DevOp {
    int L = 0, n = 0, nblk = 0, nq = 0;
    double reg_nu = 0;
    bool reg_defined = false;
    int2* rows = nullptr;
    int* row_blk = nullptr;
    int* grp_param = nullptr;
    int2* grp_rng = nullptr;
    int* tcol = nullptr;
    double* tval = nullptr;
    float* tval_f = nullptr;
    double* A2 = nullptr;
    double* A4 = nullptr;
    double* A4F = nullptr;
    double* reg_shape = nullptr;
    std::vector<double> reg_shape_h;
    int* r2c = nullptr;
    int* rpart = nullptr;
    int *cre = nullptr, *cim = nullptr, *csgn = nullptr, *cneg = nullptr;
    int* cl = nullptr;   // l of each complex index
    std::vector<void*> owned;
};

}  // namespace

struct sfg_ctx {
    int device = 0;
    int stepper = 0;      // SFG_STEPPER: 0 = fastest available, 1 = generic only
    int variant_env = -1; // SFG_VARIANT: force one generated stepper variant
    std::map<std::pair<int, int>, int> variant;   // (L, precision) -> tuned variant
    bool sync_check = false;  // SFG_SYNC_CHECK=1: synchronise and check after each launch
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::map<int, DevOp> ops;
    // coupling-table CSR by output row i (all four tables)
    int* csr_row_ptr = nullptr;
    int4* csr_ent = nullptr;
    double* csr_val = nullptr;
    int2* csr_seg = nullptr;
    // reduced-block coefficient maps per (L, mode), ELL tables (sfg::RMap)
    std::map<std::pair<int, int>, sfg::RMap> rmap;
    // workspace, shared by all calls on this context: every user waits for
    // the previous user's stream through ws_event before touching it
    void* ws = nullptr;
    size_t ws_bytes = 0;
    hipEvent_t ws_event = nullptr;
    bool ws_event_recorded = false;
    // rotation blocks of recently used direction sets (wave velocities)
    struct DirSet { std::vector<double> th, ph; double* R = nullptr; };
    std::vector<DirSet> dirs;
    // Wigner sum-formula coefficients (rotate_nlm), uploaded on first use
    double* wigner = nullptr;
    // stepping-kernel timing (sfg_ctx_set_timing)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev;
};

namespace {

int get_op(sfg_ctx* ctx, int L, const DevOp** out) {
    if (L < 0 || L > sfg::LMAX_TABLE || (L % 2)) return fail(SFG_ERR_L, "L must be even and 0 <= L <= 20");
    auto it = ctx->ops.find(L);
    if (it != ctx->ops.end()) {
        *out = &it->second;
        return SFG_OK;
    }
    const sfg::RealOperator& op = sfg::real_operator(L);
    DevOp d;
    d.L = L;
    d.n = op.n;
    d.nblk = L / 2 + 1;
    d.nq = std::min(op.n, 15);
    d.reg_nu = op.reg_nu;
    d.reg_shape_h = op.reg_shape;
    d.reg_defined = op.reg_defined;
    std::vector<int2> rows;
    std::vector<int> row_blk, gparam, tcol, r2c, rpart;
    std::vector<int2> grng;
    std::vector<double> tval;
    for (auto& r : op.rows) { rows.push_back(make_int2(r.g0, r.g1)); row_blk.push_back(r.l / 2); }
    for (auto& g : op.groups) { gparam.push_back(g.param); grng.push_back(make_int2(g.t0, g.t1)); }
    for (auto& t : op.terms) { tcol.push_back(t.col); tval.push_back(t.val); }
    auto lm = sfg::lm_list(L);
    for (auto& ri : op.ridx) {
        int j = -1;
        for (int q = 0; q < op.n; q++)
            if (lm[q].l == ri.l && lm[q].m == ri.m) j = q;
        r2c.push_back(j);
        rpart.push_back(ri.c);
    }
    std::vector<double> A2(6 * 15), A4(36 * 15);
    for (int i = 0; i < 6; i++) for (int q = 0; q < 15; q++) A2[i * 15 + q] = op.A2[i][q];
    for (int i = 0; i < 36; i++) for (int q = 0; q < 15; q++) A4[i * 15 + q] = op.A4[i][q];
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(upload(&d.rows, rows));
    HIPCHK(upload(&d.row_blk, row_blk));
    HIPCHK(upload(&d.grp_param, gparam));
    HIPCHK(upload(&d.grp_rng, grng));
    HIPCHK(upload(&d.tcol, tcol));
    HIPCHK(upload(&d.tval, tval));
    std::vector<float> tvalf(tval.begin(), tval.end());
    HIPCHK(upload(&d.tval_f, tvalf));
    HIPCHK(upload(&d.A2, A2));
    HIPCHK(upload(&d.A4, A4));
    {
        double m[15][15];
        sfg::a4full_map(m);
        std::vector<double> A4F(15 * 15);
        for (int e = 0; e < 15; e++)
            for (int q = 0; q < 15; q++) A4F[e * 15 + q] = q < d.nq ? m[e][q] : 0.0;
        HIPCHK(upload(&d.A4F, A4F));
    }
    HIPCHK(upload(&d.reg_shape, op.reg_shape));
    HIPCHK(upload(&d.r2c, r2c));
    HIPCHK(upload(&d.rpart, rpart));
    HIPCHK(upload(&d.cre, op.cre));
    HIPCHK(upload(&d.cim, op.cim));
    HIPCHK(upload(&d.csgn, op.csgn));
    HIPCHK(upload(&d.cneg, op.cneg));
    std::vector<int> cl;
    for (auto& v : lm) cl.push_back(v.l);
    HIPCHK(upload(&d.cl, cl));
    d.owned = {d.rows, d.row_blk, d.grp_param, d.grp_rng, d.tcol, d.tval, d.tval_f, d.A2, d.A4, d.A4F, d.reg_shape,
               d.r2c, d.rpart, d.cre, d.cim, d.csgn, d.cneg, d.cl};
    auto res = ctx->ops.emplace(L, std::move(d));
    *out = &res.first->second;
    return SFG_OK;
}

int ensure_csr(sfg_ctx* ctx) {
    if (ctx->csr_row_ptr) return SFG_OK;
    const auto& ents = sfg::gaunt_entries();
    std::vector<const sfg::GauntEntry*> v;
    for (auto& e : ents) v.push_back(&e);
    std::stable_sort(v.begin(), v.end(), [](const sfg::GauntEntry* a, const sfg::GauntEntry* b) {
        if (a->i != b->i) return a->i < b->i;
        if (a->j != b->j) return a->j < b->j;
        if (a->t != b->t) return a->t < b->t;
        return a->k < b->k;
    });
    std::vector<int> rp(sfg::NMAX + 1, 0);
    std::vector<int4> en;
    std::vector<double> va;
    for (auto* e : v) {
        rp[e->i + 1]++;
        en.push_back(make_int4(e->t, e->j, e->k, 0));
        va.push_back(e->v);
    }
    for (int i = 0; i < sfg::NMAX; i++) rp[i + 1] += rp[i];
    // [begin, end) of each (row, column) run; the sort makes every run contiguous
    std::vector<int2> seg((size_t)sfg::NMAX * sfg::NMAX, make_int2(0, 0));
    for (int e = 0; e < (int)v.size(); e++) {
        int2& r = seg[(size_t)v[e]->i * sfg::NMAX + v[e]->j];
        if (r.y == 0) r.x = e;
        r.y = e + 1;
    }
    HIPCHK(upload(&ctx->csr_seg, seg));
    HIPCHK(upload(&ctx->csr_row_ptr, rp));
    HIPCHK(upload(&ctx->csr_ent, en));
    HIPCHK(upload(&ctx->csr_val, va));
    return SFG_OK;
}

// Acquire the context workspace for work enqueued on stream s; the caller
// enqueues its kernels and then calls ws_release(ctx, s).
int ws_acquire(sfg_ctx* ctx, size_t bytes, hipStream_t s, void** out) {
    if (bytes > ctx->ws_bytes) {
        if (ctx->ws) {
            // the old buffer may still be read by work on any stream
            HIPCHK(hipDeviceSynchronize());
            HIPCHK(hipFree(ctx->ws));
            ctx->ws = nullptr;
            ctx->ws_bytes = 0;
        }
        HIPCHK(hipMalloc(&ctx->ws, bytes));
        ctx->ws_bytes = bytes;
    }
    if (ctx->ws_event_recorded) HIPCHK(hipStreamWaitEvent(s, ctx->ws_event, 0));
    *out = ctx->ws;
    return SFG_OK;
}

int ws_release(sfg_ctx* ctx, hipStream_t s) {
    HIPCHK(hipEventRecord(ctx->ws_event, s));
    ctx->ws_event_recorded = true;
    return SFG_OK;
}

// The workspace held by one call: released (the event recorded on the call's
// stream) on every path out of the call, also the error returns, since a
// kernel already queued on that stream may still be using it.
struct WsLease {
    sfg_ctx* ctx;
    hipStream_t s;
    bool held = false;
    WsLease(sfg_ctx* c, hipStream_t st) : ctx(c), s(st) {}
    int acquire(size_t bytes, void** out) {
        int rc = ws_acquire(ctx, bytes, s, out);
        held = rc == SFG_OK;
        return rc;
    }
    int release() {
        held = false;
        return ws_release(ctx, s);
    }
    ~WsLease() {
        if (held) (void)ws_release(ctx, s);
    }
};

// Device rotation blocks (ROT4_STRIDE doubles per direction) of the direction
// set (theta, phi); each direction rotates by (-theta, -phi) as
// Vi_*_tranisotropic do (src/wavepropagation.f90:139, 235).  Recent sets are
// kept on the context.
int get_dirs(sfg_ctx* ctx, int ndir, const double* theta, const double* phi, const double** R) {
    for (auto& ds : ctx->dirs)
        if ((int)ds.th.size() == ndir && std::equal(ds.th.begin(), ds.th.end(), theta) &&
            std::equal(ds.ph.begin(), ds.ph.end(), phi)) {
            *R = ds.R;
            return SFG_OK;
        }
    // the rotation blocks, then the unit directions (sin t cos p, sin t sin p, cos t)
    std::vector<double> host((size_t)ndir * (sfg::ROT4_STRIDE + 3));
    for (int d = 0; d < ndir; d++) {
        double R2[5][5], R4[9][9];
        sfg::rotation_l4(-theta[d], -phi[d], R2, R4);
        double* o = host.data() + (size_t)d * sfg::ROT4_STRIDE;
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) o[5 * i + j] = R2[i][j];
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) o[25 + 9 * i + j] = R4[i][j];
        double* k = host.data() + (size_t)ndir * sfg::ROT4_STRIDE + 3 * d;
        k[0] = std::sin(theta[d]) * std::cos(phi[d]);
        k[1] = std::sin(theta[d]) * std::sin(phi[d]);
        k[2] = std::cos(theta[d]);
    }
    if (ctx->dirs.size() >= 16) {
        // the oldest set may still be read by queued kernels
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(ctx->dirs.front().R));
        ctx->dirs.erase(ctx->dirs.begin());
    }
    sfg_ctx::DirSet ds;
    ds.th.assign(theta, theta + ndir);
    ds.ph.assign(phi, phi + ndir);
    HIPCHK(upload(&ds.R, host));
    ctx->dirs.push_back(ds);
    *R = ctx->dirs.back().R;
    return SFG_OK;
}

// unit propagation directions (ndir, 3) of the direction set, stored after its
// rotation blocks
int get_kdirs(sfg_ctx* ctx, int ndir, const double* theta, const double* phi, const double** K) {
    const double* R;
    int rc = get_dirs(ctx, ndir, theta, phi, &R);
    if (rc) return rc;
    *K = R + (size_t)ndir * sfg::ROT4_STRIDE;
    return SFG_OK;
}

// Batched calls run on the caller's stream; NULL is HIP's default stream (the
// convention of the HIP runtime and of PyTorch's default stream).
hipStream_t pick(sfg_ctx*, void* stream) { return (hipStream_t)stream; }

}  // namespace

extern "C" {

const char* sfg_last_error(void) { return g_err.c_str(); }

int sfg_ctx_set_timing(sfg_ctx* ctx, int enable) {
    if (!ctx) return fail(SFG_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->timing = enable != 0;
    return SFG_OK;
}

int sfg_ctx_read_timing(sfg_ctx* ctx, double* ms_total, int* count) {
    if (!ctx || !ms_total || !count) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    double tot = 0;
    for (auto& t : ctx->tev) {
        HIPCHK(hipEventSynchronize(t.second));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.first, t.second));
        tot += ms;
    }
    *ms_total = tot;
    *count = (int)ctx->tev.size();
    for (auto& t : ctx->tev) { (void)hipEventDestroy(t.first); (void)hipEventDestroy(t.second); }
    ctx->tev.clear();
    return SFG_OK;
}

int sfg_nlm_len(int L) {
    if (L < 0 || (L % 2)) return fail(SFG_ERR_L, "L must be even and >= 0");
    return sfg::nlm_len(L);
}

int sfg_lm(int L, int* lm) {
    if (!lm) return fail(SFG_ERR_ARG, "lm is NULL");
    if (L < 0 || (L % 2)) return fail(SFG_ERR_L, "L must be even and >= 0");
    int j = 0;
    for (auto& v : sfg::lm_list(L)) { lm[2 * j] = v.l; lm[2 * j + 1] = v.m; j++; }
    return SFG_OK;
}

int sfg_ctx_create(int device, sfg_ctx** ctx) {
    if (!ctx) return fail(SFG_ERR_ARG, "ctx is NULL");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SFG_ERR_ARG, "no such HIP device");
    HIPCHK(hipSetDevice(device));
    auto* c = new sfg_ctx();
    c->device = device;
    if (const char* e = std::getenv("SFG_STEPPER")) c->stepper = (std::string(e) == "generic") ? 1 : 0;
    if (const char* e = std::getenv("SFG_SYNC_CHECK")) c->sync_check = std::string(e) == "1";
    if (const char* e = std::getenv("SFG_VARIANT")) c->variant_env = std::atoi(e);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ws_event, hipEventDisableTiming);
    if (e != hipSuccess) { delete c; return fail(SFG_ERR_HIP, hipGetErrorString(e)); }
    // the no-allocation workspace of this device (sfg_evolve_noalloc)
    if (int rc = sfg::noalloc_setup(device, nullptr)) {
        (void)hipEventDestroy(c->ws_event);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *ctx = c;
    return SFG_OK;
}

int sfg_ctx_destroy(sfg_ctx* ctx) {
    if (!ctx) return SFG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto& t : ctx->tev) { (void)hipEventDestroy(t.first); (void)hipEventDestroy(t.second); }
    if (ctx->ws_event) (void)hipEventDestroy(ctx->ws_event);
    for (auto& kv : ctx->ops)
        for (void* p : kv.second.owned) (void)hipFree(p);
    (void)hipFree(ctx->csr_row_ptr);
    (void)hipFree(ctx->csr_ent);
    (void)hipFree(ctx->csr_val);
    (void)hipFree(ctx->csr_seg);
    for (auto& kv : ctx->rmap) {
        (void)hipFree((void*)kv.second.ci);
        (void)hipFree((void*)kv.second.cp);
        (void)hipFree((void*)kv.second.cm);
        (void)hipFree((void*)kv.second.nt);
    }
    if (ctx->ws) (void)hipFree(ctx->ws);
    for (auto& ds : ctx->dirs) (void)hipFree(ds.R);
    if (ctx->wigner) (void)hipFree(ctx->wigner);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SFG_OK;
}

// ---------------------------------------------------------------- conversions

int sfg_nlm_to_state(sfg_ctx* ctx, int L, int64_t P, const double* nlm_dev, double* x_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!nlm_dev || !x_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(sfg::launch_nlm_to_state(P, op->n, op->r2c, op->rpart, nlm_dev, x_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_state_to_nlm(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* nlm_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!nlm_dev || !x_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(sfg::launch_state_to_nlm(P, op->n, op->cre, op->cim, op->csgn, op->cneg, x_dev, nlm_dev,
                                    pick(ctx, stream)));
    return SFG_OK;
}

// ---------------------------------------------------------------- evolution

namespace {

// regularization law of a launch: the calibrated M_REG of L
// (src/include/regcalib.f90), or with reg_expo > 0 the law
// -reg_nu |D| (l(l+1) / (L(L+1)))^reg_expo (src/specfabpy/integrator.py:85-97)
void reg_law(const DevOp* op, const sfg_evolve_desc* d, sfg::ParamArgs& pa) {
    const bool custom = d->reg_expo > 0;
    pa.reg_nu = custom ? d->reg_nu : op->reg_nu;
    const int L = op->L;
    for (int b = 0; b < op->nblk && b < 11; b++) {
        const int l = 2 * b;
        const double Ld = -(double)(l * (l + 1));
        pa.reg_shape[b] = custom ? std::pow(std::fabs(Ld / (L * (L + 1))), d->reg_expo) : op->reg_shape_h[b];
    }
}

int variant_of(sfg_ctx* ctx, int L, int prec) {
    const int nv = sfg::fast_variants(L, prec);
    if (ctx->variant_env >= 0 && ctx->variant_env < nv) return ctx->variant_env;
    auto it = ctx->variant.find({L, prec});
    return (it != ctx->variant.end() && it->second < nv) ? it->second : 0;
}

// How one evolve call steps: the generated stepper variant (or the
// table-driven one), the stage buffers that variant uses, the parcels per
// launch, and the workspace: per-parcel launch constants, the used stage
// buffers (n x slice doubles each) and, for a batch in slices, the slice's
// state.  The register-resident steppers address SoA rows with 32-bit buffer
// offsets (and use offsets >= 2^31 as "out of range"), so one launch steps at
// most (2^31 - 1) / (8 n) parcels; larger batches are stepped slice by slice,
// each slice's state copied into the workspace and back.
struct StepPlan {
    const DevOp* op = nullptr;
    bool fast = false;
    int prec = 0, variant = 0, mask = 0;
    int64_t slice = 0;
    bool sliced = false;
    size_t ws_doubles = 0;
};

// resident_only: fail instead of uploading the tables of L (no allocation)
int plan_step(sfg_ctx* ctx, const sfg_evolve_desc* d, int force_variant, bool resident_only, StepPlan& sp) {
    if (resident_only) {
        auto it = ctx->ops.find(d->L);
        if (it == ctx->ops.end())
            return fail(SFG_ERR_ARG, "the tables of this L are not on the device yet: make one allocating call at "
                                     "this L first (e.g. sfg_evolve with P = 0)");
        sp.op = &it->second;
    } else {
        int rc = get_op(ctx, d->L, &sp.op);
        if (rc) return rc;
    }
    const DevOp* op = sp.op;
    if ((d->procs & SFG_REG) && !op->reg_defined && !(d->reg_expo > 0))
        return fail(SFG_ERR_L, "M_REG is calibrated only for L = 4, 6, ..., 20 (src/include/regcalib.f90)");
    const int64_t P = d->P;
    const int n = op->n;
    sp.prec = d->precision == SFG_FP64 ? 0 : 1;
    sp.fast = sfg::fast_available(d->L, sp.prec) && (sp.prec == 1 || ctx->stepper == 0);
    if (!sp.fast && sp.prec == 1) return fail(SFG_ERR_ARG, "FP32C stepping needs a generated stepper for L");
    const int64_t maxp = (((int64_t(1) << 31) - 1) / (8 * (int64_t)n)) / 256 * 256;
    sp.slice = (sp.fast && P > maxp) ? maxp : std::max<int64_t>(P, 1);
    sp.sliced = sp.slice < P;
    if (sp.fast) {
        sp.variant = force_variant >= 0 ? force_variant : variant_of(ctx, d->L, sp.prec);
        sp.mask = sfg::fast_buffers(d->L, sp.prec, sp.variant);
    } else {
        sp.mask = sfg::BUF_YS | sfg::BUF_XA | sfg::BUF_XB;
    }
    const int planes = __builtin_popcount(sp.mask) + (sp.sliced ? 1 : 0);
    sp.ws_doubles = (size_t)(sfg::NPAR_K + 15 + op->nblk) * sp.slice + (size_t)planes * n * sp.slice;
    return SFG_OK;
}

// enqueue the stepping of d on s with workspace ws (sp.ws_doubles doubles)
int run_step(sfg_ctx* ctx, const sfg_evolve_desc* d, hipStream_t s, const StepPlan& sp, double* ws, bool timing) {
    const DevOp* op = sp.op;
    const int64_t P = d->P;
    const int n = op->n;
    std::pair<hipEvent_t, hipEvent_t> tv{nullptr, nullptr};
    if (timing) {
        HIPCHK(hipEventCreate(&tv.first));
        HIPCHK(hipEventCreate(&tv.second));
        ctx->tev.push_back(tv);
        HIPCHK(hipEventRecord(tv.first, s));
    }
    for (int64_t p0 = 0; p0 < P; p0 += sp.slice) {
        const int64_t Ps = std::min(sp.slice, P - p0);
        double* theta = ws;
        double* omega = theta + (size_t)sfg::NPAR_K * Ps;
        double* sl = omega + (size_t)15 * Ps;
        double* plane = sl + (size_t)op->nblk * Ps;
        double* bufs[3] = {nullptr, nullptr, nullptr};
        for (int b = 0; b < 3; b++)
            if (sp.mask & (1 << b)) {
                bufs[b] = plane;
                plane += (size_t)n * Ps;
            }
        double* X = d->x_dev;
        if (sp.sliced) {
            X = plane;
            HIPCHK(hipMemcpy2DAsync(X, (size_t)Ps * 8, d->x_dev + p0, (size_t)P * 8, (size_t)Ps * 8, n,
                                    hipMemcpyDeviceToDevice, s));
        }
        sfg::ParamArgs pa{};
        pa.P = Ps; pa.nblk = op->nblk; pa.nq = op->nq; pa.procs = d->procs;
        pa.iota = d->iota; pa.zeta = d->zeta;
        reg_law(op, d, pa);
        pa.ugrad = d->ugrad_dev + 9 * p0;
        pa.tau = d->tau_dev ? d->tau_dev + 9 * p0 : nullptr;
        pa.gam0 = d->gam0_dev ? d->gam0_dev + p0 : nullptr;
        pa.lam = d->lam_dev ? d->lam_dev + p0 : nullptr;
        pa.A2 = op->A2; pa.A4 = op->A4;
        pa.theta = theta; pa.omega = omega; pa.sl = sl;
        HIPCHK(sfg::launch_params(pa, s));
        sfg::EvolveArgs ea{};
        ea.P = Ps; ea.n = n; ea.nq = op->nq; ea.scheme = d->scheme; ea.nsteps = d->nsteps;
        ea.use_ddrx = (d->procs & SFG_DDRX) ? 1 : 0;
        ea.steps = d->steps_dev ? d->steps_dev + p0 : nullptr;
        ea.dt = d->dt;
        ea.rows = op->rows; ea.row_blk = op->row_blk; ea.grp_param = op->grp_param; ea.grp_rng = op->grp_rng;
        ea.tcol = op->tcol; ea.tval = op->tval; ea.tval_f = op->tval_f;
        ea.theta = theta; ea.omega = omega; ea.sl = sl;
        ea.X = X; ea.YS = bufs[0]; ea.XA = bufs[1]; ea.XB = bufs[2];
        ea.status = d->status_dev ? d->status_dev + p0 : nullptr;
        ea.zero = 0;
        if (sp.fast) HIPCHK(sfg::launch_evolve_fast(d->L, sp.prec, sp.variant, ea, s));
        else HIPCHK(sfg::launch_evolve_generic(ea, s));
        if (sp.sliced)
            HIPCHK(hipMemcpy2DAsync(d->x_dev + p0, (size_t)P * 8, X, (size_t)Ps * 8, (size_t)Ps * 8, n,
                                    hipMemcpyDeviceToDevice, s));
    }
    if (timing) HIPCHK(hipEventRecord(tv.second, s));
    if (ctx->sync_check) {
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipGetLastError());
    }
    return SFG_OK;
}

// body of sfg_evolve; ctx->mu held.  force_variant >= 0 selects a generated
// stepper variant (autotuner).
int evolve_locked(sfg_ctx* ctx, const sfg_evolve_desc* d, hipStream_t s, int force_variant) {
    StepPlan sp;
    int rc = plan_step(ctx, d, force_variant, false, sp);
    if (rc) return rc;
    if (d->P == 0) return SFG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    void* ws;
    WsLease lease(ctx, s);
    rc = lease.acquire(sp.ws_doubles * sizeof(double), &ws);
    if (rc) return rc;
    rc = run_step(ctx, d, s, sp, (double*)ws, ctx->timing);
    if (rc) return rc;
    return lease.release();
}

}  // namespace

extern "C++" {
namespace sfg {

int set_error(int code, const char* msg) { return fail(code, msg); }

int ctx_device(const sfg_ctx* ctx) { return ctx->device; }

int evolve_in_workspace(sfg_ctx* ctx, const sfg_evolve_desc* d, void* stream, void* ws, size_t ws_bytes) {
    if (!ctx || !d) return fail(SFG_ERR_ARG, "ctx/desc is NULL");
    if (d->P < 0 || (d->P > 0 && (!d->ugrad_dev || !d->x_dev))) return fail(SFG_ERR_ARG, "ugrad_dev/x_dev missing");
    if (d->scheme != SFG_EULER && d->scheme != SFG_RK4) return fail(SFG_ERR_ARG, "unknown scheme");
    if (d->precision != SFG_FP64 && d->precision != SFG_FP32C) return fail(SFG_ERR_ARG, "unknown precision");
    if (d->nsteps < 0) return fail(SFG_ERR_ARG, "nsteps < 0");
    if (d->procs & ~(SFG_LROT | SFG_DDRX | SFG_CDRX | SFG_REG)) return fail(SFG_ERR_ARG, "unknown process bit");
    std::lock_guard<std::mutex> lk(ctx->mu);
    StepPlan sp;
    int rc = plan_step(ctx, d, -1, true, sp);
    if (rc) return rc;
    if (d->P == 0) return SFG_OK;
    const size_t need = sp.ws_doubles * sizeof(double);
    if (need > ws_bytes) {
        char msg[256];
        snprintf(msg, sizeof(msg), "the batch needs %zu bytes of workspace, more than the %zu available: split it",
                 need, ws_bytes);
        return fail(SFG_ERR_ARG, msg);
    }
    HIPCHK(hipSetDevice(ctx->device));
    return run_step(ctx, d, pick(ctx, stream), sp, (double*)ws, false);
}

}  // namespace sfg
}  // extern "C++"

int sfg_evolve(sfg_ctx* ctx, const sfg_evolve_desc* d, void* stream) {
    if (!ctx || !d) return fail(SFG_ERR_ARG, "ctx/desc is NULL");
    if (d->P < 0 || (d->P > 0 && (!d->ugrad_dev || !d->x_dev))) return fail(SFG_ERR_ARG, "ugrad_dev/x_dev missing");
    if (d->scheme != SFG_EULER && d->scheme != SFG_RK4) return fail(SFG_ERR_ARG, "unknown scheme");
    if (d->precision != SFG_FP64 && d->precision != SFG_FP32C) return fail(SFG_ERR_ARG, "unknown precision");
    if (d->nsteps < 0) return fail(SFG_ERR_ARG, "nsteps < 0");
    if (d->procs & ~(SFG_LROT | SFG_DDRX | SFG_CDRX | SFG_REG)) return fail(SFG_ERR_ARG, "unknown process bit");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return evolve_locked(ctx, d, pick(ctx, stream), -1);
}

int sfg_autotune(sfg_ctx* ctx, int L, int precision, int64_t P, int steps, void* stream, int* best, double* ms,
                 int ms_len) {
    if (!ctx || P <= 0 || steps <= 0 || (precision != SFG_FP64 && precision != SFG_FP32C))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int prec = precision == SFG_FP64 ? 0 : 1;
    const int nv = sfg::fast_variants(L, prec);
    if (best) *best = -1;
    if (nv == 0) return SFG_OK;
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = pick(ctx, stream);
    const int n = op->n;
    // synthetic batch: perturbed isotropic states, traceless gradients
    std::vector<double> hx((size_t)n * P), hg((size_t)9 * P), hgam((size_t)P);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return ((double)(st >> 11) * (1.0 / 9007199254740992.0)) - 0.5;
    };
    for (int64_t p = 0; p < P; p++) {
        hx[p] = 0.28209479177387814;
        for (int a = 1; a < n; a++) hx[(size_t)a * P + p] = 2e-3 * rnd();
        double t = 0;
        for (int i = 0; i < 9; i++) hg[9 * p + i] = 2e-3 * rnd();
        for (int i = 0; i < 3; i++) t += hg[9 * p + 4 * i];
        for (int i = 0; i < 3; i++) hg[9 * p + 4 * i] -= t / 3;
        hgam[p] = 1e-3;
    }
    DevBuf bx, bg, bgam, bst;
    HIPCHK(hipMalloc(&bx.p, hx.size() * 8));
    HIPCHK(hipMalloc(&bg.p, hg.size() * 8));
    HIPCHK(hipMalloc(&bgam.p, hgam.size() * 8));
    HIPCHK(hipMalloc(&bst.p, (size_t)P * 4));
    HIPCHK(hipMemcpy(bx.p, hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(bg.p, hg.data(), hg.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(bgam.p, hgam.data(), hgam.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(bst.p, 0, (size_t)P * 4));
    sfg_evolve_desc d{};
    d.L = L; d.P = P; d.procs = SFG_LROT | SFG_REG | SFG_DDRX; d.scheme = SFG_RK4; d.precision = precision;
    d.nsteps = steps; d.dt = 0.5; d.iota = 1.0; d.zeta = 0.0;
    d.ugrad_dev = (double*)bg.p; d.gam0_dev = (double*)bgam.p; d.x_dev = (double*)bx.p; d.status_dev = (int32_t*)bst.p;
    const bool timing = ctx->timing;
    ctx->timing = false;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double tbest = 1e300;
    int vbest = 0;
    for (int v = 0; v < nv; v++) {
        rc = evolve_locked(ctx, &d, s, v);   // warm-up (code object load)
        if (rc) break;
        HIPCHK(hipEventRecord(e0, s));
        for (int k = 0; k < 3 && !rc; k++) rc = evolve_locked(ctx, &d, s, v);
        if (rc) break;
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        if (ms && v < ms_len) ms[v] = t / (3 * steps);
        if (t < tbest) { tbest = t; vbest = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timing = timing;
    if (rc) return rc;
    ctx->variant[{L, prec}] = vbest;
    if (best) *best = vbest;
    return SFG_OK;
}

// ---------------------------------------------------------------- moments

int sfg_a2_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a2_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !a2_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (L < 2) return fail(SFG_ERR_L, "a2 needs L >= 2");
    HIPCHK(sfg::launch_a2(P, x_dev, a2_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_a4mandel_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a4v_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !a4v_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(sfg::launch_a4mandel(P, op->nq, op->A4, x_dev, a4v_dev, pick(ctx, stream)));
    return SFG_OK;
}

// ---------------------------------------------------------------- dense operators

namespace {


int dense_batch(sfg_ctx* ctx, int L, int64_t P, int mode, const double* A_dev, const double* B_dev, double iota,
                double zeta, double* M_dev, hipStream_t s) {
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    rc = ensure_csr(ctx);
    if (rc) return rc;
    sfg::DenseArgs a{};
    a.P = P; a.n = op->n; a.mode = mode; a.iota = iota; a.zeta = zeta;
    a.A = A_dev; a.B = B_dev; a.row_ptr = ctx->csr_row_ptr; a.ent = ctx->csr_ent; a.val = ctx->csr_val; a.seg = ctx->csr_seg; a.segw = sfg::NMAX;
    a.M = M_dev;
    HIPCHK(sfg::launch_dense_op(a, s));
    return SFG_OK;
}

// single-parcel helper: stage 3x3 inputs, build on the GPU, copy M back
int dense_single(sfg_ctx* ctx, int L, int mode, const double* A, const double* B, double iota, double zeta,
                 double* M) {
    if (!ctx || !A || !M || (mode == 0 && !B)) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    int n = sfg::nlm_len(L);
    DevBuf in, out;
    HIPCHK(hipMalloc(&in.p, 18 * sizeof(double)));
    HIPCHK(hipMalloc(&out.p, (size_t)n * n * 2 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(in.p, A, 9 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (B) HIPCHK(hipMemcpyAsync((double*)in.p + 9, B, 9 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = dense_batch(ctx, L, 1, mode, (double*)in.p, (double*)in.p + 9, iota, zeta, (double*)out.p, ctx->stream);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(M, out.p, (size_t)n * n * 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return SFG_OK;
}

}  // namespace

int sfg_M_LROT_batch(sfg_ctx* ctx, int L, int64_t P, const double* eps_dev, const double* omg_dev, double iota,
                     double zeta, double* M_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!eps_dev || !omg_dev || !M_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return dense_batch(ctx, L, P, 0, eps_dev, omg_dev, iota, zeta, M_dev, pick(ctx, stream));
}

// the coefficient maps Cp, Cm (K, rn^2) of the reduced blocks of mode's
// operator at truncation L (see launch_reduced_map), from the coupling-table
// entries: A(i, c)[comp] = sum of the values of the entries of element (i, c)
// with that weight component; kept as the ELL table of their nonzero terms
static int get_rmap(sfg_ctx* ctx, int L, int mode, const sfg::RMap** out) {
    auto key = std::make_pair(L, mode);
    auto it = ctx->rmap.find(key);
    if (it != ctx->rmap.end()) {
        *out = &it->second;
        return SFG_OK;
    }
    const int K = mode == 0 ? 24 : 15, kmax = mode == 0 ? 6 : 15;
    const double sg = mode == 0 ? -1.0 : 1.0;
    const int n = sfg::nlm_len(L), rn = sfg::rnlm_len(L), rn2 = rn * rn;
    // A for the rows at m >= 0 positions: (n, n, K) restricted to needed rows
    std::vector<double> A((size_t)n * n * K, 0.0);
    for (auto& e : sfg::gaunt_entries()) {
        if (e.i >= n || e.j >= n || e.k >= kmax || (mode == 1 && e.t != 0)) continue;
        const int comp = mode == 0 ? e.t * 6 + e.k : e.k;
        A[((size_t)e.i * n + e.j) * K + comp] += e.v;
    }
    auto lm = [](int k, int* l, int* m) {
        int h = (int)std::sqrt((double)k);
        while ((h + 1) * (h + 1) <= k) ++h;
        while (h * h > k) --h;
        *l = 2 * h;
        *m = k - h * h;
    };
    std::vector<double> C((size_t)2 * K * rn2, 0.0);
    for (int i = 0; i < rn; i++) {
        int li, mi;
        lm(i, &li, &mi);
        const int fi = li * (li - 1) / 2 + li + mi;
        for (int j = 0; j < rn; j++) {
            int lj, mj;
            lm(j, &lj, &mj);
            const int cp = lj * (lj - 1) / 2 + lj + mj, cn = cp - 2 * mj;
            const double sgn = mj == 0 ? 0.0 : ((mj & 1) ? -1.0 : 1.0);
            for (int c = 0; c < K; c++) {
                const double ap = A[((size_t)fi * n + cp) * K + c], an = A[((size_t)fi * n + cn) * K + c];
                C[(size_t)c * rn2 + i * rn + j] = sg * (ap + sgn * an);
                C[(size_t)(K + c) * rn2 + i * rn + j] = sg * (ap - sgn * an);
            }
        }
    }
    // ELL: entry r's components c with Cp or Cm nonzero, in increasing c
    const int S = mode == 0 ? 10 : 5, nw = (rn2 + 63) / 64;
    std::vector<int> ci((size_t)S * rn2, 0), nt(nw, 0);
    std::vector<double> cp((size_t)S * rn2, 0.0), cm((size_t)S * rn2, 0.0);
    for (int r = 0; r < rn2; r++) {
        int k = 0;
        for (int c = 0; c < K; c++) {
            const double vp = C[(size_t)c * rn2 + r], vm = C[(size_t)(K + c) * rn2 + r];
            if (vp == 0.0 && vm == 0.0) continue;
            if (k == S) return fail(SFG_ERR_L, "reduced map: more nonzero terms per entry than slots");
            ci[(size_t)k * rn2 + r] = c;
            cp[(size_t)k * rn2 + r] = vp;
            cm[(size_t)k * rn2 + r] = vm;
            k++;
        }
        nt[r >> 6] = std::max(nt[r >> 6], k);
    }
    sfg::RMap m{};
    m.rn2 = rn2;
    m.S = S;
    int *dci = nullptr, *dnt = nullptr;
    double *dcp = nullptr, *dcm = nullptr;
    HIPCHK(upload(&dci, ci));
    HIPCHK(upload(&dcp, cp));
    HIPCHK(upload(&dcm, cm));
    HIPCHK(upload(&dnt, nt));
    m.ci = dci; m.cp = dcp; m.cm = dcm; m.nt = dnt;
    auto ins = ctx->rmap.emplace(key, m).first;
    *out = &ins->second;
    return SFG_OK;
}

// the reduced blocks of M_LROT / M_DDRX_src of every node without the full
// matrices: the nodes' coupling weights (dense_weights_kernel), then the
// blocks as a linear map of them (reduced_map_kernel)
static int dense_reduced_batch(sfg_ctx* ctx, int L, int64_t P, int mode, const double* A_dev, const double* B_dev,
                               double iota, double zeta, double* Mrr, double* Mri, double* Mir, double* Mii,
                               hipStream_t s) {
    if (P < 0 || (P > 0 && (!A_dev || (mode == 0 && !B_dev) || !Mrr || !Mri || !Mir || !Mii)))
        return fail(SFG_ERR_ARG, "bad argument");
    if (P == 0) return SFG_OK;
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    rc = ensure_csr(ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const sfg::RMap* rm;
    rc = get_rmap(ctx, L, mode, &rm);
    if (rc) return rc;
    sfg::DenseArgs a{};
    a.P = P; a.n = op->n; a.mode = mode; a.iota = iota; a.zeta = zeta;
    a.A = A_dev; a.B = B_dev;
    void* ws;
    WsLease lease(ctx, s);
    rc = lease.acquire((size_t)P * sfg::RMAP_W * sizeof(double), &ws);
    if (rc) return rc;
    HIPCHK(sfg::launch_dense_weights(a, (double*)ws, s));
    HIPCHK(sfg::launch_reduced_map(P, *rm, (const double*)ws, Mrr, Mri, Mir, Mii, s));
    return lease.release();
}

int sfg_M_LROT_reduced_batch(sfg_ctx* ctx, int L, int64_t P, const double* eps_dev, const double* omg_dev,
                             double iota, double zeta, double* Mrr_dev, double* Mri_dev, double* Mir_dev,
                             double* Mii_dev, void* stream) {
    if (!ctx) return fail(SFG_ERR_ARG, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return dense_reduced_batch(ctx, L, P, 0, eps_dev, omg_dev, iota, zeta, Mrr_dev, Mri_dev, Mir_dev, Mii_dev,
                               pick(ctx, stream));
}

int sfg_M_DDRX_src_reduced_batch(sfg_ctx* ctx, int L, int64_t P, const double* tau_dev, double* Mrr_dev,
                                 double* Mri_dev, double* Mir_dev, double* Mii_dev, void* stream) {
    if (!ctx) return fail(SFG_ERR_ARG, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return dense_reduced_batch(ctx, L, P, 1, tau_dev, nullptr, 0, 0, Mrr_dev, Mri_dev, Mir_dev, Mii_dev,
                               pick(ctx, stream));
}

int sfg_M_DDRX_src_batch(sfg_ctx* ctx, int L, int64_t P, const double* tau_dev, double* M_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!tau_dev || !M_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return dense_batch(ctx, L, P, 1, tau_dev, nullptr, 0, 0, M_dev, pick(ctx, stream));
}

// ---------------------------------------------------------------- reduced form (src/reducedform.f90:76-183)

namespace {

int reduced_L(int L) {
    if (L < 0 || L > 30 || (L % 2)) return fail(SFG_ERR_L, "L must be even and 0 <= L <= 30 (src/reducedform.f90:12)");
    return SFG_OK;
}

// single node: stage `in_len` doubles, run `fn` on the device copies, copy `out_len` doubles back
extern "C++" template <class F>
int reduced_single(sfg_ctx* ctx, const double* in, size_t in_len, double* out, size_t out_len, F fn) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (in_len + out_len) * sizeof(double)));
    double* din = (double*)buf.p;
    double* dout = din + in_len;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(din, in, in_len * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(fn(din, dout, s));
    HIPCHK(hipMemcpyAsync(out, dout, out_len * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

}  // namespace

int sfg_rnlm_len(int L) {
    int rc = reduced_L(L);
    return rc ? rc : sfg::rnlm_len(L);
}

int sfg_reduce_M_batch(sfg_ctx* ctx, int L, int64_t P, const double* M_dev, double* Mrr_dev, double* Mri_dev,
                       double* Mir_dev, double* Mii_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!M_dev || !Mrr_dev || !Mri_dev || !Mir_dev || !Mii_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    if (int rc = reduced_L(L)) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_reduce_M(P, L, M_dev, Mrr_dev, Mri_dev, Mir_dev, Mii_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_nlm_to_rnlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* nlm_dev, double* rnlm_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!nlm_dev || !rnlm_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (int rc = reduced_L(L)) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_nlm_to_rnlm(P, L, nlm_dev, rnlm_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_rnlm_to_nlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* rnlm_dev, double* nlm_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!nlm_dev || !rnlm_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (int rc = reduced_L(L)) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_rnlm_to_nlm(P, L, rnlm_dev, nlm_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_reduce_M(sfg_ctx* ctx, int L, const double* M, double* Mrr, double* Mri, double* Mir, double* Mii) {
    if (!ctx || !M || !Mrr || !Mri || !Mir || !Mii) return fail(SFG_ERR_ARG, "NULL argument");
    if (int rc = reduced_L(L)) return rc;
    size_t n = sfg::nlm_len(L), r = sfg::rnlm_len(L), rr = r * r;
    std::vector<double> out(4 * rr);
    int rc = reduced_single(ctx, M, 2 * n * n, out.data(), 4 * rr, [&](const double* din, double* dout, hipStream_t s) {
        return sfg::launch_reduce_M(1, L, din, dout, dout + rr, dout + 2 * rr, dout + 3 * rr, s);
    });
    if (rc) return rc;
    std::copy(out.begin(), out.begin() + rr, Mrr);
    std::copy(out.begin() + rr, out.begin() + 2 * rr, Mri);
    std::copy(out.begin() + 2 * rr, out.begin() + 3 * rr, Mir);
    std::copy(out.begin() + 3 * rr, out.end(), Mii);
    return SFG_OK;
}

int sfg_nlm_to_rnlm(sfg_ctx* ctx, int L, const double* nlm, double* rnlm) {
    if (!ctx || !nlm || !rnlm) return fail(SFG_ERR_ARG, "NULL argument");
    if (int rc = reduced_L(L)) return rc;
    return reduced_single(ctx, nlm, 2 * (size_t)sfg::nlm_len(L), rnlm, 2 * (size_t)sfg::rnlm_len(L),
                          [&](const double* din, double* dout, hipStream_t s) {
                              return sfg::launch_nlm_to_rnlm(1, L, din, dout, s);
                          });
}

int sfg_rnlm_to_nlm(sfg_ctx* ctx, int L, const double* rnlm, double* nlm) {
    if (!ctx || !nlm || !rnlm) return fail(SFG_ERR_ARG, "NULL argument");
    if (int rc = reduced_L(L)) return rc;
    return reduced_single(ctx, rnlm, 2 * (size_t)sfg::rnlm_len(L), nlm, 2 * (size_t)sfg::nlm_len(L),
                          [&](const double* din, double* dout, hipStream_t s) {
                              return sfg::launch_rnlm_to_nlm(1, L, din, dout, s);
                          });
}

int sfg_M_LROT(sfg_ctx* ctx, int L, const double eps[9], const double omg[9], double iota, double zeta, double* M) {
    return dense_single(ctx, L, 0, eps, omg, iota, zeta, M);
}

int sfg_M_DDRX_src(sfg_ctx* ctx, int L, const double tau[9], double* M) {
    return dense_single(ctx, L, 1, tau, nullptr, 0, 0, M);
}

// <D> and a2 of one host nlm, evaluated on the GPU (state conversion + kernel)
static int single_state_eval(sfg_ctx* ctx, int L, const double* nlm, const double* tau, double* D, double* a2) {
    if (!ctx || !nlm) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(2 * n + n + 9 + 9 + 1) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* dt = dx + n;
    double* da2 = dt + 9;
    double* dD = da2 + 9;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    if (D) {
        HIPCHK(hipMemcpyAsync(dt, tau, 9 * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(sfg::launch_evD2(1, op->nq, op->A2, op->A4, dx, dt, dD, s));
        HIPCHK(hipMemcpyAsync(D, dD, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (a2) {
        HIPCHK(sfg::launch_a2(1, dx, da2, s));
        HIPCHK(hipMemcpyAsync(a2, da2, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

// ---------------------------------------------------------------- initial states, Gamma0, bounds

int sfg_reg_calibration(int L, double* expo, double* nu) {
    if (!expo || !nu) return fail(SFG_ERR_ARG, "NULL argument");
    if (L < 4 || L > sfg::LMAX_TABLE || (L & 1))
        return fail(SFG_ERR_L, "M_REG is calibrated only for L = 4, 6, ..., 20 (src/include/regcalib.f90)");
    const sfg::RealOperator& op = sfg::real_operator(L);
    if (!op.reg_defined)
        return fail(SFG_ERR_L, "M_REG is calibrated only for L = 4, 6, ..., 20 (src/include/regcalib.f90)");
    *expo = op.reg_expo;
    *nu = op.reg_nu;
    return SFG_OK;
}

int sfg_a2_to_state_batch(sfg_ctx* ctx, int L, int64_t P, const double* a2_dev, double* x_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!a2_dev || !x_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_a2_to_state(P, op->n, a2_dev, x_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_a4_to_state_batch(sfg_ctx* ctx, int L, int64_t P, const double* a4_dev, double* x_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!a4_dev || !x_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (L < 4) return fail(SFG_ERR_L, "a4_to_nlm fills l <= 4: needs L >= 4");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_a4_to_state(P, op->n, a4_dev, x_dev, pick(ctx, stream)));
    return SFG_OK;
}

namespace {
// one tensor -> complex nlm (first ncoef coefficients, specfab order) through
// the batched kernel and the L = 4 state conversion
int tensor_to_nlm(sfg_ctx* ctx, const double* t, int tlen, int ncoef, double* nlm) {
    if (!ctx || !t || !nlm) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, 4, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(tlen + n + 2 * n) * sizeof(double)));
    double* dt = (double*)buf.p;
    double* dx = dt + tlen;
    double* dn = dx + n;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dt, t, tlen * sizeof(double), hipMemcpyHostToDevice, s));
    if (tlen == 9) HIPCHK(sfg::launch_a2_to_state(1, n, dt, dx, s));
    else HIPCHK(sfg::launch_a4_to_state(1, n, dt, dx, s));
    HIPCHK(sfg::launch_state_to_nlm(1, n, op->cre, op->cim, op->csgn, op->cneg, dx, dn, s));
    HIPCHK(hipMemcpyAsync(nlm, dn, 2 * ncoef * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}
}  // namespace

int sfg_a2_to_nlm(sfg_ctx* ctx, const double a2[9], double* nlm) { return tensor_to_nlm(ctx, a2, 9, 6, nlm); }

int sfg_a4_to_nlm(sfg_ctx* ctx, const double a4[81], double* nlm) { return tensor_to_nlm(ctx, a4, 81, 15, nlm); }

int sfg_Gamma0_batch(sfg_ctx* ctx, int64_t P, const double* eps_dev, const double* T_dev, double T, double A,
                     double Q, double* G_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!eps_dev || !G_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_gamma0(P, eps_dev, T_dev, T, A, Q, G_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_Gamma0(sfg_ctx* ctx, const double eps[9], double T, double A, double Q, double* G) {
    if (!ctx || !eps || !G) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, 10 * sizeof(double)));
    double* de = (double*)buf.p;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(de, eps, 9 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_gamma0(1, de, nullptr, T, A, Q, de + 9, s));
    HIPCHK(hipMemcpyAsync(G, de + 9, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

int sfg_apply_bounds_batch(sfg_ctx* ctx, int L, int64_t P, double* x_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && !x_dev)) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_bounds(P, op->n, x_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_ev_D2(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], double* D) {
    if (!tau || !D) return fail(SFG_ERR_ARG, "NULL argument");
    if (L < 4) return fail(SFG_ERR_L, "<D> needs L >= 4");
    return single_state_eval(ctx, L, nlm, tau, D, nullptr);
}

int sfg_a2(sfg_ctx* ctx, int L, const double* nlm, double a2[9]) {
    if (!a2) return fail(SFG_ERR_ARG, "NULL argument");
    if (L < 2) return fail(SFG_ERR_L, "a2 needs L >= 2");
    return single_state_eval(ctx, L, nlm, nullptr, nullptr, a2);
}

int sfg_M_DDRX(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], double* M) {
    if (!nlm) return fail(SFG_ERR_ARG, "nlm is NULL");
    int rc = dense_single(ctx, L, 1, tau, nullptr, 0, 0, M);
    if (rc) return rc;
    // M_DDRX = M_DDRX_src - <D> I (src/dynamics.f90:266-273)
    double D = 0;
    rc = sfg_ev_D2(ctx, L, nlm, tau, &D);
    if (rc) return rc;
    int n = sfg::nlm_len(L);
    for (int i = 0; i < n; i++) M[((size_t)i * n + i) * 2] -= D;
    return SFG_OK;
}

static int diag_single(sfg_ctx* ctx, int L, int mode, const double* D, double* M) {
    if (!ctx || !M || (mode == 1 && !D)) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    if (mode == 1 && !op->reg_defined) return fail(SFG_ERR_L, "M_REG is calibrated only for L = 4, 6, ..., 20");
    HIPCHK(hipSetDevice(ctx->device));
    int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, ((size_t)n * n + 9) * sizeof(double)));
    double* dM = (double*)buf.p;
    double* dD = dM + (size_t)n * n;
    hipStream_t s = ctx->stream;
    if (D) HIPCHK(hipMemcpyAsync(dD, D, 9 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_diag_op(1, n, mode, op->cl, op->reg_shape, op->reg_nu, dD, dM, s));
    HIPCHK(hipMemcpyAsync(M, dM, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

int sfg_M_CDRX(sfg_ctx* ctx, int L, double* M) { return diag_single(ctx, L, 0, nullptr, M); }

int sfg_M_REG_batch(sfg_ctx* ctx, int L, int64_t P, const double* D_dev, double* M_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!D_dev || !M_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    if (!op->reg_defined) return fail(SFG_ERR_L, "M_REG is calibrated only for L = 4, 6, ..., 20");
    if (P > (int64_t(1) << 40) / ((int64_t)op->n * op->n)) return fail(SFG_ERR_ARG, "P too large for one launch");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_diag_op(P, op->n, 1, op->cl, op->reg_shape, op->reg_nu, D_dev, M_dev, pick(ctx, stream)));
    return SFG_OK;
}

// ---------------------------------------------------------------- enhancement factors

static int eij_batch_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int frame, const double* frames_dev,
                            const double Eg[2], double alpha, int n_grain, double* Eij_dev, double* lami_dev,
                            double* ei_dev, int32_t* status_dev, hipStream_t s) {
    if (L < 4) return fail(SFG_ERR_L, "Eij needs the l <= 4 coefficients (L >= 4)");
    if (n_grain != 1 && n_grain != -3 && n_grain != 3) return fail(SFG_ERR_ARG, "n_grain must be 1, -3 or 3");
    if (n_grain == 3 && L < 8)   // src/homogenizations.f90:94-96
        return fail(SFG_ERR_L, "Sachs homogenization with n'=3 requires L >= 8");
    if (frame < 0 || frame > 2 || (frame == 1 && !frames_dev)) return fail(SFG_ERR_ARG, "bad frame selection");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    sfg::EijArgs a{};
    a.P = P; a.nq = op->nq; a.frame = frame; a.n_grain = n_grain;
    a.Ecc = Eg[0]; a.Eca = Eg[1]; a.alpha = alpha;
    a.x = x_dev; a.A2 = op->A2; a.A4 = op->A4; a.frames = frames_dev;
    a.Eij = Eij_dev; a.lami = lami_dev; a.ei = ei_dev; a.status = status_dev;
    if (n_grain != 3 || alpha > 0) {
        // one flag per wavefront of 64 rows for the regularized second pass
        void* ws;
        WsLease lease(ctx, s);
        rc = lease.acquire((size_t)((P + 63) / 64) * sizeof(int), &ws);
        if (rc) return rc;
        a.wave_flag = (int*)ws;
        if (n_grain == 3) a.alpha = 1.0;   // the Taylor values (n' = 3 parameters), mixed in below
        HIPCHK(sfg::launch_eij(a, s));
        rc = lease.release();
        if (rc) return rc;
        a.wave_flag = nullptr;
        a.alpha = alpha;
    }
    if (n_grain == 3) {
        // n' = 3 Sachs (src/homogenizations.f90:93-103) and the mix with the Taylor values
        sfg::Sachs3Args b{};
        b.A4F = op->A4F;
        b.nrows = op->n;
        double m[15][15];
        sfg::a4full_map(m);
        for (int e = 0; e < 15; e++) b.u_iso[e] = m[e][0];   // the isotropic state x = (x0, 0, ...)
        HIPCHK(sfg::launch_eij_sachs3(a, b, s));
    }
    return SFG_OK;
}

int sfg_Eij_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int frame,
                                const double* frames_dev, const double Eij_grain[2], double alpha, int n_grain,
                                double* Eij_dev, double* lami_dev, double* ei_dev, int32_t* status_dev, void* stream) {
    if (!ctx || !Eij_grain || P < 0 || (P > 0 && (!x_dev || !Eij_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    return eij_batch_locked(ctx, L, P, x_dev, frame, frames_dev, Eij_grain, alpha, n_grain, Eij_dev, lami_dev, ei_dev,
                            status_dev, pick(ctx, stream));
}

// Evw of nstates stress states states_dev (S, 15) for P parcels: Evw_dev (P, S)
static int evw_batch_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int nstates,
                            const double* states_dev, const double Eg[2], double alpha, int n_grain, double* Evw_dev,
                            int32_t* status_dev, hipStream_t s) {
    if (L < 4) return fail(SFG_ERR_L, "Evw needs the l <= 4 coefficients (L >= 4)");
    if (n_grain != 1 && n_grain != -3 && n_grain != 3) return fail(SFG_ERR_ARG, "n_grain must be 1, -3 or 3");
    if (n_grain == 3 && L < 8)   // src/homogenizations.f90:94-96
        return fail(SFG_ERR_L, "Sachs homogenization with n'=3 requires L >= 8");
    if (nstates < 0 || (nstates > 0 && !states_dev)) return fail(SFG_ERR_ARG, "bad stress states");
    if (nstates == 0) return SFG_OK;
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    sfg::EijArgs a{};
    a.P = P; a.nq = op->nq; a.frame = 2; a.n_grain = n_grain;
    a.Ecc = Eg[0]; a.Eca = Eg[1]; a.alpha = n_grain == 3 ? 1.0 : alpha;
    a.x = x_dev; a.A2 = op->A2; a.A4 = op->A4;
    a.Eij = Evw_dev; a.status = status_dev; a.states = states_dev; a.nstates = nstates;
    if (n_grain != 3 || alpha > 0) HIPCHK(sfg::launch_evw(a, s));   // (the Taylor part for n' = 3)
    if (n_grain == 3) {
        a.alpha = alpha;
        sfg::Sachs3Args b{};
        b.A4F = op->A4F;
        b.nrows = op->n;
        double m[15][15];
        sfg::a4full_map(m);
        for (int e = 0; e < 15; e++) b.u_iso[e] = m[e][0];
        HIPCHK(sfg::launch_evw_sachs3(a, b, s));
    }
    return SFG_OK;
}

int sfg_Evw_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int nstates,
                                const double* states_dev, const double Eij_grain[2], double alpha, int n_grain,
                                double* Evw_dev, int32_t* status_dev, void* stream) {
    if (!ctx || !Eij_grain || P < 0 || (P > 0 && nstates > 0 && (!x_dev || !Evw_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    return evw_batch_locked(ctx, L, P, x_dev, nstates, states_dev, Eij_grain, alpha, n_grain, Evw_dev, status_dev,
                            pick(ctx, stream));
}

// (v, w, tau) of one stress state as the 15 doubles the kernels read; tau must be symmetric
static int pack_state(const double v[3], const double w[3], const double tau[9], double st[15]) {
    if (!v || !w || !tau) return fail(SFG_ERR_ARG, "NULL argument");
    double nrm = 0, asym = 0;
    for (int i = 0; i < 9; i++) nrm = std::max(nrm, std::fabs(tau[i]));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) asym = std::max(asym, std::fabs(tau[3 * i + j] - tau[3 * j + i]));
    if (asym > 1e-12 * nrm) return fail(SFG_ERR_ARG, "tau must be symmetric (a stress)");
    for (int i = 0; i < 3; i++) st[i] = v[i], st[3 + i] = w[i];
    for (int i = 0; i < 9; i++) st[6 + i] = tau[i];
    return SFG_OK;
}

int sfg_Evw_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, const double v[3], const double w[3],
                          const double tau[9], const double Eij_grain[2], double alpha, int n_grain, double* Evw) {
    if (!ctx || !nlm || !Eij_grain || !Evw) return fail(SFG_ERR_ARG, "NULL argument");
    double st[15];
    int rc = pack_state(v, w, tau, st);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(3 * n + 16) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* dst = dx + n;
    double* dout = dst + 15;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dst, st, 15 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    rc = evw_batch_locked(ctx, L, 1, dx, 1, dst, Eij_grain, alpha, n_grain, dout, nullptr, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(Evw, dout, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

// single parcel: host nlm -> device state -> kernel -> host outputs
static int eij_single(sfg_ctx* ctx, int L, const double* nlm, int frame, const double* frame9, const double Eg[2],
                      double alpha, int ng, double* Eij, double* ei, double* lami) {
    if (!ctx || !nlm) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(2 * n + n + 9 + 6 + 3 + 9 + 1) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* df = dx + n;
    double* dE = df + 9;
    double* dl = dE + 6;
    double* de = dl + 3;
    int32_t* dst = (int32_t*)(de + 9);
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    if (frame9) HIPCHK(hipMemcpyAsync(df, frame9, 9 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(dst, 0, sizeof(int32_t), s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    double Eg0[2] = {Eg ? Eg[0] : 1.0, Eg ? Eg[1] : 1.0};
    rc = eij_batch_locked(ctx, L, 1, dx, frame, df, Eg0, alpha, ng, dE, dl, de, dst, s);
    if (rc) return rc;
    if (Eij) HIPCHK(hipMemcpyAsync(Eij, dE, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ei) HIPCHK(hipMemcpyAsync(ei, de, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
    if (lami) HIPCHK(hipMemcpyAsync(lami, dl, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

int sfg_eig(sfg_ctx* ctx, int L, const double* nlm, double ei[9], double lami[3]) {
    if (!ei || !lami) return fail(SFG_ERR_ARG, "NULL argument");
    return eij_single(ctx, L, nlm, 0, nullptr, nullptr, 0.0, 1, nullptr, ei, lami);
}

int sfg_moments68_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* m6_dev, double* m8_dev,
                        void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !m6_dev || !m8_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (L < 6) return fail(SFG_ERR_L, "sixth moments need the l <= 6 coefficients (L >= 6)");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_moments68(P, std::min(op->n, 45), x_dev, m6_dev, m8_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_eig3(sfg_ctx* ctx, const double M[9], double ei[9], double lami[3]) {
    if (!ctx || !M || !ei || !lami) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, 21 * sizeof(double)));
    double* d = (double*)buf.p;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(d, M, 9 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_eig3(1, d, d + 9, d + 18, s));
    HIPCHK(hipMemcpyAsync(ei, d + 9, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(lami, d + 18, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

int sfg_eig3_batch(sfg_ctx* ctx, int64_t P, const double* M_dev, double* ei_dev, double* lami_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!M_dev || !ei_dev || !lami_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(sfg::launch_eig3(P, M_dev, ei_dev, lami_dev, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_Eij_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, const double e1[3], const double e2[3],
                          const double e3[3], const double Eij_grain[2], double alpha, int n_grain, double Eij[6]) {
    if (!e1 || !e2 || !e3 || !Eij_grain || !Eij) return fail(SFG_ERR_ARG, "NULL argument");
    double f[9] = {e1[0], e1[1], e1[2], e2[0], e2[1], e2[2], e3[0], e3[1], e3[2]};
    return eij_single(ctx, L, nlm, 1, f, Eij_grain, alpha, n_grain, Eij, nullptr, nullptr);
}

int sfg_M_REG(sfg_ctx* ctx, int L, const double D[9], double* M) { return diag_single(ctx, L, 1, D, M); }

// ---------------------------------------------------------------- deformability, CAFFE

namespace {
// <D> of P states with per-parcel tau_dev (P, 3, 3): n_RSS = 2 (ev_D2) or 4 (ev_D4)
int ev_D_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* tau_dev, int n_rss,
                double* D_dev, hipStream_t s) {
    if (n_rss != 2 && n_rss != 4) return fail(SFG_ERR_ARG, "n_RSS must be 2 or 4 (src/dynamics.f90:385-400)");
    if (L < 4) return fail(SFG_ERR_L, "<D> needs the l <= 4 coefficients (L >= 4)");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    if (n_rss == 2) HIPCHK(sfg::launch_evD2(P, op->nq, op->A2, op->A4, x_dev, tau_dev, D_dev, s));
    else HIPCHK(sfg::launch_ev_D4(P, op->n, op->A4F, x_dev, tau_dev, D_dev, s));
    return SFG_OK;
}

// one host nlm and host tau through the batched path; caffe: E_CAFFE of <D>
int ev_D_single(sfg_ctx* ctx, int L, const double* nlm, const double* tau, int n_rss, bool caffe, double Emin,
                double Emax, double* out) {
    if (!ctx || !nlm || !tau || !out) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(3 * n + 9 + 1) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* dt = dx + n;
    double* dD = dt + 9;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dt, tau, 9 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    rc = ev_D_locked(ctx, L, 1, dx, dt, n_rss, dD, s);
    if (rc) return rc;
    if (caffe) HIPCHK(sfg::launch_caffe(1, n_rss, Emin, Emax, dD, s));
    HIPCHK(hipMemcpyAsync(out, dD, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}
}  // namespace

int sfg_ev_D(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], int n_RSS, double* D) {
    return ev_D_single(ctx, L, nlm, tau, n_RSS, false, 0.0, 0.0, D);
}

int sfg_ev_D_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* tau_dev, int n_RSS,
                   double* D_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !tau_dev || !D_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    return ev_D_locked(ctx, L, P, x_dev, tau_dev, n_RSS, D_dev, pick(ctx, stream));
}

int sfg_E_CAFFE(sfg_ctx* ctx, int L, const double* nlm, const double eps[9], double Emin, double Emax, int n_grain,
                double* E) {
    return ev_D_single(ctx, L, nlm, eps, n_grain == 3 ? 4 : 2, true, Emin, Emax, E);
}

int sfg_E_CAFFE_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* eps_dev, double Emin,
                      double Emax, int n_grain, double* E_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !eps_dev || !E_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    hipStream_t s = pick(ctx, stream);
    const int n_rss = n_grain == 3 ? 4 : 2;   // src/enhancementfactors.f90:317-321
    int rc = ev_D_locked(ctx, L, P, x_dev, eps_dev, n_rss, E_dev, s);
    if (rc) return rc;
    HIPCHK(sfg::launch_caffe(P, n_rss, Emin, Emax, E_dev, s));
    return SFG_OK;
}

// ---------------------------------------------------------------- full a4, wave velocities

int sfg_Cij_to_Lame_tranisotropic(const double Cij[5], double Lame[5]) {
    if (!Cij || !Lame) return fail(SFG_ERR_ARG, "NULL argument");
    // src/elasticities.f90:107-120
    const double C66 = (Cij[0] - Cij[3]) / 2;
    Lame[0] = Cij[3];
    Lame[1] = C66;
    Lame[2] = Cij[4] / Cij[3];
    Lame[3] = Cij[2] / C66;
    Lame[4] = Cij[1] / Cij[0];
    return SFG_OK;
}

int sfg_Cij_to_Lame_orthotropic(const double Cij[9], double Lame[9]) {
    if (!Cij || !Lame) return fail(SFG_ERR_ARG, "NULL argument");
    // src/elasticities.f90:122-139
    Lame[6] = Cij[4] + Cij[5] - Cij[3];
    Lame[7] = Cij[3] + Cij[5] - Cij[4];
    Lame[8] = Cij[3] + Cij[4] - Cij[5];
    Lame[0] = Cij[0] - 2 * Lame[6];
    Lame[1] = Cij[1] - 2 * Lame[7];
    Lame[2] = Cij[2] - 2 * Lame[8];
    Lame[3] = Cij[6];
    Lame[4] = Cij[7];
    Lame[5] = Cij[8];
    return SFG_OK;
}

static int a4_batch_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a4_dev, hipStream_t s) {
    if (L < 4) return fail(SFG_ERR_L, "a4 needs the l <= 4 coefficients (L >= 4)");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    sfg::A4FullArgs a{};
    a.P = P; a.nq = op->nq; a.x = x_dev; a.A4F = op->A4F; a.a4 = a4_dev;
    HIPCHK(sfg::launch_a4full(a, s));
    return SFG_OK;
}

int sfg_a4_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a4_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !a4_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return a4_batch_locked(ctx, L, P, x_dev, a4_dev, pick(ctx, stream));
}

static int vi_el_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double alpha, const double Lame[5],
                        double rho, int ndir, const double* theta, const double* phi, double* Vi_dev,
                        int32_t* status_dev, hipStream_t s, double* Qn_dev = nullptr) {
    if (L < 4) return fail(SFG_ERR_L, "wave velocities need the l <= 4 coefficients (L >= 4)");
    if (ndir < 0 || (ndir > 0 && (!theta || !phi)) || !Lame) return fail(SFG_ERR_ARG, "bad direction arguments");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    if (P == 0 || ndir == 0) return SFG_OK;
    sfg::ViElasticArgs a{};
    rc = get_dirs(ctx, ndir, theta, phi, &a.R);
    if (rc) return rc;
    a.P = P; a.ndir = ndir; a.nq = op->nq;
    a.x = x_dev; a.A2 = op->A2; a.A4 = op->A4; a.A4F = op->A4F;
    a.alpha = alpha; a.rho = rho;
    sfg::elas_params(Lame, a.kf, a.kr);
    a.Vi = Vi_dev; a.status = status_dev; a.Qn = Qn_dev;
    HIPCHK(sfg::launch_vi_elastic(a, s));
    return SFG_OK;
}

static int vi_em_locked(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double eps_m, double eps_t, double mu,
                        int ndir, const double* theta, const double* phi, double* Vi_dev, hipStream_t s) {
    if (L < 2) return fail(SFG_ERR_L, "EM velocities need the l <= 2 coefficients (L >= 2)");
    if (ndir < 0 || (ndir > 0 && (!theta || !phi))) return fail(SFG_ERR_ARG, "bad direction arguments");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    if (P == 0 || ndir == 0) return SFG_OK;
    sfg::ViEMArgs a{};
    rc = get_dirs(ctx, ndir, theta, phi, &a.R);
    if (rc) return rc;
    a.P = P; a.ndir = ndir; a.x = x_dev; a.A2 = op->A2;
    a.eps_m = eps_m; a.eps_t = eps_t; a.mu = mu; a.Vi = Vi_dev;
    HIPCHK(sfg::launch_vi_em(a, s));
    return SFG_OK;
}

int sfg_Vi_elastic_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double alpha,
                                       const double Lame[5], double rho, int ndir, const double* theta,
                                       const double* phi, double* Vi_dev, int32_t* status_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && ndir > 0 && (!x_dev || !Vi_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return vi_el_locked(ctx, L, P, x_dev, alpha, Lame, rho, ndir, theta, phi, Vi_dev, status_dev, pick(ctx, stream));
}

int sfg_Vi_electromagnetic_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double eps_m,
                                               double eps_t, double mu, int ndir, const double* theta,
                                               const double* phi, double* Vi_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && ndir > 0 && (!x_dev || !Vi_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return vi_em_locked(ctx, L, P, x_dev, eps_m, eps_t, mu, ndir, theta, phi, Vi_dev, pick(ctx, stream));
}

// single parcel: host nlm -> device state -> kernel -> host output
namespace {
int single_parcel(sfg_ctx* ctx, int L, const double* nlm, size_t out_doubles, double* out,
                  int (*body)(sfg_ctx*, int, const double*, double*, void*), void* user) {
    if (!ctx || !nlm || !out) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(3 * n + out_doubles + 1) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* dout = dx + n;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    rc = body(ctx, L, dx, dout, user);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, out_doubles * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

struct ViElUser { double alpha, rho; const double* Lame; int ndir; const double *th, *ph; };
struct ViEmUser { double eps_m, eps_t, mu; int ndir; const double *th, *ph; };
}  // namespace

int sfg_apply_bounds(sfg_ctx* ctx, int L, const double* nlm, double* out) {
    if (!out) return fail(SFG_ERR_ARG, "NULL argument");
    const int n = sfg::nlm_len(L);
    return single_parcel(ctx, L, nlm, (size_t)2 * n, out,
                         [](sfg_ctx* c, int L_, const double* dx, double* dout, void*) {
                             const DevOp* o = nullptr;
                             int r = get_op(c, L_, &o);
                             if (r) return r;
                             // dx is the converted input state (scratch of this call)
                             HIPCHK(sfg::launch_bounds(1, o->n, (double*)dx, c->stream));
                             HIPCHK(sfg::launch_state_to_nlm(1, o->n, o->cre, o->cim, o->csgn, o->cneg, dx, dout,
                                                             c->stream));
                             return SFG_OK;
                         },
                         nullptr);
}

int sfg_a4(sfg_ctx* ctx, int L, const double* nlm, double a4[81]) {
    return single_parcel(ctx, L, nlm, 81, a4,
                         [](sfg_ctx* c, int L_, const double* dx, double* dout, void*) {
                             return a4_batch_locked(c, L_, 1, dx, dout, c->stream);
                         },
                         nullptr);
}

int sfg_Vi_elastic_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double alpha, const double Lame[5],
                                 double rho, int ndir, const double* theta, const double* phi, double* Vi) {
    if (ndir <= 0) return ndir == 0 ? SFG_OK : fail(SFG_ERR_ARG, "ndir < 0");
    ViElUser u{alpha, rho, Lame, ndir, theta, phi};
    return single_parcel(ctx, L, nlm, (size_t)3 * ndir, Vi,
                         [](sfg_ctx* c, int L_, const double* dx, double* dout, void* up) {
                             auto* v = (ViElUser*)up;
                             return vi_el_locked(c, L_, 1, dx, v->alpha, v->Lame, v->rho, v->ndir, v->th, v->ph,
                                                 dout, nullptr, c->stream);
                         },
                         &u);
}

// Qnorm_tranisotropic: the acoustic tensor for k = z, i.e. the velocity
// kernel's tensor for the direction (0, 0), whose rotation blocks are the
// identity (Wigner-D at theta = 0 with the REAL(4) rule is exactly I)
int sfg_Qnorm_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double alpha, const double Lame[5],
                            double Qnorm[9]) {
    if (!Lame || !Qnorm) return fail(SFG_ERR_ARG, "NULL argument");
    const double zero = 0.0;
    ViElUser u{alpha, 1.0, Lame, 1, &zero, &zero};
    double out[12];
    int rc = single_parcel(ctx, L, nlm, 12, out,
                           [](sfg_ctx* c, int L_, const double* dx, double* dout, void* up) {
                               auto* v = (ViElUser*)up;
                               return vi_el_locked(c, L_, 1, dx, v->alpha, v->Lame, v->rho, 1, v->th, v->ph, dout,
                                                   nullptr, c->stream, dout + 3);
                           },
                           &u);
    if (rc) return rc;
    for (int i = 0; i < 9; i++) Qnorm[i] = out[3 + i];
    return SFG_OK;
}

int sfg_moments68(sfg_ctx* ctx, int L, const double* nlm, double m6[28], double m8[45]) {
    if (!m6 || !m8) return fail(SFG_ERR_ARG, "NULL argument");
    if (L < 6) return fail(SFG_ERR_L, "sixth moments need the l <= 6 coefficients (L >= 6)");
    double out[73];
    int rc = single_parcel(ctx, L, nlm, 73, out,
                           [](sfg_ctx* c, int L_, const double* dx, double* dout, void*) {
                               const DevOp* op;
                               int r = get_op(c, L_, &op);
                               if (r) return r;
                               HIPCHK(sfg::launch_moments68(1, std::min(op->n, 45), dx, dout, dout + 28, c->stream));
                               return SFG_OK;
                           },
                           nullptr);
    if (rc) return rc;
    for (int i = 0; i < 28; i++) m6[i] = out[i];
    for (int i = 0; i < 45; i++) m8[i] = out[28 + i];
    return SFG_OK;
}

int sfg_Qnorm_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double alpha,
                                  const double Lame[5], double* Qn_dev, int32_t* status_dev, void* stream) {
    if (!ctx || !Lame || P < 0 || (P > 0 && (!x_dev || !Qn_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    hipStream_t s = pick(ctx, stream);
    HIPCHK(hipSetDevice(ctx->device));
    void* ws;
    WsLease lease(ctx, s);
    int rc = lease.acquire((size_t)P * 3 * sizeof(double), &ws);   // the velocities, not returned
    if (rc) return rc;
    const double zero = 0.0;
    rc = vi_el_locked(ctx, L, P, x_dev, alpha, Lame, 1.0, 1, &zero, &zero, (double*)ws, status_dev, s, Qn_dev);
    if (rc) return rc;
    return lease.release();
}

int sfg_Vi_electromagnetic_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double eps_m, double eps_t,
                                         double mu, int ndir, const double* theta, const double* phi, double* Vi) {
    if (ndir <= 0) return ndir == 0 ? SFG_OK : fail(SFG_ERR_ARG, "ndir < 0");
    ViEmUser u{eps_m, eps_t, mu, ndir, theta, phi};
    return single_parcel(ctx, L, nlm, (size_t)2 * ndir, Vi,
                         [](sfg_ctx* c, int L_, const double* dx, double* dout, void* up) {
                             auto* v = (ViEmUser*)up;
                             return vi_em_locked(c, L_, 1, dx, v->eps_m, v->eps_t, v->mu, v->ndir, v->th, v->ph,
                                                 dout, c->stream);
                         },
                         &u);
}

// ---------------------------------------------------------------- orthotropic grains

extern "C++" {
namespace {
// rheo_params_orthotropic (src/rheologies.f90:186-206) -> ci of the Sachs tensor
// (src/homogenizations.f90:301-303)
void ortho_ci(const double Eg[6], int n_grain, double ci[6]) {
    double B[6];
    for (int i = 0; i < 6; i++) B[i] = std::pow(Eg[i], 2.0 / (n_grain + 1));
    const double lam[6] = {-B[0] + B[1] + B[2], B[0] - B[1] + B[2], B[0] + B[1] - B[2], B[3], B[4], B[5]};
    for (int i = 0; i < 3; i++) ci[i] = 4.0 / 3 * lam[i];
    for (int i = 3; i < 6; i++) ci[i] = 2 * lam[i];
}

int eij_ortho_locked(sfg_ctx* ctx, int L, int64_t P, const double* x1, const double* x2, const double* x3, int frame,
                     const double* frames_dev, const double Eg[6], double alpha, int n_grain, double* Eij_dev,
                     hipStream_t s) {
    (void)alpha;   // Sachs only, as the reference (src/enhancementfactors.f90:162)
    if (L < 4) return fail(SFG_ERR_L, "orthotropic Eij needs the l <= 4 coefficients (L >= 4)");
    if (n_grain != 1)   // the reference returns eps = 0 (Eij = 0/0) for other n'
        return fail(SFG_ERR_ARG, "orthotropic Sachs homogenization supports n_grain = 1 only");
    if (frame != 1 && frame != 2) return fail(SFG_ERR_ARG, "frame must be 1 (given frames) or 2 (Cartesian)");
    if (frame == 1 && !frames_dev) return fail(SFG_ERR_ARG, "frames missing");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    sfg::OrthoArgs a{};
    a.P = P; a.nq = op->nq; a.x1 = x1; a.x2 = x2; a.x3 = x3;
    a.frames = frame == 1 ? frames_dev : nullptr;
    ortho_ci(Eg, n_grain, a.ci);
    a.Eij = Eij_dev;
    HIPCHK(sfg::launch_eij_ortho(a, s));
    return SFG_OK;
}

int evw_ortho_locked(sfg_ctx* ctx, int L, int64_t P, const double* x1, const double* x2, const double* x3,
                     int nstates, const double* states_dev, const double Eg[6], double alpha, int n_grain,
                     double* Evw_dev, hipStream_t s) {
    (void)alpha;   // Sachs only, as the reference (src/enhancementfactors.f90:162)
    if (L < 4) return fail(SFG_ERR_L, "orthotropic Evw needs the l <= 4 coefficients (L >= 4)");
    if (n_grain != 1) return fail(SFG_ERR_ARG, "orthotropic Sachs homogenization supports n_grain = 1 only");
    if (nstates < 0 || (nstates > 0 && !states_dev)) return fail(SFG_ERR_ARG, "bad stress states");
    if (nstates == 0) return SFG_OK;
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    sfg::OrthoArgs a{};
    a.P = P; a.nq = op->nq; a.x1 = x1; a.x2 = x2; a.x3 = x3;
    ortho_ci(Eg, n_grain, a.ci);
    a.Eij = Evw_dev; a.states = states_dev; a.nstates = nstates;
    HIPCHK(sfg::launch_evw_ortho(a, s));
    return SFG_OK;
}

int vi_ortho_locked(sfg_ctx* ctx, int L, int64_t P, const double* x1, const double* x2, const double* x3,
                    const double Lame[9], double rho, int ndir, const double* theta, const double* phi, double* Vi_dev,
                    hipStream_t s) {
    if (L < 4) return fail(SFG_ERR_L, "orthotropic Vi needs the l <= 4 coefficients (L >= 4)");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const double* R;
    rc = get_dirs(ctx, ndir, theta, phi, &R);
    if (rc) return rc;
    sfg::OrthoArgs a{};
    a.P = P; a.nq = op->nq; a.ndir = ndir; a.x1 = x1; a.x2 = x2; a.x3 = x3; a.R = R;
    for (int i = 0; i < 6; i++) a.lam[i] = Lame[i];
    for (int i = 0; i < 3; i++) a.mu[i] = Lame[6 + i];
    a.rho = rho;
    a.Vi = Vi_dev;
    HIPCHK(sfg::launch_vi_ortho(a, s));
    return SFG_OK;
}

// three host nlm vectors (n complex each; nlm3 may be zero: v = b x n) as
// device states, then fn on them with P = 1, out (nout doubles) back to the host
template <class Fn>
int ortho_single(sfg_ctx* ctx, int L, const double* n1, const double* n2, const double* n3, size_t nout, double* out,
                 Fn fn) {
    if (!ctx || !n1 || !n2 || !out) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, ((size_t)9 * n + nout + 15) * sizeof(double)));
    double* dn = (double*)buf.p;             // 3 x 2n
    double* dx = dn + 6 * n;                 // 3 x n
    double* dout = dx + 3 * n;
    double* dfr = dout + nout;               // a frame (9) or a stress state (15)
    hipStream_t s = ctx->stream;
    const double* ns[3] = {n1, n2, n3};
    for (int m = 0; m < 3; m++) {
        if (ns[m]) HIPCHK(hipMemcpyAsync(dn + 2 * n * m, ns[m], 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
        else HIPCHK(hipMemsetAsync(dn + 2 * n * m, 0, 2 * n * sizeof(double), s));
        HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn + 2 * n * m, dx + n * m, s));
    }
    rc = fn(dx, dx + n, dx + 2 * n, dout, dfr, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}
}  // namespace
}  // extern "C++"

int sfg_Eij_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                              const double* x3_dev, int frame, const double* frames_dev, const double Eij_grain[6],
                              double alpha, int n_grain, double* Eij_dev, void* stream) {
    if (!ctx || !Eij_grain || P < 0 || (P > 0 && (!x1_dev || !x2_dev || !Eij_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    return eij_ortho_locked(ctx, L, P, x1_dev, x2_dev, x3_dev, frame, frames_dev, Eij_grain, alpha, n_grain, Eij_dev,
                            pick(ctx, stream));
}

int sfg_Eij_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                        const double e1[3], const double e2[3], const double e3[3], const double Eij_grain[6],
                        double alpha, int n_grain, double Eij[6]) {
    if (!e1 || !e2 || !e3 || !Eij_grain || !Eij) return fail(SFG_ERR_ARG, "NULL argument");
    const double f[9] = {e1[0], e1[1], e1[2], e2[0], e2[1], e2[2], e3[0], e3[1], e3[2]};
    return ortho_single(ctx, L, nlm1, nlm2, nlm3, 6, Eij,
                        [&](const double* x1, const double* x2, const double* x3, double* dout, double* dfr,
                            hipStream_t s) {
                            HIPCHK(hipMemcpyAsync(dfr, f, 9 * sizeof(double), hipMemcpyHostToDevice, s));
                            return eij_ortho_locked(ctx, L, 1, x1, x2, x3, 1, dfr, Eij_grain, alpha, n_grain, dout,
                                                    s);
                        });
}

int sfg_Evw_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                              const double* x3_dev, int nstates, const double* states_dev, const double Eij_grain[6],
                              double alpha, int n_grain, double* Evw_dev, void* stream) {
    if (!ctx || !Eij_grain || P < 0 || (P > 0 && nstates > 0 && (!x1_dev || !x2_dev || !Evw_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0) return SFG_OK;
    return evw_ortho_locked(ctx, L, P, x1_dev, x2_dev, x3_dev, nstates, states_dev, Eij_grain, alpha, n_grain,
                            Evw_dev, pick(ctx, stream));
}

int sfg_Evw_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                        const double v[3], const double w[3], const double tau[9], const double Eij_grain[6],
                        double alpha, int n_grain, double* Evw) {
    if (!Eij_grain || !Evw) return fail(SFG_ERR_ARG, "NULL argument");
    double st[15];
    int rc = pack_state(v, w, tau, st);
    if (rc) return rc;
    return ortho_single(ctx, L, nlm1, nlm2, nlm3, 1, Evw,
                        [&](const double* x1, const double* x2, const double* x3, double* dout, double* dfr,
                            hipStream_t s) {
                            HIPCHK(hipMemcpyAsync(dfr, st, 15 * sizeof(double), hipMemcpyHostToDevice, s));
                            return evw_ortho_locked(ctx, L, 1, x1, x2, x3, 1, dfr, Eij_grain, alpha, n_grain, dout,
                                                    s);
                        });
}

int sfg_Vi_elastic_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                                     const double* x3_dev, double alpha, const double Lame[9], double rho, int ndir,
                                     const double* theta, const double* phi, double* Vi_dev, void* stream) {
    (void)alpha;   // Voigt only, as the reference (src/wavepropagation.f90:44)
    if (!ctx || !Lame || P < 0 || ndir < 0 || (P > 0 && ndir > 0 && (!x1_dev || !x2_dev || !Vi_dev || !theta || !phi)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (P == 0 || ndir == 0) return SFG_OK;
    return vi_ortho_locked(ctx, L, P, x1_dev, x2_dev, x3_dev, Lame, rho, ndir, theta, phi, Vi_dev,
                           pick(ctx, stream));
}

int sfg_Vi_elastic_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                               double alpha, const double Lame[9], double rho, int ndir, const double* theta,
                               const double* phi, double* Vi) {
    (void)alpha;
    if (ndir <= 0) return ndir == 0 ? SFG_OK : fail(SFG_ERR_ARG, "ndir < 0");
    if (!Lame || !theta || !phi) return fail(SFG_ERR_ARG, "NULL argument");
    return ortho_single(ctx, L, nlm1, nlm2, nlm3, (size_t)3 * ndir, Vi,
                        [&](const double* x1, const double* x2, const double* x3, double* dout, double*,
                            hipStream_t s) {
                            return vi_ortho_locked(ctx, L, 1, x1, x2, x3, Lame, rho, ndir, theta, phi, dout, s);
                        });
}

// ---------------------------------------------------------------- orthotropic grains, discrete axes

extern "C++" {
namespace {
// grain averages (E, 3, 57) of E ensembles of N grains into the lease's workspace
int ortho_disc_moments(sfg_ctx* ctx, int64_t E, int64_t N, const double* b, const double* n, const double* v,
                       WsLease& lease, size_t extra, double** mom, double** rest, hipStream_t s) {
    if (N < 1) return fail(SFG_ERR_ARG, "an ensemble needs at least one grain");
    if (!b || !n || !v) return fail(SFG_ERR_ARG, "NULL axis array");
    HIPCHK(hipSetDevice(ctx->device));
    // grain chunks of >= 1024 grains per workgroup, >= ~1024 workgroups overall
    int64_t G = std::min<int64_t>((N + 1023) / 1024, (1024 + E - 1) / E);
    G = std::max<int64_t>(G, 1);
    if (E * G > 0x7fffffff) return fail(SFG_ERR_ARG, "too many ensembles");
    sfg::OrthoDiscArgs a{};
    a.E = E; a.N = N; a.G = (int)G; a.b = b; a.n = n; a.v = v;
    void* ws;
    const size_t nm = (size_t)E * 3 * sfg::ORTHO_NMOM;
    int rc = lease.acquire((nm * (G + 1) + extra) * sizeof(double), &ws);
    if (rc) return rc;
    a.part = (double*)ws;
    a.mom = a.part + nm * G;
    HIPCHK(sfg::launch_ortho_moments(a, s));
    *mom = a.mom;
    if (rest) *rest = a.mom + nm;
    return SFG_OK;
}

int eij_ortho_disc_locked(sfg_ctx* ctx, int64_t E, int64_t N, const double* b, const double* n, const double* v,
                          int frame, const double* frames_dev, const double Eg[6], double alpha, int n_grain,
                          double* Eij_dev, hipStream_t s) {
    (void)alpha;   // Sachs only, as the reference (src/enhancementfactors.f90:258-266)
    if (n_grain != 1) return fail(SFG_ERR_ARG, "orthotropic Sachs homogenization supports n_grain = 1 only");
    if (frame != 1 && frame != 2) return fail(SFG_ERR_ARG, "frame must be 1 (given frames) or 2 (Cartesian)");
    if (frame == 1 && !frames_dev) return fail(SFG_ERR_ARG, "frames missing");
    WsLease lease(ctx, s);
    double* mom;
    int rc = ortho_disc_moments(ctx, E, N, b, n, v, lease, 0, &mom, nullptr, s);
    if (rc) return rc;
    sfg::OrthoMomArgs a{};
    a.E = E; a.mom = mom; a.frames = frame == 1 ? frames_dev : nullptr; a.Eij = Eij_dev;
    ortho_ci(Eg, n_grain, a.ci);
    HIPCHK(sfg::launch_eij_ortho_mom(a, s));
    return lease.release();
}

int vi_ortho_disc_locked(sfg_ctx* ctx, int64_t E, int64_t N, const double* b, const double* n, const double* v,
                         const double Lame[9], double rho, int ndir, const double* theta, const double* phi,
                         double* Vi_dev, hipStream_t s) {
    if (ndir > 65535) return fail(SFG_ERR_ARG, "at most 65535 directions per call");
    const double* K;
    int rc = get_kdirs(ctx, ndir, theta, phi, &K);
    if (rc) return rc;
    WsLease lease(ctx, s);
    double* mom;
    rc = ortho_disc_moments(ctx, E, N, b, n, v, lease, 0, &mom, nullptr, s);
    if (rc) return rc;
    sfg::OrthoMomArgs a{};
    a.E = E; a.mom = mom; a.ndir = ndir; a.kdir = K; a.rho = rho; a.Vi = Vi_dev;
    for (int i = 0; i < 6; i++) a.lam[i] = Lame[i];
    for (int i = 0; i < 3; i++) a.mu[i] = Lame[6 + i];
    HIPCHK(sfg::launch_vi_ortho_mom(a, s));
    return lease.release();
}

// host axes (N, 3) each -> one device ensemble, fn on it, nout doubles back
template <class Fn>
int ortho_disc_single(sfg_ctx* ctx, int64_t N, const double* b, const double* n, const double* v, size_t nout,
                      double* out, Fn fn) {
    if (!ctx || !b || !n || !v || !out) return fail(SFG_ERR_ARG, "NULL argument");
    if (N < 1) return fail(SFG_ERR_ARG, "an ensemble needs at least one grain");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, ((size_t)9 * N + nout + 9) * sizeof(double)));
    double* d = (double*)buf.p;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(d, b, 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d + 3 * N, n, 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d + 6 * N, v, 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
    double* dout = d + 9 * N;
    int rc = fn(d, d + 3 * N, d + 6 * N, dout, dout + nout, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}
}  // namespace
}  // extern "C++"

int sfg_Eij_orthotropic_discrete_batch(sfg_ctx* ctx, int64_t E, int64_t N, const double* b_dev, const double* n_dev,
                                       const double* v_dev, int frame, const double* frames_dev,
                                       const double Eij_grain[6], double alpha, int n_grain, double* Eij_dev,
                                       void* stream) {
    if (!ctx || !Eij_grain || E < 0 || (E > 0 && !Eij_dev)) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (E == 0) return SFG_OK;
    return eij_ortho_disc_locked(ctx, E, N, b_dev, n_dev, v_dev, frame, frames_dev, Eij_grain, alpha, n_grain,
                                 Eij_dev, pick(ctx, stream));
}

int sfg_Eij_orthotropic_discrete(sfg_ctx* ctx, int64_t N, const double* bi, const double* ni, const double* vi,
                                 const double e1[3], const double e2[3], const double e3[3],
                                 const double Eij_grain[6], double alpha, int n_grain, double Eij[6]) {
    if (!e1 || !e2 || !e3 || !Eij_grain) return fail(SFG_ERR_ARG, "NULL argument");
    const double f[9] = {e1[0], e1[1], e1[2], e2[0], e2[1], e2[2], e3[0], e3[1], e3[2]};
    return ortho_disc_single(ctx, N, bi, ni, vi, 6, Eij,
                             [&](const double* b, const double* n, const double* v, double* dout, double* dfr,
                                 hipStream_t s) {
                                 HIPCHK(hipMemcpyAsync(dfr, f, 9 * sizeof(double), hipMemcpyHostToDevice, s));
                                 return eij_ortho_disc_locked(ctx, 1, N, b, n, v, 1, dfr, Eij_grain, alpha, n_grain,
                                                              dout, s);
                             });
}

int sfg_Vi_elastic_orthotropic_discrete_batch(sfg_ctx* ctx, int64_t E, int64_t N, const double* b_dev,
                                              const double* n_dev, const double* v_dev, double alpha,
                                              const double Lame[9], double rho, int ndir, const double* theta,
                                              const double* phi, double* Vi_dev, void* stream) {
    (void)alpha;   // Voigt only, as the reference (src/wavepropagation.f90:73)
    if (!ctx || !Lame || E < 0 || ndir < 0 || (E > 0 && ndir > 0 && (!Vi_dev || !theta || !phi)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (E == 0 || ndir == 0) return SFG_OK;
    return vi_ortho_disc_locked(ctx, E, N, b_dev, n_dev, v_dev, Lame, rho, ndir, theta, phi, Vi_dev,
                                pick(ctx, stream));
}

int sfg_Vi_elastic_orthotropic_discrete(sfg_ctx* ctx, int64_t N, const double* bi, const double* ni,
                                        const double* vi, double alpha, const double Lame[9], double rho, int ndir,
                                        const double* theta, const double* phi, double* Vi) {
    (void)alpha;
    if (ndir <= 0) return ndir == 0 ? SFG_OK : fail(SFG_ERR_ARG, "ndir < 0");
    if (!Lame || !theta || !phi) return fail(SFG_ERR_ARG, "NULL argument");
    return ortho_disc_single(ctx, N, bi, ni, vi, (size_t)3 * ndir, Vi,
                             [&](const double* b, const double* n, const double* v, double* dout, double*,
                                 hipStream_t s) {
                                 return vi_ortho_disc_locked(ctx, 1, N, b, n, v, Lame, rho, ndir, theta, phi, dout,
                                                             s);
                             });
}

// ---------------------------------------------------------------- rotation, a4 eigentensors, validity

namespace {

int ensure_wigner(sfg_ctx* ctx) {
    if (ctx->wigner) return SFG_OK;
    std::vector<sfg::WignerTerm> terms;
    std::vector<int> first;
    sfg::wigner_terms(sfg::ROT_LMAX, terms, first);
    if ((int)terms.size() != sfg::wigner_offset(sfg::ROT_LMAX + 2, 0, -(sfg::ROT_LMAX + 2)))
        return fail(SFG_ERR_ARG, "Wigner term table does not match the kernels' layout");
    std::vector<double> c;
    for (auto& t : terms) c.push_back(t.coef);
    HIPCHK(upload(&ctx->wigner, c));
    return SFG_OK;
}

int rotate_locked(sfg_ctx* ctx, int L, int64_t P, const double* x, const double* th, const double* ph, double* out,
                  hipStream_t s) {
    if (L < 0 || L > sfg::ROT_LMAX || (L % 2)) return fail(SFG_ERR_L, "rotate_nlm needs even L <= 12 (src/rotation.f90:37)");
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_wigner(ctx);
    if (rc) return rc;
    if (L == 0) {   // only n_0^0, which rotations leave alone
        if (out != x && P > 0) HIPCHK(hipMemcpyAsync(out, x, (size_t)P * 8, hipMemcpyDeviceToDevice, s));
        return SFG_OK;
    }
    sfg::RotateArgs a{};
    a.P = P; a.L = L; a.x = x; a.out = out; a.theta = th; a.phi = ph; a.coef = ctx->wigner;
    HIPCHK(sfg::launch_rotate(a, s));
    return SFG_OK;
}

}  // namespace

int sfg_rotate_nlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* theta_dev,
                         const double* phi_dev, double* out_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !theta_dev || !phi_dev || !out_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return rotate_locked(ctx, L, P, x_dev, theta_dev, phi_dev, out_dev, pick(ctx, stream));
}

int sfg_rotate_nlm(sfg_ctx* ctx, int L, const double* nlm, double theta, double phi, double* out) {
    if (!ctx || !nlm || !out) return fail(SFG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)(2 * n + 2 * n + 2) * sizeof(double)));
    double* dn = (double*)buf.p;
    double* dx = dn + 2 * n;
    double* dang = dx + n;
    double* dy = dang + 2;
    const double ang[2] = {theta, phi};
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dn, nlm, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dang, ang, 2 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(sfg::launch_nlm_to_state(1, n, op->r2c, op->rpart, dn, dx, s));
    rc = rotate_locked(ctx, L, 1, dx, dang, dang + 1, dy, s);
    if (rc) return rc;
    HIPCHK(sfg::launch_state_to_nlm(1, n, op->cre, op->cim, op->csgn, op->cneg, dy, dn, s));
    HIPCHK(hipMemcpyAsync(out, dn, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

/* ---- discrete grains ---- */

int sfg_ri_LROT_batch(sfg_ctx* ctx, int64_t N, int Nt, const double* ri0_dev, double dt, const double* D_dev,
                      const double* W_dev, double iota, double* ri_dev, void* stream) {
    if (!ctx || N < 0 || Nt < 1 || (N > 0 && (!ri0_dev || !ri_dev || (Nt > 1 && (!D_dev || !W_dev)))))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    sfg::RiLrotArgs a{};
    a.N = N; a.Nt = Nt; a.dt = dt; a.iota = iota; a.ri0 = ri0_dev; a.D = D_dev; a.W = W_dev; a.ri = ri_dev;
    HIPCHK(sfg::launch_ri_lrot(a, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_ri_LROT(sfg_ctx* ctx, int64_t N, int Nt, const double* ri0, double dt, const double* D, const double* W,
                double iota, double* ri) {
    if (!ctx || N < 0 || Nt < 1 || (N > 0 && (!ri0 || !ri || (Nt > 1 && (!D || !W)))))
        return fail(SFG_ERR_ARG, "bad argument");
    if (N == 0) return SFG_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nr = (size_t)N * 3, nd = (size_t)Nt * 9, nt = (size_t)Nt * N * 3;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, (nr + 2 * nd + nt) * sizeof(double)));
    double* d0 = (double*)buf.p;
    double* dD = d0 + nr;
    double* dW = dD + nd;
    double* dri = dW + nd;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(d0, ri0, nr * sizeof(double), hipMemcpyHostToDevice, s));
    if (Nt > 1) {
        HIPCHK(hipMemcpyAsync(dD, D, nd * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dW, W, nd * sizeof(double), hipMemcpyHostToDevice, s));
    }
    sfg::RiLrotArgs a{};
    a.N = N; a.Nt = Nt; a.dt = dt; a.iota = iota; a.ri0 = d0; a.D = dD; a.W = dW; a.ri = dri;
    HIPCHK(sfg::launch_ri_lrot(a, s));
    HIPCHK(hipMemcpyAsync(ri, dri, nt * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

namespace {
// ri_to_nlm of E ensembles into the real state x (n(L), E) on stream s
int ri_to_state_locked(sfg_ctx* ctx, int L, int64_t E, int64_t N, const double* ri, const double* wi, double* x,
                       hipStream_t s) {
    if (L < 2) return fail(SFG_ERR_L, "ri_to_nlm needs L >= 2 (src/moments.f90:111-130)");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    sfg::RiMomentArgs a{};
    a.E = E; a.N = N; a.ri = ri; a.wi = wi; a.n = op->n;
    a.order = L == 2 ? 2 : (L == 4 ? 4 : 6);
    if (a.order == 6) {
        a.out = x;
        HIPCHK(sfg::launch_ri_moments(a, s));
        return SFG_OK;
    }
    // the a2 / a4 tensors, then the reference's tensor-to-harmonics bodies
    void* ws;
    WsLease lease(ctx, s);
    rc = lease.acquire((size_t)E * (a.order == 2 ? 9 : 81) * sizeof(double), &ws);
    if (rc) return rc;
    a.out = (double*)ws;
    HIPCHK(sfg::launch_ri_moments(a, s));
    if (a.order == 2) HIPCHK(sfg::launch_a2_to_state(E, op->n, a.out, x, s));
    else HIPCHK(sfg::launch_a4_to_state(E, op->n, a.out, x, s));
    return lease.release();
}
}  // namespace

int sfg_ri_to_state_batch(sfg_ctx* ctx, int L, int64_t E, int64_t N, const double* ri_dev, const double* wi_dev,
                          double* x_dev, void* stream) {
    if (!ctx || E < 0 || N < 0 || (E > 0 && (!x_dev || (N > 0 && (!ri_dev || !wi_dev)))))
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (E == 0) return SFG_OK;
    return ri_to_state_locked(ctx, L, E, N, ri_dev, wi_dev, x_dev, pick(ctx, stream));
}

int sfg_ri_to_nlm(sfg_ctx* ctx, int L, int64_t N, const double* ri, const double* wi, double* nlm) {
    if (!ctx || N < 0 || !nlm || (N > 0 && (!ri || !wi))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int n = op->n;
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, ((size_t)N * 4 + 3 * n + 1) * sizeof(double)));
    double* dr = (double*)buf.p;
    double* dw = dr + (size_t)N * 3;
    double* dx = dw + N;
    double* dn = dx + n;
    hipStream_t s = ctx->stream;
    if (N > 0) {
        HIPCHK(hipMemcpyAsync(dr, ri, (size_t)N * 3 * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dw, wi, (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
    }
    rc = ri_to_state_locked(ctx, L, 1, N, dr, dw, dx, s);
    if (rc) return rc;
    HIPCHK(sfg::launch_state_to_nlm(1, n, op->cre, op->cim, op->csgn, op->cneg, dx, dn, s));
    HIPCHK(hipMemcpyAsync(nlm, dn, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

namespace {
int a4eig_locked(sfg_ctx* ctx, int L, int64_t P, const double* x, double* Q, double* ev, int32_t* valid,
                 hipStream_t s) {
    if (L < 4) return fail(SFG_ERR_L, "a4 needs the l <= 4 coefficients (L >= 4)");
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    sfg::A4EigArgs a{};
    a.P = P; a.x = x; a.A2 = op->A2; a.A4F = op->A4F; a.Q = Q; a.ev = ev; a.valid = valid;
    HIPCHK(sfg::launch_a4eig(a, s));
    return SFG_OK;
}
}  // namespace

int sfg_a4_eigentensors_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* Q_dev, double* ev_dev,
                              void* stream) {
    if (!ctx || P < 0 || (P > 0 && !x_dev)) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return a4eig_locked(ctx, L, P, x_dev, Q_dev, ev_dev, nullptr, pick(ctx, stream));
}

int sfg_state_isvalid_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int32_t* valid_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !valid_dev))) return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return a4eig_locked(ctx, L, P, x_dev, nullptr, nullptr, valid_dev, pick(ctx, stream));
}

int sfg_a4_eigentensors(sfg_ctx* ctx, int L, const double* nlm, double Q[54], double ev[6]) {
    if (!Q || !ev) return fail(SFG_ERR_ARG, "NULL argument");
    std::vector<double> o(60);
    int rc = single_parcel(ctx, L, nlm, 60, o.data(),
                           [](sfg_ctx* c, int L_, const double* dx, double* dout, void*) {
                               return a4eig_locked(c, L_, 1, dx, dout, dout + 54, nullptr, c->stream);
                           },
                           nullptr);
    if (rc) return rc;
    std::copy(o.begin(), o.begin() + 54, Q);
    std::copy(o.begin() + 54, o.end(), ev);
    return SFG_OK;
}

int sfg_nlm_isvalid(sfg_ctx* ctx, int64_t K, const double* nhat20, const double* nhat40, int32_t* isvalid) {
    if (!ctx || K < 0 || (K > 0 && (!nhat20 || !nhat40 || !isvalid))) return fail(SFG_ERR_ARG, "NULL argument");
    if (K == 0) return SFG_OK;
    // the states (n_0^0 = 1, n_2^0, n_4^0) at L = 4 in the real basis: rows 0, 1, 6
    std::vector<double> hx((size_t)15 * K, 0.0);
    for (int64_t k = 0; k < K; k++) {
        hx[k] = 1.0;
        hx[(size_t)1 * K + k] = nhat20[2 * k];
        hx[(size_t)6 * K + k] = nhat40[2 * k];
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    HIPCHK(hipMalloc(&buf.p, hx.size() * sizeof(double) + (size_t)K * 4));
    double* dx = (double*)buf.p;
    int32_t* dv = (int32_t*)(dx + hx.size());
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dx, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = a4eig_locked(ctx, 4, K, dx, nullptr, nullptr, dv, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(isvalid, dv, (size_t)K * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

// ---------------------------------------------------------------- surrogate network

}  // extern "C"

struct sfg_mlp {
    sfg_ctx* ctx = nullptr;
    sfg::MlpNet net;
    // surrogate workspace (grow-only)
    int64_t wcap = 0;
    int* flags = nullptr;
    int* idx = nullptr;
    int* count = nullptr;
    double *Ep = nullptr, *Vp = nullptr, *errE = nullptr, *errV = nullptr;
    sfg::AuditSums* sums = nullptr;
    std::vector<void*> wbufs;
    // the network's buffers (activations, work list, statistics) are shared by
    // every call on this handle: each call waits for the previous caller's
    // stream through this event and records it on its own stream when done
    hipEvent_t ev = nullptr;
    bool ev_recorded = false;
};

namespace {
struct MlpLease {
    sfg_mlp* m;
    hipStream_t s;
    bool held = false;
    MlpLease(sfg_mlp* net, hipStream_t st) : m(net), s(st) {}
    int acquire() {
        if (m->ev_recorded) HIPCHK(hipStreamWaitEvent(s, m->ev, 0));
        held = true;
        return SFG_OK;
    }
    ~MlpLease() {
        if (held && hipEventRecord(m->ev, s) == hipSuccess) m->ev_recorded = true;
    }
};
}  // namespace

namespace {

int mlp_ws(sfg_mlp* m, int64_t P) {
    if (P <= m->wcap) return SFG_OK;
    HIPCHK(hipDeviceSynchronize());
    for (void* p : m->wbufs) (void)hipFree(p);
    m->wbufs.clear();
    auto A = [&](void** p, size_t bytes) {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (e == hipSuccess) m->wbufs.push_back(*p);
        return e;
    };
    const size_t Pz = (size_t)P;
    HIPCHK(A((void**)&m->flags, Pz * 4));
    HIPCHK(A((void**)&m->idx, Pz * 4));
    HIPCHK(A((void**)&m->count, 64));
    HIPCHK(A((void**)&m->Ep, Pz * 6 * 8));
    HIPCHK(A((void**)&m->Vp, Pz * 3 * 8));
    HIPCHK(A((void**)&m->errE, Pz * 8));
    HIPCHK(A((void**)&m->errV, Pz * 8));
    HIPCHK(A((void**)&m->sums, sizeof(sfg::AuditSums)));
    m->wcap = P;
    return SFG_OK;
}

constexpr int64_t MLP_CHUNK = 1 << 16;   // rows per forward pass of the surrogate

}  // namespace

extern "C" {

int64_t sfg_mlp_nparams(void) { return sfg::MLP_NPARAM; }

int sfg_mlp_create(sfg_ctx* ctx, uint64_t seed, sfg_mlp** net) {
    if (!ctx || !net) return fail(SFG_ERR_ARG, "NULL argument");
    auto* m = new sfg_mlp();
    m->ctx = ctx;
    hipError_t e = m->net.init(ctx->device, seed);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete m;
        return fail(SFG_ERR_HIP, std::string("sfg_mlp_create: ") + hipGetErrorString(e));
    }
    *net = m;
    return SFG_OK;
}

int sfg_mlp_destroy(sfg_mlp* net) {
    if (!net) return SFG_OK;
    (void)hipSetDevice(net->net.device);
    (void)hipDeviceSynchronize();
    for (void* p : net->wbufs) (void)hipFree(p);
    if (net->ev) (void)hipEventDestroy(net->ev);
    delete net;
    return SFG_OK;
}

int sfg_mlp_get_params(sfg_mlp* net, float* params) {
    if (!net || !params) return fail(SFG_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(net->net.device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(params, net->net.w, (size_t)sfg::MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    return SFG_OK;
}

int sfg_mlp_set_params(sfg_mlp* net, const float* params) {
    if (!net || !params) return fail(SFG_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(net->net.device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(net->net.set_params(params));
    net->net.trained = true;
    return SFG_OK;
}

int sfg_mlp_set_normalization(sfg_mlp* net, const float in_mean[36], const float in_std[36], const float out_mean[9],
                              const float out_std[9]) {
    if (!net || !in_mean || !in_std || !out_mean || !out_std) return fail(SFG_ERR_ARG, "NULL argument");
    std::vector<float> nm(2 * (sfg::MLP_IN + sfg::MLP_OUT));
    for (int i = 0; i < sfg::MLP_IN; i++) {
        if (!(in_std[i] > 0)) return fail(SFG_ERR_ARG, "in_std must be positive");
        nm[i] = in_mean[i];
        nm[sfg::MLP_IN + i] = in_std[i];
    }
    for (int i = 0; i < sfg::MLP_OUT; i++) {
        if (!(out_std[i] > 0)) return fail(SFG_ERR_ARG, "out_std must be positive");
        nm[2 * sfg::MLP_IN + i] = out_mean[i];
        nm[2 * sfg::MLP_IN + sfg::MLP_OUT + i] = out_std[i];
    }
    HIPCHK(hipSetDevice(net->net.device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(net->net.norm, nm.data(), nm.size() * 4, hipMemcpyHostToDevice));
    net->net.norm_host = nm;
    return SFG_OK;
}

int sfg_mlp_get_normalization(sfg_mlp* net, float* norm90) {
    if (!net || !norm90) return fail(SFG_ERR_ARG, "NULL argument");
    std::copy(net->net.norm_host.begin(), net->net.norm_host.end(), norm90);
    return SFG_OK;
}

int sfg_mlp_set_envelope(sfg_mlp* net, const double lo[3], const double hi[3]) {
    if (!net || !lo || !hi) return fail(SFG_ERR_ARG, "NULL argument");
    for (int i = 0; i < 3; i++) {
        net->net.env_lo[i] = lo[i];
        net->net.env_hi[i] = hi[i];
    }
    net->net.trained = true;
    return SFG_OK;
}

int sfg_mlp_set_error_model(sfg_mlp* net, const double taylor_rel_err[6]) {
    if (!net || !taylor_rel_err) return fail(SFG_ERR_ARG, "NULL argument");
    for (int i = 0; i < 6; i++)
        if (!(taylor_rel_err[i] >= 0) || !std::isfinite(taylor_rel_err[i]))
            return fail(SFG_ERR_ARG, "taylor_rel_err must be finite and >= 0");
    for (int i = 0; i < 6; i++) net->net.taylor_err[i] = taylor_rel_err[i];
    return SFG_OK;
}

static int mlp_loss(sfg_mlp* net, int64_t B, hipStream_t s, float* loss) {
    std::vector<float> rows((size_t)B);
    HIPCHK(hipMemcpyAsync(rows.data(), net->net.loss_rows, (size_t)B * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double acc = 0;
    for (float r : rows) acc += r;
    *loss = (float)(acc / ((double)B * sfg::MLP_OUT));
    return SFG_OK;
}

int sfg_mlp_train_step(sfg_mlp* net, int64_t B, const float* X_dev, const float* T_dev, float lr, float* loss,
                       void* stream) {
    if (!net || B <= 0 || !X_dev || !T_dev) return fail(SFG_ERR_ARG, "bad argument");
    if (B > INT32_MAX) return fail(SFG_ERR_ARG, "batch too large");
    HIPCHK(hipSetDevice(net->net.device));
    hipStream_t s = (hipStream_t)stream;
    MlpLease lease(net, s);
    if (int rc = lease.acquire()) return rc;
    HIPCHK(net->net.ensure_batch(B));
    HIPCHK(net->net.forward(B, X_dev, s));
    HIPCHK(net->net.backward(B, T_dev, s));
    if (loss) {
        int rc = mlp_loss(net, B, s, loss);
        if (rc) return rc;
    }
    HIPCHK(net->net.adam(lr, s));
    return SFG_OK;
}

int sfg_mlp_gradients(sfg_mlp* net, int64_t B, const float* X_dev, const float* T_dev, float* grad, float* loss) {
    if (!net || B <= 0 || !X_dev || !T_dev || !grad) return fail(SFG_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(net->net.device));
    hipStream_t s = net->ctx->stream;
    MlpLease lease(net, s);
    if (int rc = lease.acquire()) return rc;
    HIPCHK(net->net.ensure_batch(B));
    HIPCHK(net->net.forward(B, X_dev, s));
    HIPCHK(net->net.backward(B, T_dev, s));
    if (loss) {
        int rc = mlp_loss(net, B, s, loss);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(grad, net->net.g, (size_t)sfg::MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SFG_OK;
}

int sfg_mlp_forward(sfg_mlp* net, int64_t B, const float* X_dev, float* Y_dev, void* stream) {
    if (!net || B < 0 || (B > 0 && (!X_dev || !Y_dev))) return fail(SFG_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(net->net.device));
    hipStream_t s = (hipStream_t)stream;
    MlpLease lease(net, s);
    if (int rc = lease.acquire()) return rc;
    HIPCHK(net->net.ensure_batch(std::min<int64_t>(B, MLP_CHUNK)));
    for (int64_t off = 0; off < B; off += MLP_CHUNK) {
        const int64_t n = std::min<int64_t>(MLP_CHUNK, B - off);
        HIPCHK(net->net.forward(n, X_dev + off * sfg::MLP_IN, s));
        HIPCHK(sfg::launch_denormalize_out(n, net->net.yout, net->net.norm, Y_dev + off * sfg::MLP_OUT, s));
    }
    return SFG_OK;
}

int sfg_surrogate_features(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* ugrad_dev,
                           float* feat_dev, void* stream) {
    if (!ctx || P < 0 || (P > 0 && (!x_dev || !ugrad_dev || !feat_dev))) return fail(SFG_ERR_ARG, "bad argument");
    if (L < 4) return fail(SFG_ERR_L, "surrogate features need the l <= 4 coefficients (L >= 4)");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    sfg::SurrogateFeatArgs a{};
    a.P = P; a.x = x_dev; a.ugrad = ugrad_dev; a.A2 = op->A2; a.feat = feat_dev; a.flags = nullptr;
    HIPCHK(sfg::launch_features(a, pick(ctx, stream)));
    return SFG_OK;
}

int sfg_surrogate_targets(int64_t P, const double* Eij_dev, const double* Vi_dev, float* T_dev, void* stream) {
    if (P < 0 || (P > 0 && (!Eij_dev || !Vi_dev || !T_dev))) return fail(SFG_ERR_ARG, "bad argument");
    HIPCHK(sfg::launch_targets(P, Eij_dev, Vi_dev, T_dev, (hipStream_t)stream));
    return SFG_OK;
}

int sfg_surrogate_homogenize(sfg_ctx* ctx, sfg_mlp* net, int L, int64_t P, const double* x_dev,
                             const double* ugrad_dev, const sfg_homog_params* hp, double audit_frac, uint64_t seed,
                             double* Eij_dev, double* Vi_dev, int32_t* status_dev, sfg_surrogate_stats* stats,
                             void* stream) {
    if (!ctx || !net || !hp || P < 0 || (P > 0 && (!x_dev || !ugrad_dev || !Eij_dev || !Vi_dev)))
        return fail(SFG_ERR_ARG, "bad argument");
    if (L < 4) return fail(SFG_ERR_L, "the surrogate needs the l <= 4 coefficients (L >= 4)");
    if (hp->n_grain != 1 && hp->n_grain != -3) return fail(SFG_ERR_ARG, "n_grain must be 1 or -3");
    if (!(audit_frac >= 0 && audit_frac <= 1)) return fail(SFG_ERR_ARG, "audit_frac must be in [0, 1]");
    if (!net->net.trained) return fail(SFG_ERR_ARG, "surrogate network has no trained parameters/envelope");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    const DevOp* op;
    int rc = get_op(ctx, L, &op);
    if (rc) return rc;
    sfg_surrogate_stats st{};
    if (P == 0) {
        if (stats) *stats = st;
        return SFG_OK;
    }
    hipStream_t s = pick(ctx, stream);
    rc = mlp_ws(net, P);
    if (rc) return rc;
    MlpLease lease(net, s);
    if ((rc = lease.acquire())) return rc;
    // 1. features, envelope / audit flags, network and decoding in one kernel
    sfg::SurrogateInferArgs ia{};
    sfg::SurrogateFeatArgs& fa = ia.fa;
    fa.P = P; fa.x = x_dev; fa.ugrad = ugrad_dev; fa.A2 = op->A2; fa.feat = nullptr; fa.flags = net->flags;
    for (int i = 0; i < 3; i++) { fa.env_lo[i] = net->net.env_lo[i]; fa.env_hi[i] = net->net.env_hi[i]; }
    fa.audit_frac = audit_frac; fa.seed = seed;
    for (int i = 0; i < sfg::MLP_NLAYER; i++) { ia.wf[i] = net->net.wf[i]; ia.bias[i] = net->net.bias_pad[i]; }
    ia.norm = net->net.norm; ia.gtab = net->net.gtab; ia.Eij = Eij_dev; ia.Vi = Vi_dev; ia.idx = net->idx; ia.count = net->count;
    {   // Sachs term and error bound (rheo_params_tranisotropic, src/rheologies.f90:123-132, ef = +1)
        sfg::SurrogateMixArgs& mx = ia.mix;
        const double ne = 2.0 / (hp->n_grain + 1), pc = std::pow(hp->Eij_grain[0], ne),
                     pa = std::pow(hp->Eij_grain[1], ne);
        mx.A4 = op->A4;
        mx.sA = (pc - 1) / 2;
        mx.sB = (3 * (pc + 1) - 2) / 2 - 2 * pa;
        mx.sC = pa - 1;
        mx.snorm = hp->n_grain == 1 ? 1 + 2.0 / 15 * mx.sB + 2.0 / 3 * mx.sC : 1.0;
        mx.alpha = hp->alpha;
        mx.tol = hp->eij_tol > 0 ? hp->eij_tol : 1e-3;
        for (int i = 0; i < 6; i++) mx.dT[i] = net->net.taylor_err[i];
    }
    HIPCHK(sfg::launch_reset_count(net->count, s));
    HIPCHK(sfg::launch_surrogate_infer(ia, s));
    // 2. plain homogenization of the flagged parcels: the work list and its
    // length stay on the device (grid-stride kernels read the count), so the
    // sequence is stream-ordered without a host round trip
    sfg::EijArgs ea{};
    ea.P = P; ea.nq = op->nq; ea.frame = 2; ea.n_grain = hp->n_grain;
    ea.Ecc = hp->Eij_grain[0]; ea.Eca = hp->Eij_grain[1]; ea.alpha = hp->alpha;
    ea.x = x_dev; ea.A2 = op->A2; ea.A4 = op->A4; ea.Eij = net->Ep;
    ea.idx = net->idx; ea.nidx = P; ea.nidx_dev = net->count;
    HIPCHK(sfg::launch_eij(ea, s));
    sfg::ViElasticArgs va{};
    const double zero = 0.0;
    rc = get_dirs(ctx, 1, &zero, &zero, &va.R);
    if (rc) return rc;
    va.P = P; va.ndir = 1; va.nq = op->nq; va.x = x_dev; va.A2 = op->A2; va.A4 = op->A4; va.A4F = op->A4F;
    va.alpha = hp->vi_alpha; va.rho = hp->rho;
    sfg::elas_params(hp->Lame, va.kf, va.kr);
    va.Vi = net->Vp; va.idx = net->idx; va.nidx = P; va.nidx_dev = net->count;
    HIPCHK(sfg::launch_vi_elastic(va, s));
    HIPCHK(sfg::launch_merge(net->count, P, net->idx, net->flags, net->Ep, net->Vp, Eij_dev, Vi_dev, status_dev,
                             net->errE, net->errV, s));
    if (!stats) return SFG_OK;
    // 3. audit statistics, reduced on the device; the one synchronisation
    HIPCHK(sfg::launch_audit_stats(net->count, P, net->errE, net->errV, net->sums, s));
    sfg::AuditSums h{};
    HIPCHK(hipMemcpyAsync(&h, net->sums, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    st.n_fallback = h.n_fallback;
    st.n_audit = h.n_audit;
    st.n_bound = h.n_bound;
    st.audit_max_err_Eij = h.max_E;
    st.audit_max_err_Vi = h.max_V;
    if (st.n_audit) {
        st.audit_mean_err_Eij = h.sum_E / (double)st.n_audit;
        st.audit_mean_err_Vi = h.sum_V / (double)st.n_audit;
    }
    *stats = st;
    return SFG_OK;
}

// ---------------------------------------------------------------- uneven batches: graph plans

}  // extern "C"

struct sfg_plan {
    sfg_ctx* ctx = nullptr;
    struct Bucket {
        sfg::EvolveArgs ea{};
        const void* fn = nullptr;
        unsigned grid = 0, block = 64;
        void* ws = nullptr;
    };
    std::vector<Bucket> b;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nchunks = 0;
};

extern "C" {

int sfg_plan_destroy(sfg_plan* plan) {
    if (!plan) return SFG_OK;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipDeviceSynchronize();
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    for (auto& bk : plan->b) if (bk.ws) (void)hipFree(bk.ws);
    delete plan;
    return SFG_OK;
}

int sfg_plan_create(sfg_ctx* ctx, int nbuckets, const sfg_evolve_desc* buckets, int chunk, int nchunks,
                    void* stream, sfg_plan** out) {
    if (!ctx || !out || nbuckets <= 0 || !buckets || chunk <= 0 || nchunks <= 0)
        return fail(SFG_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = pick(ctx, stream);
    auto* plan = new sfg_plan();
    plan->ctx = ctx;
    plan->nchunks = nchunks;
    auto bail = [&](int rc) {
        for (auto& bk : plan->b) if (bk.ws) (void)hipFree(bk.ws);
        delete plan;
        return rc;
    };
    for (int i = 0; i < nbuckets; i++) {
        const sfg_evolve_desc* d = buckets + i;
        if (d->P < 0 || (d->P > 0 && (!d->ugrad_dev || !d->x_dev || !d->steps_dev)))
            return bail(fail(SFG_ERR_ARG, "bucket needs ugrad_dev, x_dev and steps_dev (remaining steps)"));
        if (d->scheme != SFG_EULER && d->scheme != SFG_RK4) return bail(fail(SFG_ERR_ARG, "unknown scheme"));
        if (d->precision != SFG_FP64 && d->precision != SFG_FP32C) return bail(fail(SFG_ERR_ARG, "unknown precision"));
        const DevOp* op;
        int rc = get_op(ctx, d->L, &op);
        if (rc) return bail(rc);
        if ((d->procs & SFG_REG) && !op->reg_defined) return bail(fail(SFG_ERR_L, "M_REG needs L = 4..20"));
        if (d->P == 0) continue;
        sfg_plan::Bucket bk;
        const int64_t P = d->P;
        const int n = op->n;
        // the stepper of the bucket first: its stage buffers size the workspace
        const int prec = d->precision == SFG_FP64 ? 0 : 1;
        const bool fits = (int64_t)n * P * 8 < (int64_t(1) << 31);
        int per_block = 64;   // parcels stepped by one workgroup
        int mask = sfg::BUF_YS | sfg::BUF_XA | sfg::BUF_XB;
        if (fits && sfg::fast_available(d->L, prec) && (prec == 1 || ctx->stepper == 0)) {
            int blk = 64;
            const int v = variant_of(ctx, d->L, prec);
            bk.fn = sfg::fast_kernel(d->L, prec, v, &blk, &per_block);
            bk.block = (unsigned)blk;
            mask = sfg::fast_buffers(d->L, prec, v);
        } else if (prec == 0) {
            bk.fn = sfg::generic_kernel();
            bk.block = sfg::GENERIC_BLOCK;
            per_block = sfg::GENERIC_BLOCK;
        } else {
            return bail(fail(SFG_ERR_ARG, "FP32C stepping in a plan needs a generated stepper for L and n * P * 8 < 2^31"));
        }
        bk.grid = (unsigned)((P + per_block - 1) / per_block);
        const size_t nth = (size_t)sfg::NPAR_K * P, nom = (size_t)15 * P, nsl = (size_t)op->nblk * P;
        const size_t nscr = (size_t)__builtin_popcount(mask) * n * P;
        hipError_t e = hipMalloc(&bk.ws, (nth + nom + nsl + nscr) * sizeof(double));
        if (e != hipSuccess) return bail(fail(SFG_ERR_HIP, hipGetErrorString(e)));
        plan->b.push_back(bk);
        sfg_plan::Bucket& B = plan->b.back();
        double* theta = (double*)B.ws;
        double* omega = theta + nth;
        double* sl = omega + nom;
        double* plane = sl + nsl;
        double* bufs[3] = {nullptr, nullptr, nullptr};
        for (int k = 0; k < 3; k++)
            if (mask & (1 << k)) {
                bufs[k] = plane;
                plane += (size_t)n * P;
            }
        sfg::ParamArgs pa{};
        pa.P = P; pa.nblk = op->nblk; pa.nq = op->nq; pa.procs = d->procs;
        pa.iota = d->iota; pa.zeta = d->zeta;
        reg_law(op, d, pa);
        pa.ugrad = d->ugrad_dev; pa.tau = d->tau_dev; pa.gam0 = d->gam0_dev; pa.lam = d->lam_dev;
        pa.A2 = op->A2; pa.A4 = op->A4;
        pa.theta = theta; pa.omega = omega; pa.sl = sl;
        e = sfg::launch_params(pa, s);
        if (e != hipSuccess) return bail(fail(SFG_ERR_HIP, hipGetErrorString(e)));
        sfg::EvolveArgs& ea = B.ea;
        ea.P = P; ea.n = n; ea.nq = op->nq; ea.scheme = d->scheme; ea.nsteps = chunk;
        ea.use_ddrx = (d->procs & SFG_DDRX) ? 1 : 0;
        ea.steps = d->steps_dev; ea.steps_dec = 1; ea.dt = d->dt;
        ea.rows = op->rows; ea.row_blk = op->row_blk; ea.grp_param = op->grp_param; ea.grp_rng = op->grp_rng;
        ea.tcol = op->tcol; ea.tval = op->tval; ea.tval_f = op->tval_f;
        ea.theta = theta; ea.omega = omega; ea.sl = sl;
        ea.X = d->x_dev; ea.YS = bufs[0]; ea.XA = bufs[1]; ea.XB = bufs[2];
        ea.status = d->status_dev;
        ea.zero = 0;
    }
    // the time loop as one graph: chunk c of every bucket depends on chunk
    // c - 1 of the same bucket only, so the buckets run concurrently.  The
    // runtime spreads independent chains over a few streams and queues a
    // surplus chain behind another one, so the chains are added costliest
    // first (n^2 P: the coupling terms grow about as n^2): the longest chain
    // then starts at once and the cheap ones share streams.
    std::vector<size_t> order(plan->b.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    auto cost = [&](size_t i) { return (double)plan->b[i].ea.n * plan->b[i].ea.n * (double)plan->b[i].ea.P; };
    std::stable_sort(order.begin(), order.end(), [&](size_t u, size_t v) { return cost(u) > cost(v); });
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGraphCreate(&plan->graph, 0);
    std::vector<hipGraphNode_t> prev(plan->b.size(), nullptr);
    for (int c = 0; c < nchunks && e == hipSuccess; c++) {
        for (size_t i : order) {
            if (e != hipSuccess) break;
            sfg_plan::Bucket& B = plan->b[i];
            void* args[] = {&B.ea};
            hipKernelNodeParams kp{};
            kp.func = const_cast<void*>(B.fn);
            kp.gridDim = dim3(B.grid);
            kp.blockDim = dim3(B.block);
            kp.sharedMemBytes = 0;
            kp.kernelParams = args;
            kp.extra = nullptr;
            hipGraphNode_t node;
            e = hipGraphAddKernelNode(&node, plan->graph, prev[i] ? &prev[i] : nullptr, prev[i] ? 1 : 0, &kp);
            prev[i] = node;
        }
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (plan->graph) (void)hipGraphDestroy(plan->graph);
        return bail(fail(SFG_ERR_HIP, std::string("sfg_plan_create: ") + hipGetErrorString(e)));
    }
    *out = plan;
    return SFG_OK;
}

int sfg_plan_run(sfg_plan* plan, void* stream) {
    if (!plan) return fail(SFG_ERR_ARG, "plan is NULL");
    HIPCHK(hipSetDevice(plan->ctx->device));
    HIPCHK(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    if (plan->ctx->sync_check) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return SFG_OK;
}

}  // extern "C"
