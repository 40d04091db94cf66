// C ABI of the MI355X LBM hot path (declarations and conventions: lbm_amd.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_internal.h"
#include "lbm_multi.h"

using namespace lbm;

namespace {

thread_local std::string g_create_error;

struct ProfRec {
    const char* name;
    hipEvent_t a, b;
};

}  // namespace

struct lbm_ctx {
    lbm_config cfg{};
    Dims d{};
    hipStream_t stream = nullptr;
    std::string err;
    size_t esize = 8;  // bytes per stored population

    uint8_t* type = nullptr;
    uint8_t* nbflag = nullptr;
    void* pf[2] = {nullptr, nullptr};  // buffer A/B, f-type populations [27][n]
    void* pg[2] = {nullptr, nullptr};  // buffer A/B, g-type populations
    double* fld[LBM_FIELD_COUNT] = {};
    double* fld_block = nullptr;
    void* qd[3] = {};
    void* qd_block = nullptr;
    unsigned long long* tau_count = nullptr;
    double* diag_part = nullptr;
    double* diag_out = nullptr;
    int diag_npart = 0;
    int64_t bytes = 0;

    // Fused-mode state machine.  `cur` holds the latest populations; they are
    // explicit streamed populations (after Init or an upload) or
    // post-collision ones whose streaming is implicit in the next pull.
    enum Kind { EXPLICIT_F, POST_COLLISION };
    int cur = 0;
    Kind kind = EXPLICIT_F;
    bool q_valid = false;
    bool fields_valid = true;
    bool f_materialized = false;
    bool types_set = false;
    bool needs_tms = false;

    double fref[27] = {}, gref[27] = {}, fref_sum[5] = {};

    // multicomponent state (ns > 0): species populations A/B, species fields
    int ns = 0;
    void* sf[2] = {nullptr, nullptr};
    double* sp_block = nullptr;  // X, rho_a [NS][n] and u_a [3][NS][n]
    chem::Units units{};
    // TMS_BC outflow interface cells (cell, direction) and unsupported boundary kinds
    int* tms_cell = nullptr;
    int* tms_dir = nullptr;
    int tms_count = 0;
    std::string tms_unsupported;

    bool prof = false;
    std::vector<ProfRec> pending;
    std::map<std::string, std::pair<double, int64_t>> prof_acc;
};

namespace {

int fail(lbm_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(ctx, LBM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

template <typename F>
int launch(lbm_ctx* c, const char* name, F&& body)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (c->prof) {
        HIP_TRY(c, hipEventCreate(&a));
        HIP_TRY(c, hipEventCreate(&b));
        HIP_TRY(c, hipEventRecord(a, c->stream));
    }
    body();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, LBM_E_HIP, std::string("launch ") + name + ": " + hipGetErrorString(e));
    if (c->prof) {
        HIP_TRY(c, hipEventRecord(b, c->stream));
        c->pending.push_back({name, a, b});
    }
    return LBM_OK;
}

template <typename P>
SingleView<P> make_view(lbm_ctx* c, int src, int dst)
{
    SingleView<P> v{};
    v.d = c->d;
    v.type = c->type;
    v.nbflag = c->nbflag;
    v.f = static_cast<P*>(c->pf[src]);
    v.g = static_cast<P*>(c->pg[src]);
    v.fo = static_cast<P*>(c->pf[dst]);
    v.go = static_cast<P*>(c->pg[dst]);
    for (int i = 0; i < LBM_FIELD_COUNT; ++i) v.fld[i] = c->fld[i];
    for (int d = 0; d < 3; ++d) v.qd[d] = static_cast<P*>(c->qd[d]);
    const double R = c->cfg.gas_const, gam = c->cfg.gamma;
    v.R = R;
    v.gamma = gam;
    v.cv = R / (gam - 1.0);
    v.cp = v.cv + R;
    v.nu = c->cfg.nu;
    v.prtl = c->cfg.prtl;
    for (int l = 0; l < Q; ++l) {
        v.fref[l] = c->fref[l];
        v.gref[l] = c->gref[l];
    }
    for (int i = 0; i < 5; ++i) v.fref_sum[i] = c->fref_sum[i];
    v.tau_count = c->tau_count;
    return v;
}

MultiView<double> make_mview(lbm_ctx* c, int src, int dst)
{
    MultiView<double> v{};
    v.d = c->d;
    v.type = c->type;
    v.f = static_cast<double*>(c->sf[src]);
    v.fo = static_cast<double*>(c->sf[dst]);
    v.g = static_cast<double*>(c->pg[src]);
    v.go = static_cast<double*>(c->pg[dst]);
    for (int i = 0; i < LBM_FIELD_COUNT; ++i) v.fld[i] = c->fld[i];
    v.hrr = c->fld[LBM_FIELD_HRR];
    const int64_t n = c->d.n;
    v.X = c->sp_block;
    v.rhoa = c->sp_block + (size_t)c->ns * n;
    v.ua = c->sp_block + (size_t)2 * c->ns * n;
    v.u = c->units;
    v.reaction = c->cfg.reaction;
    v.tau_count = c->tau_count;
    return v;
}

int run_mc(lbm_ctx* c, McKernel k, const char* name, int src, int dst)
{
    return launch(c, name, [&] { launch_mc<double>(k, make_mview(c, src, dst), c->stream); });
}

bool multi(const lbm_ctx* c) { return c->ns > 0; }

bool fp32(const lbm_ctx* c) { return c->cfg.precision == LBM_FP32_SHIFTED; }
bool fused(const lbm_ctx* c) { return c->cfg.mode == LBM_MODE_FUSED; }

// Runs `fn.template operator()<P>()` with the context's population type.
#define DISPATCH(c, body)          \
    (fp32(c) ? [&] { using P = float; return body; }() : [&] { using P = double; return body; }())

int run_init(lbm_ctx* c) { return DISPATCH(c, launch(c, "init", [&] { launch_init<P>(make_view<P>(c, 0, 0), c->stream); })); }

int run_collide_explicit(lbm_ctx* c, int src, int dst)
{
    return DISPATCH(c, launch(c, "collide_explicit",
                              [&] { launch_collide_explicit<P>(make_view<P>(c, src, dst), c->stream); }));
}
int run_stream(lbm_ctx* c, int src, int dst)
{
    return DISPATCH(c, launch(c, "stream", [&] { launch_stream<P>(make_view<P>(c, src, dst), c->stream); }));
}
int run_moments(lbm_ctx* c, int src)
{
    return DISPATCH(c, launch(c, "moments", [&] { launch_moments<P>(make_view<P>(c, src, src), c->stream); }));
}
int run_fused_collide(lbm_ctx* c, int src, int dst)
{
    return DISPATCH(c, launch(c, "fused_collide_stream",
                              [&] { launch_fused_collide<P>(make_view<P>(c, src, dst), c->stream); }));
}
int run_fused_moments(lbm_ctx* c, int src, int dst, bool mat)
{
    return DISPATCH(c, launch(c, mat ? "fused_moments_materialize" : "fused_moments",
                              [&] { launch_fused_moments<P>(make_view<P>(c, src, dst), mat, c->stream); }));
}

// Brings the host fields (and, in fused mode, the streamed populations) up to date.
int materialize(lbm_ctx* c)
{
    if (!fused(c) || c->kind == lbm_ctx::EXPLICIT_F || c->fields_valid) return LBM_OK;
    int rc = run_fused_moments(c, c->cur, 1 - c->cur, true);
    if (rc) return rc;
    c->fields_valid = true;
    c->q_valid = true;
    c->f_materialized = true;
    return LBM_OK;
}

// Rest state of the FP32 deviations: f_ref = rho0 prod Psi(c; 0, theta0),
// g_ref = g_eq(rho0, cv T0, theta0, 0).
void set_reference_state(lbm_ctx* c)
{
    for (int l = 0; l < Q; ++l) c->fref[l] = c->gref[l] = 0.0;
    for (double& s : c->fref_sum) s = 0.0;
    if (!fp32(c)) return;
    const double R = c->cfg.gas_const, cv = R / (c->cfg.gamma - 1.0);
    const double rho0 = c->cfg.ref_rho > 0 ? c->cfg.ref_rho : 1.0;
    const double T0 = c->cfg.ref_T > 0 ? c->cfg.ref_T : 1.0;
    const double th = R * T0;
    const double zero[3] = {0.0, 0.0, 0.0};
    const Axis3<double> z = psi_axis<double>(0.0, th);
    const GeqAxes<double> G = geq_axes<double>(th, zero);
    for (int l = 0; l < Q; ++l) {
        c->fref[l] = rho0 * pick(z, cx(l)) * pick(z, cy(l)) * pick(z, cz(l));
        c->gref[l] = geq_dir<double>(G, l, rho0, cv * T0 - 1.5 * th);
        c->fref_sum[0] += c->fref[l];
        c->fref_sum[1] += cx(l) * cx(l) * c->fref[l];
        c->fref_sum[2] += cy(l) * cy(l) * c->fref[l];
        c->fref_sum[3] += cz(l) * cz(l) * c->fref[l];
        c->fref_sum[4] += c->gref[l];
    }
}

void release(lbm_ctx* c)
{
    for (auto& r : c->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->pending.clear();
    void* ptrs[] = {c->type, c->nbflag, c->pf[0], c->pf[1], c->pg[0], c->pg[1], c->fld_block, c->qd_block,
                    c->tau_count, c->diag_part, c->diag_out, c->sf[0], c->sf[1], c->sp_block, c->tms_cell,
                    c->tms_dir};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int check_field(lbm_ctx* c, int field)
{
    if (field < 0 || field >= LBM_FIELD_COUNT) return fail(c, LBM_E_ARG, "field id out of range");
    return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_create_error(void) { return g_create_error.c_str(); }
const char* lbm_last_error(const lbm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lbm_create(const lbm_config* cfg, lbm_ctx** out)
{
    g_create_error.clear();
    if (!cfg || !out) {
        g_create_error = "null argument";
        return LBM_E_ARG;
    }
    *out = nullptr;
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) {
        g_create_error = "grid extents must be >= 1";
        return LBM_E_ARG;
    }
    if (cfg->n_species != 0 && cfg->n_species != mech::NS) {
        g_create_error = "n_species must be 0 (single component) or 9 (the H2-air mechanism)";
        return LBM_E_UNSUPPORTED;
    }
    if (cfg->n_species == mech::NS) {
        if (cfg->mode != LBM_MODE_STRICT || cfg->precision != LBM_FP64) {
            g_create_error = "the multicomponent path runs in LBM_MODE_STRICT with LBM_FP64";
            return LBM_E_UNSUPPORTED;
        }
        if (!(cfg->dx_si > 0.0) || !(cfg->dt_si > 0.0)) {
            g_create_error = "multicomponent state needs dx_si > 0 and dt_si > 0 (units.set_m_kg_s)";
            return LBM_E_ARG;
        }
    }
    if (cfg->precision != LBM_FP64 && cfg->precision != LBM_FP32_SHIFTED) {
        g_create_error = "unknown precision";
        return LBM_E_ARG;
    }
    if (cfg->mode != LBM_MODE_FUSED && cfg->mode != LBM_MODE_STRICT) {
        g_create_error = "unknown mode";
        return LBM_E_ARG;
    }
    if (cfg->n_species == 0 &&
        (!(cfg->gamma > 1.0) || !(cfg->gas_const > 0.0) || !(cfg->prtl > 0.0) || !(cfg->nu >= 0.0))) {
        g_create_error = "gas constants out of range (gamma > 1, R > 0, Pr > 0, nu >= 0)";
        return LBM_E_ARG;
    }
    const int64_t n = (int64_t)(cfg->nx + 2) * (cfg->ny + 2) * (cfg->nz + 2);
    if (n * Q >= ((int64_t)1 << 40) || n >= ((int64_t)1 << 31)) {
        g_create_error = "grid too large";
        return LBM_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "no HIP device visible";
        return LBM_E_HIP;
    }
    if (cfg->device < 0 || cfg->device >= ndev) {
        g_create_error = "device ordinal out of range";
        return LBM_E_ARG;
    }
    lbm_ctx* c = new lbm_ctx();
    c->cfg = *cfg;
    c->d = {cfg->nx + 2, cfg->ny + 2, cfg->nz + 2, n};
    c->esize = fp32(c) ? 4 : 8;
    auto bail = [&](const char* what, hipError_t e) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(e);
        release(c);
        delete c;
        return LBM_E_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(cfg->device)) != hipSuccess) return bail("hipSetDevice", e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail("hipStreamCreate", e);
    const size_t pop_bytes = (size_t)n * Q * c->esize;
    for (int b = 0; b < 2; ++b) {
        if (cfg->n_species == 0) {  // the mixture f exists only without species (lbm.hpp:62-64)
            if ((e = hipMalloc(&c->pf[b], pop_bytes)) != hipSuccess) return bail("hipMalloc populations", e);
            (void)hipMemsetAsync(c->pf[b], 0, pop_bytes, c->stream);
            c->bytes += pop_bytes;
        }
        if ((e = hipMalloc(&c->pg[b], pop_bytes)) != hipSuccess) return bail("hipMalloc populations", e);
        (void)hipMemsetAsync(c->pg[b], 0, pop_bytes, c->stream);
    }
    if ((e = hipMalloc(&c->type, n)) != hipSuccess) return bail("hipMalloc types", e);
    if ((e = hipMalloc(&c->nbflag, n)) != hipSuccess) return bail("hipMalloc types", e);
    (void)hipMemsetAsync(c->type, TYPE_F, n, c->stream);
    (void)hipMemsetAsync(c->nbflag, 0, n, c->stream);
    if ((e = hipMalloc(&c->fld_block, sizeof(double) * n * LBM_FIELD_COUNT)) != hipSuccess)
        return bail("hipMalloc fields", e);
    for (int i = 0; i < LBM_FIELD_COUNT; ++i) c->fld[i] = c->fld_block + (size_t)i * n;
    (void)hipMemsetAsync(c->fld_block, 0, sizeof(double) * n * LBM_FIELD_COUNT, c->stream);
    if ((e = hipMalloc(&c->qd_block, c->esize * n * 3)) != hipSuccess) return bail("hipMalloc fluxes", e);
    for (int d = 0; d < 3; ++d) c->qd[d] = static_cast<char*>(c->qd_block) + (size_t)d * n * c->esize;
    (void)hipMemsetAsync(c->qd_block, 0, c->esize * n * 3, c->stream);
    if ((e = hipMalloc(&c->tau_count, sizeof(unsigned long long))) != hipSuccess) return bail("hipMalloc", e);
    (void)hipMemsetAsync(c->tau_count, 0, sizeof(unsigned long long), c->stream);
    if (cfg->n_species > 0) {
        c->ns = cfg->n_species;
        c->units = chem::make_units(cfg->dx_si, cfg->dt_si);
        const size_t sbytes = (size_t)n * Q * c->ns * sizeof(double);
        for (int b = 0; b < 2; ++b) {
            if ((e = hipMalloc(&c->sf[b], sbytes)) != hipSuccess) return bail("hipMalloc species populations", e);
            (void)hipMemsetAsync(c->sf[b], 0, sbytes, c->stream);
        }
        const size_t fb = sizeof(double) * (size_t)n * c->ns * 5;
        if ((e = hipMalloc((void**)&c->sp_block, fb)) != hipSuccess) return bail("hipMalloc species fields", e);
        (void)hipMemsetAsync(c->sp_block, 0, fb, c->stream);
        c->bytes += 2 * sbytes + fb;
    }
    c->diag_npart = diag_partials(c->d);
    if ((e = hipMalloc(&c->diag_part, sizeof(double) * 8 * c->diag_npart)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMalloc(&c->diag_out, sizeof(double) * 8)) != hipSuccess) return bail("hipMalloc", e);
    c->bytes += 2 * pop_bytes + 2 * n + sizeof(double) * n * LBM_FIELD_COUNT + c->esize * n * 3 +
               sizeof(double) * (8 * c->diag_npart + 9);
    // MIXTURE defaults (lbm.hpp:70-83): rho = 1, p = T = 1/3.
    std::vector<double> tmp(n, 1.0);
    (void)hipMemcpyAsync(c->fld[LBM_FIELD_RHO], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    std::fill(tmp.begin(), tmp.end(), 1.0 / 3.0);
    (void)hipMemcpyAsync(c->fld[LBM_FIELD_P], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(c->fld[LBM_FIELD_T], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail("initial upload", e);
    set_reference_state(c);
    *out = c;
    return LBM_OK;
}

void lbm_destroy(lbm_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->cfg.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
}

int lbm_set_types(lbm_ctx* c, const uint8_t* types)
{
    if (!c || !types) return fail(c, LBM_E_ARG, "null argument");
    const Dims& d = c->d;
    for (int64_t i = 0; i < d.n; ++i)
        if (types[i] > TYPE_Q) return fail(c, LBM_E_ARG, "cell type out of range 0..9");
    std::vector<uint8_t> flag(d.n, 0);
    bool needs_tms = false, fused_ok = true;
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) {
                const int64_t cidx = ((int64_t)i * d.ny + j) * d.nz + k;
                if (types[cidx] != TYPE_F) continue;
                if (i == 0 || j == 0 || k == 0 || i == d.nx - 1 || j == d.ny - 1 || k == d.nz - 1)
                    return fail(c, LBM_E_ARG, "ghost-layer cell marked TYPE_F");
                bool simple = true, keep = false;
                for (int l = 1; l < Q; ++l) {
                    const uint8_t t = types[((int64_t)(i - cx(l)) * d.ny + (j - cy(l))) * d.nz + (k - cz(l))];
                    if (t != TYPE_F && t != TYPE_P) simple = false;
                    if (t == TYPE_S || t == TYPE_O || t == TYPE_O_C || t == TYPE_I || t == TYPE_I_C) keep = true;
                    if (t == TYPE_Q) fused_ok = false;
                }
                if (keep) needs_tms = true;
                flag[cidx] = (simple ? NB_SIMPLE : 0) | (keep ? NB_KEEP : 0);
            }
    std::vector<int> tcell, tdir;
    std::string tms_unsupported;
    if (multi(c)) {
        for (int64_t i = 0; i < d.n; ++i)
            if (types[i] == TYPE_FS)
                return fail(c, LBM_E_UNSUPPORTED, "free-slip walls are not provided for the multicomponent state");
        // the outflow sweep's interface detection (lbm-bc.cpp:141-155): the first
        // direction whose upstream cell is an outflow ghost and whose downstream
        // cell is fluid; only axis-aligned interfaces are reconstructed
        auto at = [&](int i, int j, int k) { return types[((int64_t)i * d.ny + j) * d.nz + k]; };
        for (int i = 1; i < d.nx - 1; ++i)
            for (int j = 1; j < d.ny - 1; ++j)
                for (int k = 1; k < d.nz - 1; ++k) {
                    if (at(i, j, k) != TYPE_F) continue;
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = at(i - cx(l), j - cy(l), k - cz(l));
                        if (tn == TYPE_I || tn == TYPE_I_C || tn == TYPE_S)
                            tms_unsupported = "inflow (TYPE_I, TYPE_I_C) and isothermal-wall (TYPE_S) TMS "
                                              "reconstruction is not provided for the multicomponent state";
                    }
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = at(i - cx(l), j - cy(l), k - cz(l));
                        const uint8_t tp = at(i + cx(l), j + cy(l), k + cz(l));
                        if ((tn == TYPE_O || tn == TYPE_O_C) && tp == TYPE_F) {
                            if (c2(l) == 1) {
                                if (tn == TYPE_O)
                                    tms_unsupported = "TYPE_O (zero-gradient) outflow is not provided; use TYPE_O_C";
                                tcell.push_back((int)(((int64_t)i * d.ny + j) * d.nz + k));
                                tdir.push_back(l);
                            }
                            break;
                        }
                    }
                }
    }
    if (fused(c) && (needs_tms || !fused_ok))
        return fail(c, LBM_E_UNSUPPORTED,
                    "fused mode handles periodic, adiabatic (TYPE_A) and free-slip (TYPE_FS) neighbours; "
                    "TYPE_S/I/O/I_C/O_C/Q neighbours need LBM_MODE_STRICT");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipMemcpyAsync(c->type, types, d.n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->nbflag, flag.data(), d.n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (multi(c)) {
        if (c->tms_cell) (void)hipFree(c->tms_cell);
        if (c->tms_dir) (void)hipFree(c->tms_dir);
        c->tms_cell = c->tms_dir = nullptr;
        c->tms_count = (int)tcell.size();
        if (c->tms_count) {
            HIP_TRY(c, hipMalloc((void**)&c->tms_cell, sizeof(int) * tcell.size()));
            HIP_TRY(c, hipMalloc((void**)&c->tms_dir, sizeof(int) * tdir.size()));
            HIP_TRY(c, hipMemcpy(c->tms_cell, tcell.data(), sizeof(int) * tcell.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->tms_dir, tdir.data(), sizeof(int) * tdir.size(), hipMemcpyHostToDevice));
        }
        c->tms_unsupported = tms_unsupported;
        needs_tms = false;
    }
    c->needs_tms = needs_tms;
    c->types_set = true;
    return LBM_OK;
}

int lbm_set_field(lbm_ctx* c, int field, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = check_field(c, field)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (int rc = materialize(c)) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->fld[field], host, sizeof(double) * c->d.n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_field(lbm_ctx* c, int field, double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = check_field(c, field)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (int rc = materialize(c)) return rc;
    HIP_TRY(c, hipMemcpyAsync(host, c->fld[field], sizeof(double) * c->d.n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

namespace {

// Which device buffer backs a population set in the current state.
int pop_slot(lbm_ctx* c, int which, bool for_write, void** buf, const double** ref)
{
    const bool gset = (which == LBM_POP_G || which == LBM_POP_GPC);
    const bool post = (which == LBM_POP_FPC || which == LBM_POP_GPC);
    int slot;
    if (!fused(c)) {
        slot = post ? 1 : 0;
    } else if (for_write) {
        slot = post ? 1 : 0;
    } else if (post) {
        if (c->kind != lbm_ctx::POST_COLLISION)
            return fail(c, LBM_E_STATE, "no post-collision populations yet (call lbm_collide first)");
        slot = c->cur;
    } else if (c->kind == lbm_ctx::EXPLICIT_F) {
        slot = c->cur;
    } else {
        if (int rc = materialize(c)) return rc;
        slot = 1 - c->cur;
    }
    *buf = gset ? c->pg[slot] : c->pf[slot];
    *ref = gset ? c->gref : c->fref;
    return slot;
}

}  // namespace

int lbm_set_populations(lbm_ctx* c, int which, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (which < 0 || which > 3) return fail(c, LBM_E_ARG, "population set out of range");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (multi(c)) {
        const int slot = (which == LBM_POP_F || which == LBM_POP_G) ? 0 : 1;
        const bool sp = (which == LBM_POP_F || which == LBM_POP_FPC);
        const size_t bytes = (size_t)c->d.n * Q * (sp ? c->ns : 1) * sizeof(double);
        HIP_TRY(c, hipMemcpyAsync(sp ? c->sf[slot] : c->pg[slot], host, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return LBM_OK;
    }
    if (fused(c)) {
        // keep the other member of the f/g pair in the same buffer
        if (c->kind == lbm_ctx::POST_COLLISION && (which == LBM_POP_F || which == LBM_POP_G)) {
            if (int rc = materialize(c)) return rc;
            const size_t bytes = c->d.n * Q * c->esize;
            const int src = 1 - c->cur;
            HIP_TRY(c, hipMemcpyAsync(c->pf[0], c->pf[src], bytes, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->pg[0], c->pg[src], bytes, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    void* buf;
    const double* ref;
    const int slot = pop_slot(c, which, true, &buf, &ref);
    if (slot < 0) return slot;
    const int64_t n = c->d.n;
    if (fp32(c)) {
        std::vector<float> tmp((size_t)n * Q);
        for (int l = 0; l < Q; ++l)
            for (int64_t i = 0; i < n; ++i) tmp[l * n + i] = (float)(host[l * n + i] - ref[l]);
        HIP_TRY(c, hipMemcpyAsync(buf, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(buf, host, (size_t)n * Q * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    if (fused(c)) {
        c->cur = slot;
        c->kind = (which == LBM_POP_F || which == LBM_POP_G) ? lbm_ctx::EXPLICIT_F : lbm_ctx::POST_COLLISION;
        c->q_valid = false;
        c->f_materialized = false;
        if (c->kind == lbm_ctx::POST_COLLISION) c->fields_valid = false;
    }
    return LBM_OK;
}

int lbm_get_populations(lbm_ctx* c, int which, double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (which < 0 || which > 3) return fail(c, LBM_E_ARG, "population set out of range");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (multi(c)) {
        const int slot = (which == LBM_POP_F || which == LBM_POP_G) ? 0 : 1;
        const bool sp = (which == LBM_POP_F || which == LBM_POP_FPC);
        const size_t bytes = (size_t)c->d.n * Q * (sp ? c->ns : 1) * sizeof(double);
        HIP_TRY(c, hipMemcpyAsync(host, sp ? c->sf[slot] : c->pg[slot], bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return LBM_OK;
    }
    void* buf;
    const double* ref;
    const int slot = pop_slot(c, which, false, &buf, &ref);
    if (slot < 0) return slot;
    const int64_t n = c->d.n;
    if (fp32(c)) {
        std::vector<float> tmp((size_t)n * Q);
        HIP_TRY(c, hipMemcpyAsync(tmp.data(), buf, tmp.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (int l = 0; l < Q; ++l)
            for (int64_t i = 0; i < n; ++i) host[l * n + i] = (double)tmp[l * n + i] + ref[l];
    } else {
        HIP_TRY(c, hipMemcpyAsync(host, buf, (size_t)n * Q * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

namespace {
int species_field_ptr(lbm_ctx* c, int field, double** dev, size_t* count)
{
    if (!multi(c)) return fail(c, LBM_E_STATE, "no species in this context");
    const int64_t n = c->d.n;
    switch (field) {
    case LBM_SP_X: *dev = c->sp_block; *count = (size_t)c->ns * n; return LBM_OK;
    case LBM_SP_RHO: *dev = c->sp_block + (size_t)c->ns * n; *count = (size_t)c->ns * n; return LBM_OK;
    case LBM_SP_U:
    case LBM_SP_V:
    case LBM_SP_W:
        *dev = c->sp_block + (size_t)(2 + field - LBM_SP_U) * c->ns * n;
        *count = (size_t)c->ns * n;
        return LBM_OK;
    default: return fail(c, LBM_E_ARG, "species field id out of range");
    }
}
}  // namespace

int lbm_set_species_field(lbm_ctx* c, int field, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    double* dev;
    size_t cnt;
    if (int rc = species_field_ptr(c, field, &dev, &cnt)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipMemcpyAsync(dev, host, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_species_field(lbm_ctx* c, int field, double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    double* dev;
    size_t cnt;
    if (int rc = species_field_ptr(c, field, &dev, &cnt)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipMemcpyAsync(host, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_init(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede lbm_init");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (multi(c)) {
        if (int rc = run_mc(c, McKernel::Init, "mc_init", 0, 0)) return rc;
        c->fields_valid = true;
        return LBM_OK;
    }
    if (int rc = materialize(c)) return rc;
    if (int rc = run_init(c)) return rc;
    c->cur = 0;
    c->kind = lbm_ctx::EXPLICIT_F;
    c->fields_valid = true;
    c->q_valid = false;
    c->f_materialized = false;
    return LBM_OK;
}

int lbm_collide_species(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!multi(c)) return LBM_OK;  // Collide_Species exists only with species (lbm-main.cpp:156-158)
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    return run_mc(c, McKernel::CollideSpecies, "mc_collide_species", 0, 1);
}

int lbm_collide(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    if (multi(c)) return run_mc(c, McKernel::Collide, "mc_collide", 0, 1);
    if (!fused(c)) return run_collide_explicit(c, 0, 1);
    int rc;
    if (c->kind == lbm_ctx::EXPLICIT_F) {
        rc = run_collide_explicit(c, c->cur, 1 - c->cur);
    } else {
        if (!c->q_valid) return fail(c, LBM_E_STATE, "lbm_moments must run between two lbm_collide calls");
        rc = run_fused_collide(c, c->cur, 1 - c->cur);
    }
    if (rc) return rc;
    c->cur = 1 - c->cur;
    c->kind = lbm_ctx::POST_COLLISION;
    c->q_valid = false;
    c->fields_valid = false;
    c->f_materialized = false;
    return LBM_OK;
}

int lbm_stream(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    if (multi(c)) return run_mc(c, McKernel::Stream, "mc_stream", 1, 0);
    if (fused(c)) return LBM_OK;  // streaming happens in the next pull
    return run_stream(c, 1, 0);
}

int lbm_tms_bc(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (multi(c)) {
        if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
        if (!c->tms_unsupported.empty()) return fail(c, LBM_E_UNSUPPORTED, c->tms_unsupported);
        return launch(c, "mc_tms_outflow", [&] {
            launch_mc_tms<double>(make_mview(c, 0, 0), c->tms_cell, c->tms_dir, c->tms_count, c->stream);
        });
    }
    if (c->needs_tms)
        return fail(c, LBM_E_UNSUPPORTED,
                    "single-component TMS reconstruction next to TYPE_S/I/O cells is not provided by this build");
    return LBM_OK;  // periodic, adiabatic and free-slip walls need no reconstruction
}

int lbm_moments(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    if (multi(c)) return run_mc(c, McKernel::Moments, "mc_moments", 0, 0);
    if (!fused(c)) {
        if (int rc = run_moments(c, 0)) return rc;
        c->fields_valid = true;
        return LBM_OK;
    }
    if (c->kind == lbm_ctx::EXPLICIT_F) {
        if (int rc = run_moments(c, c->cur)) return rc;
        c->fields_valid = true;
        return LBM_OK;
    }
    if (int rc = run_fused_moments(c, c->cur, 1 - c->cur, false)) return rc;
    c->q_valid = true;
    return LBM_OK;
}

int lbm_step(lbm_ctx* c, int nsteps)
{
    if (!c) return LBM_E_ARG;
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        if ((rc = lbm_collide_species(c)) || (rc = lbm_collide(c)) || (rc = lbm_stream(c)) || (rc = lbm_tms_bc(c)) ||
            (rc = lbm_moments(c)))
            return rc;
    }
    return LBM_OK;
}

int lbm_diagnostics(lbm_ctx* c, lbm_diag* out)
{
    if (!c || !out) return fail(c, LBM_E_ARG, "null argument");
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede lbm_diagnostics");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (int rc = materialize(c)) return rc;
    if (int rc = launch(c, "diagnostics", [&] {
            launch_diag(c->d, c->type, c->fld, c->cfg.gamma > 1.0 ? c->cfg.gamma : 1.4, c->diag_part,
                        c->diag_npart, c->diag_out, c->stream);
        }))
        return rc;
    double h[8] = {};
    unsigned long long tau = 0;
    HIP_TRY(c, hipMemcpyAsync(h, c->diag_out, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&tau, c->tau_count, sizeof(tau), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    out->mass = h[0];
    out->energy = h[1];
    out->kinetic = h[2];
    out->enstrophy = h[3];
    out->nan_count = (int64_t)h[4];
    out->max_mach = h[5];
    out->tau_warnings = (int64_t)tau;
    out->fallback_count = 0;
    return LBM_OK;
}

int lbm_sync(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

void* lbm_get_stream(lbm_ctx* c) { return c ? (void*)c->stream : nullptr; }

int lbm_profile_enable(lbm_ctx* c, int on)
{
    if (!c) return LBM_E_ARG;
    c->prof = on != 0;
    return LBM_OK;
}

int lbm_profile_read(lbm_ctx* c, const char* name, double* total_ms, int64_t* launches)
{
    if (!c || !name || !total_ms || !launches) return fail(c, LBM_E_ARG, "null argument");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& r : c->pending) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, r.a, r.b));
        auto& acc = c->prof_acc[r.name];
        acc.first += ms;
        acc.second += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->pending.clear();
    *total_ms = 0.0;
    *launches = 0;
    for (auto& kv : c->prof_acc)
        if (kv.first.find(name) != std::string::npos) {
            *total_ms += kv.second.first;
            *launches += kv.second.second;
        }
    return LBM_OK;
}

int64_t lbm_device_bytes(const lbm_ctx* c) { return c ? c->bytes : 0; }

}  // extern "C"
