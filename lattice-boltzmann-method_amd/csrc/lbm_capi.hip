// C ABI of the MI355X LBM hot path (declarations and conventions: lbm_amd.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_internal.h"
#include "lbm_multi.h"
#include "surrogate.h"

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
    void* sp_block = nullptr;  // X, rho_a [NS][n] and u_a [3][NS][n], element type of the populations
    chem::Units units{};
    // TMS_BC outflow interface cells (cell, direction) and unsupported boundary kinds
    int* tms_cell = nullptr;
    int* tms_dir = nullptr;
    int tms_count = 0;
    double* tms_rec = nullptr;  // per-cell values between the outflow sweep's launches
    int* tin_cell = nullptr;  // TMS_BC inflow interface cells (TYPE_I, TYPE_I_C)
    int* tin_dir = nullptr;
    int tin_count = 0;
    int* twall_cell = nullptr;  // TMS_BC isothermal-wall interface cells (TYPE_S)
    int twall_count = 0;
    std::string tms_unsupported;
    // single-component TMS_BC interface cells (fluid cells next to TYPE_O/I/S)
    int* sbc_cell = nullptr;
    int sbc_count = 0;
    // conjugate variant: interior TYPE_S cells conduct heat; the solid
    // interface cells of its TMS_BC part
    bool conjugate = false;
    int* cjc_cell = nullptr;
    int cjc_count = 0;
    // multicomponent fused mode: Collide_Species writes each species
    // population to its streamed slot (buffer 1) and Streaming swaps the
    // species buffers and streams only the energy populations.  Cells without
    // NB_PUSH keep their post-collision species sets in the side store.  The
    // first step after Init or an upload runs strictly (push_ready false), so
    // that the streamed populations and the side store hold the post-collision
    // values a low species keeps.
    int* side_slot = nullptr;  // [n] side slot of a cell, -1 for none
    int* side_cell = nullptr;  // [nside]
    void* fside = nullptr;     // [NS][27][nside]
    int64_t nside = 0;
    int64_t side_bytes = 0;  // side store and its index planes (counted in `bytes`)
    bool push_ready = false;
    // Collide_Species has pushed this step's post-collision species sets into
    // their streamed slots of buffer 1 and Streaming has not run yet: buffer 0
    // still holds the pre-collision f, the post-collision set lives at the
    // downstream slots of buffer 1 (and in the side store)
    bool pushed = false;
    // in-place mode (LBM_MODE_INPLACE): species buffer sf[0] only, in layout
    // aa_swapped; aa_first until the first Collide_Species after lbm_init or
    // an upload (the kept post-collision values then come from aa_pending or
    // the rest state); the step's list of low species pairs and their kept
    // values (capacity low_cap, grown when a step finds more)
    bool aa_swapped = false;
    bool aa_first = true;
    void* aa_pending = nullptr;  // [NS][27][n] uploaded LBM_POP_FPC for the first step
    int* low_cell = nullptr;
    uint8_t* low_sp = nullptr;
    void* low_keep = nullptr;
    unsigned* low_count = nullptr;
    unsigned* low_count_host = nullptr;  // pinned
    int64_t low_cap = 0;
    int64_t low_bytes = 0;  // counted in `bytes`
    hipEvent_t ev_save = nullptr;
    // SMOOTHING variant (strict multicomponent): Jacobi buffer [27][n]
    bool smoothing = false;
    double smooth_alpha = 0.1;
    void* smooth_tmp = nullptr;
    void* src = nullptr;    // [NS][n] chemistry mass sources, element type of the populations
    void* tsp = nullptr;    // [4][NS][n] Collide_Species transport (omega_a, friction impulse)
    double* omix = nullptr; // FP32: [7][n] the mixture collide's rates (transport pass) and cell terms
    bool omix_fresh = false; // omix holds the rates of the current fields: set by the transport pass
                             // of lbm_collide_species, cleared by every call that changes the fields
    // FP32 multicomponent storage: rest-state shifts (host copies and device tables)
    double mfref[mech::NS * Q] = {}, mgref[Q] = {};
    double* mref_dev = nullptr;  // [NS*27 + 27]
    double* sums_part = nullptr;

    // chemistry surrogate
    sur::Mlp net;
    bool net_ready = false, norm_ready = false, sur_on = false;
    int adam_step = 0;
    sur::Work work;
    float* sur_stats = nullptr;       // [SS_SIZE]
    double* stats_part = nullptr;
    unsigned long long* sur_count = nullptr;  // [SUR_NCOUNT]
    unsigned* chem_list = nullptr;
    unsigned* chem_count = nullptr;
    double* tchem = nullptr;
    uint16_t* sur_x = nullptr;  // in-step network inputs [sur_rows][16] BF16
    float* sur_y = nullptr;     // outputs [sur_rows][16]
    uint8_t* sur_flag = nullptr;
    int64_t sur_rows = 0;
    float* cap_buf = nullptr;
    unsigned long long* cap_count = nullptr;
    long long cap_max = 0;
    int cap_stride = 0;
    bool capturing = false;
    int audit_stride = 0;
    float audit_tol = 0.f;
    unsigned step = 0;
    uint16_t* bx = nullptr;  // training batch, BF16 features
    float* bt = nullptr;     // training batch, targets
    float* bloss = nullptr;  // per-iteration losses
    float* adam_c12 = nullptr; // per-iteration Adam bias corrections [2][lcap] (device)
    int* train_it = nullptr;   // the training loop's iteration counter (device)
    std::vector<float> c12_host;  // host copy of adam_c12 while its upload is in flight
    hipGraph_t tgraph = nullptr;  // the captured training iteration of the last call
    hipGraphExec_t texec = nullptr;
    // asynchronous refresh: a training copy of the network on its own stream
    sur::Mlp tnet;
    float* stats_next = nullptr;  // [SS_SIZE] the envelope the copy is trained with
    hipStream_t tstream = nullptr;
    hipEvent_t ev_tstart = nullptr, ev_tdone = nullptr;
    bool refreshing = false;
    // lbm_step's overlapped schedule: the chemistry on its own stream beside
    // the transport pass and the mixture collide (LBM_AMD_OVERLAP=0: serial)
    int64_t low_max = INT64_MAX;  // LBM_AMD_LOW_MAX: most floor pairs the in-place list may hold at once
    hipStream_t cstream = nullptr;
    hipEvent_t ev_cstart = nullptr, ev_cdone = nullptr;
    bool overlap = true;
    int refresh_iters = 0;
    int bcap = 0, lcap = 0;

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

// Enqueues body() (kernels on stream s); with profiling on, brackets it by
// two events on s.
template <typename F>
int launch_on(lbm_ctx* c, hipStream_t s, const char* name, F&& body)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (c->prof) {
        HIP_TRY(c, hipEventCreate(&a));
        HIP_TRY(c, hipEventCreate(&b));
        HIP_TRY(c, hipEventRecord(a, s));
    }
    body();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, LBM_E_HIP, std::string("launch ") + name + ": " + hipGetErrorString(e));
    if (c->prof) {
        HIP_TRY(c, hipEventRecord(b, s));
        c->pending.push_back({name, a, b});
    }
    return LBM_OK;
}
template <typename F>
int launch(lbm_ctx* c, const char* name, F&& body)
{
    return launch_on(c, c->stream, name, body);
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
    v.conjugate = c->conjugate ? 1 : 0;
    return v;
}

template <typename P>
MultiView<P> make_mview(lbm_ctx* c, int src, int dst)
{
    MultiView<P> v{};
    v.d = c->d;
    v.type = c->type;
    v.nbflag = c->nbflag;
    v.f = static_cast<P*>(c->sf[src]);
    v.fo = static_cast<P*>(c->sf[dst]);
    v.g = static_cast<P*>(c->pg[src]);
    v.go = static_cast<P*>(c->pg[dst]);
    v.fref = c->mref_dev;
    v.gref = c->mref_dev ? c->mref_dev + mech::NS * Q : nullptr;
    for (int i = 0; i < LBM_FIELD_COUNT; ++i) v.fld[i] = c->fld[i];
    v.hrr = c->fld[LBM_FIELD_HRR];
    const int64_t n = c->d.n;
    P* const sp = static_cast<P*>(c->sp_block);
    v.X = sp;
    v.rhoa = sp + (size_t)c->ns * n;
    v.ua = sp + (size_t)2 * c->ns * n;
    v.u = c->units;
    v.reaction = c->cfg.reaction;
    v.tau_count = c->tau_count;
    v.src = static_cast<P*>(c->src);
    v.tsp = static_cast<P*>(c->tsp);
    v.omix = c->omix;
    v.use_omix = c->omix_fresh ? 1 : 0;
    v.tchem = c->tchem;
    v.step = c->step;
    if (c->capturing) {
        v.cap_buf = c->cap_buf;
        v.cap_count = c->cap_count;
        v.cap_max = c->cap_max;
        v.cap_stride = c->cap_stride;
    }
    v.sur_stats = c->sur_stats;
    v.sur_count = c->sur_count;
    v.audit_stride = c->audit_stride;
    v.audit_tol = c->audit_tol;
    v.side_slot = c->side_slot;
    v.side_cell = c->side_cell;
    v.fside = static_cast<P*>(c->fside);
    v.nside = c->nside;
    if (c->cfg.mode == LBM_MODE_INPLACE) {
        v.fo = v.f = static_cast<P*>(c->sf[0]);
        v.swapped = c->aa_swapped ? 1 : 0;
        v.low_cell = c->low_cell;
        v.low_sp = c->low_sp;
        v.low_keep = static_cast<P*>(c->low_keep);
        v.low_count = c->low_count;
        v.low_cap = c->low_cap;
        v.pending = static_cast<const P*>(c->aa_pending);
    }
    return v;
}

bool multi(const lbm_ctx* c) { return c->ns > 0; }

bool fp32(const lbm_ctx* c) { return c->cfg.precision == LBM_FP32_SHIFTED; }

int run_mc(lbm_ctx* c, McKernel k, const char* name, int src, int dst)
{
    if (fp32(c)) return launch(c, name, [&] { launch_mc<float>(k, make_mview<float>(c, src, dst), c->stream); });
    return launch(c, name, [&] { launch_mc<double>(k, make_mview<double>(c, src, dst), c->stream); });
}
bool fused(const lbm_ctx* c) { return c->cfg.mode == LBM_MODE_FUSED; }

// SMOOTHING of every species (lbm-smoothing.cpp:95-173) on the streamed
// populations of buffer 0.
int run_smoothing(lbm_ctx* c)
{
    if (fp32(c))
        return launch(c, "mc_smoothing",
                      [&] { launch_mc_smoothing<float>(make_mview<float>(c, 0, 0), c->smooth_tmp, c->smooth_alpha, c->stream); });
    return launch(c, "mc_smoothing",
                  [&] { launch_mc_smoothing<double>(make_mview<double>(c, 0, 0), c->smooth_tmp, c->smooth_alpha, c->stream); });
}
bool inplace(const lbm_ctx* c) { return c->cfg.mode == LBM_MODE_INPLACE; }

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

// Rest state of the FP32 multicomponent deviations: the zero-velocity
// equilibria (as Init builds them, lbm-init.cpp:85-146) of the mean TYPE_F
// temperature, pressure and composition of the set-up fields.
int set_multi_reference_state(lbm_ctx* c)
{
    const int64_t n = c->d.n;
    if (fp32(c))
        launch_mc_state_sums(c->type, c->fld[LBM_FIELD_T], c->fld[LBM_FIELD_P], static_cast<const float*>(c->sp_block),
                             n, c->sums_part, c->stream);
    else
        launch_mc_state_sums(c->type, c->fld[LBM_FIELD_T], c->fld[LBM_FIELD_P], static_cast<const double*>(c->sp_block),
                             n, c->sums_part, c->stream);
    std::vector<double> part((size_t)MC_STATS * MC_SUM_BLOCKS);
    HIP_TRY(c, hipMemcpyAsync(part.data(), c->sums_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double sum[MC_SUMS] = {}, xmin[mech::NS];
    for (int a = 0; a < mech::NS; ++a) xmin[a] = HUGE_VAL;
    for (int b = 0; b < MC_SUM_BLOCKS; ++b) {
        for (int q = 0; q < MC_SUMS; ++q) sum[q] += part[(size_t)b * MC_STATS + q];
        for (int a = 0; a < mech::NS; ++a) xmin[a] = std::min(xmin[a], part[(size_t)b * MC_STATS + MC_SUMS + a]);
    }
    if (!(sum[0] > 0)) return fail(c, LBM_E_STATE, "no TYPE_F cells");
    using namespace chem;
    double X[NS], Y[NS], xs = 0;
    for (int a = 0; a < NS; ++a) xs += (X[a] = std::max(sum[3 + a], 0.0));
    if (!(xs > 0)) return fail(c, LBM_E_STATE, "mole fractions of the set-up are all zero");
    for (int a = 0; a < NS; ++a) X[a] /= xs;
    // Binary32 deviations from the mean state resolve a species population only
    // to ~6e-8 of the mean's: a species whose set-up mole fraction falls below
    // 1e-6 of its mean somewhere is rounding noise there, and the run diverges
    // within steps (measured: a premixed flame with 1e-8 traces).  Refuse it.
    constexpr double kTraceFloor = 1e-6;
    for (int a = 0; a < NS; ++a)
        if (X[a] > 0 && !(xmin[a] >= kTraceFloor * X[a])) {
            char msg[256];
            snprintf(msg, sizeof msg,
                     "LBM_FP32_SHIFTED cannot resolve %s: its smallest set-up mole fraction %.3g is below 1e-6 of "
                     "its mean %.3g (raise the trace floor of the set-up or use LBM_FP64)",
                     mech::SPECIES_NAMES[a], xmin[a], X[a]);
            return fail(c, LBM_E_UNSUPPORTED, msg);
        }
    Y_from_X(X, Y);
    const Units& u = c->units;
    const double Tsi = u.si_temp(sum[1] / sum[0]);
    const double psi = u.si_p(sum[2] / sum[0]);
    if (!(Tsi > 0) || !(psi > 0)) return fail(c, LBM_E_STATE, "set-up temperature and pressure must be > 0");
    const double Wm = mean_W(Y);
    const double rho = u.rho(psi * Wm / (R_UNIV * Tsi));
    const double U = u.energy_mass(u_mass(Tsi, Y));
    const double zero[3] = {0, 0, 0};
    const GeqAxes<double> G = geq_axes<double>(u.energy_mass(R_UNIV * Tsi / Wm), zero);
    const double theta = u.energy_mass(R_UNIV * Tsi / Wm);
    for (int l = 0; l < Q; ++l) c->mgref[l] = geq_dir<double>(G, l, rho, U - 1.5 * theta);
    for (int a = 0; a < NS; ++a) {
        const double ra = std::max(Y[a] * rho, SPECIES_MIN);
        const Axis3<double> z = psi_axis<double>(0.0, u.energy_mass(R_UNIV * Tsi / W[a]));
        for (int l = 0; l < Q; ++l) c->mfref[a * Q + l] = ra * pick(z, cx(l)) * pick(z, cy(l)) * pick(z, cz(l));
    }
    HIP_TRY(c, hipMemcpyAsync(c->mref_dev, c->mfref, sizeof(double) * NS * Q, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->mref_dev + NS * Q, c->mgref, sizeof(double) * Q, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int refresh_finish(lbm_ctx* c, bool wait, int* applied);

void release(lbm_ctx* c)
{
    for (auto& r : c->pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->pending.clear();
    void* ptrs[] = {c->type,      c->nbflag,    c->pf[0],      c->pf[1],     c->pg[0],     c->pg[1],
                    c->fld_block, c->qd_block,  c->tau_count,  c->diag_part, c->diag_out,  c->sf[0],
                    c->sf[1],     c->sp_block,  c->tms_cell,   c->tms_dir,   c->src,       c->sur_stats,
                    c->stats_part, c->sur_count, c->chem_list, c->chem_count, c->tchem,    c->cap_buf,
                    c->cap_count, c->bx,        c->bt,         c->bloss,     c->mref_dev,  c->sums_part, c->adam_c12, c->train_it,
                    c->sur_x,     c->sur_y,     c->sur_flag,  c->tin_cell,  c->tin_dir,   c->sbc_cell,  c->tsp,        c->omix,       c->twall_cell, c->cjc_cell,
                    c->side_slot, c->side_cell, c->fside, c->aa_pending, c->low_cell, c->low_sp,
                    c->low_keep,  c->low_count, c->smooth_tmp, c->tms_rec};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->low_count_host) (void)hipHostFree(c->low_count_host);
    if (c->ev_save) (void)hipEventDestroy(c->ev_save);
    if (c->tstream) {  // an asynchronous refresh may still read the capture rows
        (void)hipStreamSynchronize(c->tstream);
        (void)hipStreamDestroy(c->tstream);
    }
    if (c->texec) (void)hipGraphExecDestroy(c->texec);
    if (c->tgraph) (void)hipGraphDestroy(c->tgraph);
    if (c->cstream) {
        (void)hipStreamSynchronize(c->cstream);
        (void)hipStreamDestroy(c->cstream);
    }
    if (c->ev_cstart) (void)hipEventDestroy(c->ev_cstart);
    if (c->ev_cdone) (void)hipEventDestroy(c->ev_cdone);
    if (c->ev_tstart) (void)hipEventDestroy(c->ev_tstart);
    if (c->ev_tdone) (void)hipEventDestroy(c->ev_tdone);
    if (c->tnet.Wf) sur::mlp_free(c->tnet);
    if (c->stats_next) (void)hipFree(c->stats_next);
    if (c->net_ready) sur::mlp_free(c->net);
    if (c->work.cap) sur::work_free(c->work);
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

const char* lbm_mechanism(int* n_species)
{
    if (n_species) *n_species = mech::NS;
    return mech::MECH_NAME;
}

const char* lbm_species_name(int k) { return (k >= 0 && k < mech::NS) ? mech::SPECIES_NAMES[k] : nullptr; }
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
        g_create_error = std::string("n_species must be 0 (single component) or ") + std::to_string(mech::NS) +
                         " (the " + mech::MECH_NAME + " mechanism of this library)";
        return LBM_E_UNSUPPORTED;
    }
    if (cfg->n_species == mech::NS) {
        if (!(cfg->dx_si > 0.0) || !(cfg->dt_si > 0.0)) {
            g_create_error = "multicomponent state needs dx_si > 0 and dt_si > 0 (units.set_m_kg_s)";
            return LBM_E_ARG;
        }
    }
    if (cfg->precision != LBM_FP64 && cfg->precision != LBM_FP32_SHIFTED) {
        g_create_error = "unknown precision";
        return LBM_E_ARG;
    }
    // the FP32 moment pass stages its per-species sums in LDS, 16 species at most
    if (cfg->n_species > 16 && cfg->precision == LBM_FP32_SHIFTED) {
        g_create_error = std::string("LBM_FP32_SHIFTED holds at most 16 species; mechanism ") + mech::MECH_NAME +
                         " has " + std::to_string(mech::NS) + " (use LBM_FP64)";
        return LBM_E_UNSUPPORTED;
    }
    if (cfg->mode != LBM_MODE_FUSED && cfg->mode != LBM_MODE_STRICT && cfg->mode != LBM_MODE_INPLACE) {
        g_create_error = "unknown mode";
        return LBM_E_ARG;
    }
    if (cfg->mode == LBM_MODE_INPLACE && cfg->n_species == 0) {
        g_create_error = "LBM_MODE_INPLACE streams species populations; single-component contexts use "
                         "LBM_MODE_FUSED or LBM_MODE_STRICT";
        return LBM_E_UNSUPPORTED;
    }
    if (cfg->n_species == 0 &&
        (!(cfg->gamma > 1.0) || !(cfg->gas_const > 0.0) || !(cfg->prtl > 0.0) || !(cfg->nu >= 0.0))) {
        g_create_error = "gas constants out of range (gamma > 1, R > 0, Pr > 0, nu >= 0)";
        return LBM_E_ARG;
    }
    const int64_t n = (int64_t)(cfg->nx + 2) * (cfg->ny + 2) * (cfg->nz + 2);
    // at most 2^29 cells: a plane of doubles stays within 32-bit byte offsets
    // (kernels address planes as a uniform base plus a 32-bit offset)
    if (n >= ((int64_t)1 << 29)) {
        g_create_error = "grid too large (at most 2^29 cells with ghosts)";
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
    if (const char* ov = getenv("LBM_AMD_OVERLAP")) c->overlap = ov[0] != '0';
    if (const char* lm = getenv("LBM_AMD_LOW_MAX")) c->low_max = std::max<int64_t>(1, std::atoll(lm));
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
    if (cfg->n_species == 0) {  // Q_d planes of the single-component fused collide
        if ((e = hipMalloc(&c->qd_block, c->esize * n * 3)) != hipSuccess) return bail("hipMalloc fluxes", e);
        for (int d = 0; d < 3; ++d) c->qd[d] = static_cast<char*>(c->qd_block) + (size_t)d * n * c->esize;
        (void)hipMemsetAsync(c->qd_block, 0, c->esize * n * 3, c->stream);
        c->bytes += c->esize * n * 3;
    }
    if ((e = hipMalloc(&c->tau_count, sizeof(unsigned long long))) != hipSuccess) return bail("hipMalloc", e);
    (void)hipMemsetAsync(c->tau_count, 0, sizeof(unsigned long long), c->stream);
    if (cfg->n_species > 0) {
        c->ns = cfg->n_species;
        c->units = chem::make_units(cfg->dx_si, cfg->dt_si);
        const size_t sbytes = (size_t)n * Q * c->ns * c->esize;
        const int nbuf = cfg->mode == LBM_MODE_INPLACE ? 1 : 2;
        for (int b = 0; b < nbuf; ++b) {
            if ((e = hipMalloc(&c->sf[b], sbytes)) != hipSuccess) return bail("hipMalloc species populations", e);
            (void)hipMemsetAsync(c->sf[b], 0, sbytes, c->stream);
        }
        if (cfg->mode == LBM_MODE_INPLACE) {
            if ((e = hipMalloc((void**)&c->low_count, sizeof(unsigned))) != hipSuccess) return bail("hipMalloc", e);
            if ((e = hipHostMalloc((void**)&c->low_count_host, sizeof(unsigned))) != hipSuccess)
                return bail("hipHostMalloc", e);
            if ((e = hipEventCreateWithFlags(&c->ev_save, hipEventDisableTiming)) != hipSuccess)
                return bail("hipEventCreate", e);
        }
        const size_t fb = c->esize * (size_t)n * c->ns * 5;
        if ((e = hipMalloc(&c->sp_block, fb)) != hipSuccess) return bail("hipMalloc species fields", e);
        (void)hipMemsetAsync(c->sp_block, 0, fb, c->stream);
        const size_t srcb = c->esize * (size_t)n * c->ns;
        if ((e = hipMalloc(&c->src, srcb)) != hipSuccess) return bail("hipMalloc chemistry sources", e);
        (void)hipMemsetAsync(c->src, 0, srcb, c->stream);
        const size_t tspb = c->esize * (size_t)n * c->ns * 4;
        if ((e = hipMalloc(&c->tsp, tspb)) != hipSuccess) return bail("hipMalloc species transport", e);
        c->bytes += tspb;
        if (fp32(c)) {
            if ((e = hipMalloc((void**)&c->omix, sizeof(double) * 7 * (size_t)n)) != hipSuccess)
                return bail("hipMalloc mixture collide terms", e);
            c->bytes += sizeof(double) * 7 * (size_t)n;
        }
        if (fp32(c)) {
            if ((e = hipMalloc((void**)&c->mref_dev, sizeof(double) * (mech::NS + 1) * Q)) != hipSuccess)
                return bail("hipMalloc", e);
            (void)hipMemsetAsync(c->mref_dev, 0, sizeof(double) * (mech::NS + 1) * Q, c->stream);
            if ((e = hipMalloc((void**)&c->sums_part, sizeof(double) * MC_STATS * MC_SUM_BLOCKS)) != hipSuccess)
                return bail("hipMalloc", e);
        }
        c->bytes += nbuf * sbytes + fb + srcb;
    }
    c->diag_npart = diag_partials(c->d);
    if ((e = hipMalloc(&c->diag_part, sizeof(double) * 8 * c->diag_npart)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = hipMalloc(&c->diag_out, sizeof(double) * 8)) != hipSuccess) return bail("hipMalloc", e);
    c->bytes += 2 * pop_bytes + 2 * n + sizeof(double) * n * LBM_FIELD_COUNT + sizeof(double) * (8 * c->diag_npart + 9);
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

namespace {

// The fused mode's plan on the host.  A fluid cell gets NB_PUSH when every
// population it pulls comes from a plain move (upstream neither a wall nor
// an in/outflow ghost) and every population it sends lands in a fluid cell,
// by a plain move, in a cell TMS_BC does not rewrite; the other fluid cells
// go to the side list.  Free-slip and TYPE_Q neighbours, and periodic
// images that are not fluid, are refused.
int push_plan(lbm_ctx* c, const uint8_t* types, const std::vector<int>& tcell, const std::vector<int>& icell,
              const std::vector<int>& wcell, std::vector<uint8_t>& flag, std::vector<int>& side)
{
    const Dims& d = c->d;
    std::vector<uint8_t> tms(d.n, 0);
    for (const auto* v : {&tcell, &icell, &wcell})
        for (int x : *v) tms[x] = 1;
    auto at = [&](int i, int j, int k) { return ((int64_t)i * d.ny + j) * d.nz + k; };
    auto wrap = [](int i, int n) { return i == 0 ? n - 2 : (i == n - 1 ? 1 : i); };
    auto pullable = [](uint8_t t) {
        return !(t == TYPE_S || t == TYPE_A || t == TYPE_FS || t == TYPE_O || t == TYPE_O_C || t == TYPE_I ||
                 t == TYPE_I_C);
    };
    side.clear();
    for (int i = 1; i < d.nx - 1; ++i)
        for (int j = 1; j < d.ny - 1; ++j)
            for (int k = 1; k < d.nz - 1; ++k) {
                const int64_t cidx = at(i, j, k);
                if (types[cidx] != TYPE_F) continue;
                bool push = !tms[cidx];
                for (int l = 0; l < Q; ++l) {
                    const uint8_t tu = types[at(i - cx(l), j - cy(l), k - cz(l))];
                    if (tu == TYPE_FS || tu == TYPE_Q)
                        return fail(c, LBM_E_UNSUPPORTED,
                                    "the multicomponent fused and in-place modes handle periodic, wall (TYPE_S, "
                                    "TYPE_A) and in/outflow neighbours; TYPE_FS/Q neighbours need LBM_MODE_STRICT");
                    if (pullable(tu) &&
                        types[at(wrap(i - cx(l), d.nx), wrap(j - cy(l), d.ny), wrap(k - cz(l), d.nz))] != TYPE_F)
                        return fail(c, LBM_E_UNSUPPORTED,
                                    "a periodic image of a fluid cell's neighbour is not fluid; the fused and "
                                    "in-place modes need LBM_MODE_STRICT here");
                    if (!pullable(tu)) push = false;
                    const uint8_t tt = types[at(i + cx(l), j + cy(l), k + cz(l))];
                    const int di = wrap(i + cx(l), d.nx), dj = wrap(j + cy(l), d.ny), dk = wrap(k + cz(l), d.nz);
                    const int64_t dc = at(di, dj, dk);
                    if (tt == TYPE_S || tt == TYPE_A || types[dc] != TYPE_F || tms[dc] ||
                        !pullable(types[at(di - cx(l), dj - cy(l), dk - cz(l))]))
                        push = false;
                }
                if (push) flag[cidx] |= NB_PUSH;
                else side.push_back((int)cidx);
            }
    return LBM_OK;
}

}  // namespace

int lbm_set_types(lbm_ctx* c, const uint8_t* types)
{
    if (!c || !types) return fail(c, LBM_E_ARG, "null argument");
    c->omix_fresh = false;
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
    std::vector<int> tcell, tdir, icell, idir, wcell;
    std::string tms_unsupported;
    if (multi(c)) {
        // the outflow sweep's interface detection (lbm-bc.cpp:141-155): the first
        // direction whose upstream cell is an outflow ghost and whose downstream
        // cell is fluid; only axis-aligned interfaces are reconstructed
        auto at = [&](int i, int j, int k) { return types[((int64_t)i * d.ny + j) * d.nz + k]; };
        for (int i = 1; i < d.nx - 1; ++i)
            for (int j = 1; j < d.ny - 1; ++j)
                for (int k = 1; k < d.nz - 1; ++k) {
                    if (at(i, j, k) != TYPE_F) continue;
                    const int cidx = (int)(((int64_t)i * d.ny + j) * d.nz + k);
                    // the inflow sweep's interface detection (lbm-bc.cpp:479-494)
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = at(i - cx(l), j - cy(l), k - cz(l));
                        const uint8_t tp = at(i + cx(l), j + cy(l), k + cz(l));
                        if ((tn == TYPE_I || tn == TYPE_I_C) && tp == TYPE_F) {
                            if (c2(l) == 1) {
                                icell.push_back(cidx);
                                idir.push_back(l);
                            }
                            break;
                        }
                    }
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = at(i - cx(l), j - cy(l), k - cz(l));
                        const uint8_t tp = at(i + cx(l), j + cy(l), k + cz(l));
                        if ((tn == TYPE_O || tn == TYPE_O_C) && tp == TYPE_F) {
                            if (c2(l) == 1) {
                                tcell.push_back(cidx);
                                tdir.push_back(l);
                            }
                            break;
                        }
                    }
                    // the wall sweep's interface links (lbm-bc.cpp:957-967)
                    for (int l = 0; l < Q; ++l)
                        if (at(i - cx(l), j - cy(l), k - cz(l)) == TYPE_S && at(i + cx(l), j + cy(l), k + cz(l)) == TYPE_F) {
                            wcell.push_back(cidx);
                            break;
                        }
                }
    }
    // single-component TMS_BC interface cells: some direction with a TYPE_O,
    // TYPE_I or TYPE_S upstream neighbour and a fluid downstream neighbour
    // (the interface tests of lbm-bc.cpp:357-368, 827-838, 1128-1137)
    std::vector<int> scell;
    if (!multi(c))
        for (int i = 1; i < d.nx - 1; ++i)
            for (int j = 1; j < d.ny - 1; ++j)
                for (int k = 1; k < d.nz - 1; ++k) {
                    const int64_t cidx = ((int64_t)i * d.ny + j) * d.nz + k;
                    if (types[cidx] != TYPE_F) continue;
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = types[((int64_t)(i - cx(l)) * d.ny + (j - cy(l))) * d.nz + (k - cz(l))];
                        const uint8_t tp = types[((int64_t)(i + cx(l)) * d.ny + (j + cy(l))) * d.nz + (k + cz(l))];
                        if ((tn == TYPE_O || tn == TYPE_I || tn == TYPE_S) && tp == TYPE_F) {
                            scell.push_back((int)cidx);
                            break;
                        }
                    }
                }
    // solid cells of the conjugate variant's TMS_BC part (lbm-bc.cpp:1246-1259):
    // interior TYPE_S cells with some direction whose upstream cell is gas
    // (TYPE_F or TYPE_Q) and whose downstream cell is solid
    std::vector<int> jcell;
    if (!multi(c))
        for (int i = 1; i < d.nx - 1; ++i)
            for (int j = 1; j < d.ny - 1; ++j)
                for (int k = 1; k < d.nz - 1; ++k) {
                    const int64_t cidx = ((int64_t)i * d.ny + j) * d.nz + k;
                    if (types[cidx] != TYPE_S) continue;
                    for (int l = 0; l < Q; ++l) {
                        const uint8_t tn = types[((int64_t)(i - cx(l)) * d.ny + (j - cy(l))) * d.nz + (k - cz(l))];
                        const uint8_t tp = types[((int64_t)(i + cx(l)) * d.ny + (j + cy(l))) * d.nz + (k + cz(l))];
                        if ((tn == TYPE_F || tn == TYPE_Q) && tp == TYPE_S) {
                            jcell.push_back((int)cidx);
                            break;
                        }
                    }
                }
    std::vector<int> side;
    if (multi(c) && (fused(c) || inplace(c)))
        if (int rc = push_plan(c, types, tcell, icell, wcell, flag, side)) return rc;
    if (!multi(c) && fused(c) && (needs_tms || !fused_ok))
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
        if (c->tms_rec) (void)hipFree(c->tms_rec);
        c->tms_cell = c->tms_dir = nullptr;
        c->tms_rec = nullptr;
        c->tms_count = (int)tcell.size();
        if (c->tms_count) {
            HIP_TRY(c, hipMalloc((void**)&c->tms_cell, sizeof(int) * tcell.size()));
            HIP_TRY(c, hipMalloc((void**)&c->tms_dir, sizeof(int) * tdir.size()));
            HIP_TRY(c, hipMalloc((void**)&c->tms_rec, sizeof(double) * mc_tms_record_fields() * tcell.size()));
            HIP_TRY(c, hipMemcpy(c->tms_cell, tcell.data(), sizeof(int) * tcell.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->tms_dir, tdir.data(), sizeof(int) * tdir.size(), hipMemcpyHostToDevice));
        }
        if (c->tin_cell) (void)hipFree(c->tin_cell);
        if (c->tin_dir) (void)hipFree(c->tin_dir);
        c->tin_cell = c->tin_dir = nullptr;
        c->tin_count = (int)icell.size();
        if (c->tin_count) {
            HIP_TRY(c, hipMalloc((void**)&c->tin_cell, sizeof(int) * icell.size()));
            HIP_TRY(c, hipMalloc((void**)&c->tin_dir, sizeof(int) * idir.size()));
            HIP_TRY(c, hipMemcpy(c->tin_cell, icell.data(), sizeof(int) * icell.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->tin_dir, idir.data(), sizeof(int) * idir.size(), hipMemcpyHostToDevice));
        }
        if (c->twall_cell) (void)hipFree(c->twall_cell);
        c->twall_cell = nullptr;
        c->twall_count = (int)wcell.size();
        if (c->twall_count) {
            HIP_TRY(c, hipMalloc((void**)&c->twall_cell, sizeof(int) * wcell.size()));
            HIP_TRY(c, hipMemcpy(c->twall_cell, wcell.data(), sizeof(int) * wcell.size(), hipMemcpyHostToDevice));
        }
        c->tms_unsupported = tms_unsupported;
        needs_tms = false;
        for (void** p : {(void**)&c->side_slot, (void**)&c->side_cell, &c->fside}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        c->nside = (int64_t)side.size();
        c->bytes -= c->side_bytes;
        c->side_bytes = 0;
        c->push_ready = false;
        c->pushed = false;
        c->aa_swapped = false;
        c->aa_first = true;
        if (fused(c) || inplace(c)) {
            std::vector<int> slot(d.n, -1);
            for (size_t s = 0; s < side.size(); ++s) slot[side[s]] = (int)s;
            HIP_TRY(c, hipMalloc((void**)&c->side_slot, sizeof(int) * d.n));
            c->side_bytes = sizeof(int) * d.n + (sizeof(int) + Q * c->ns * c->esize) * c->nside;
            c->bytes += c->side_bytes;
            HIP_TRY(c, hipMemcpy(c->side_slot, slot.data(), sizeof(int) * d.n, hipMemcpyHostToDevice));
            if (c->nside) {
                const size_t sb = (size_t)c->nside * Q * c->esize;
                HIP_TRY(c, hipMalloc((void**)&c->side_cell, sizeof(int) * side.size()));
                HIP_TRY(c, hipMemcpy(c->side_cell, side.data(), sizeof(int) * side.size(), hipMemcpyHostToDevice));
                HIP_TRY(c, hipMalloc(&c->fside, sb * c->ns));
                HIP_TRY(c, hipMemsetAsync(c->fside, 0, sb * c->ns, c->stream));
                HIP_TRY(c, hipStreamSynchronize(c->stream));
            }
        }
    }
    if (!multi(c)) {
        if (c->sbc_cell) (void)hipFree(c->sbc_cell);
        c->sbc_cell = nullptr;
        c->sbc_count = (int)scell.size();
        if (c->sbc_count) {
            HIP_TRY(c, hipMalloc((void**)&c->sbc_cell, sizeof(int) * scell.size()));
            HIP_TRY(c, hipMemcpy(c->sbc_cell, scell.data(), sizeof(int) * scell.size(), hipMemcpyHostToDevice));
        }
        if (c->cjc_cell) (void)hipFree(c->cjc_cell);
        c->cjc_cell = nullptr;
        c->cjc_count = (int)jcell.size();
        if (c->cjc_count) {
            HIP_TRY(c, hipMalloc((void**)&c->cjc_cell, sizeof(int) * jcell.size()));
            HIP_TRY(c, hipMemcpy(c->cjc_cell, jcell.data(), sizeof(int) * jcell.size(), hipMemcpyHostToDevice));
        }
    }
    c->needs_tms = needs_tms;
    c->types_set = true;
    return LBM_OK;
}

int lbm_set_field(lbm_ctx* c, int field, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    c->omix_fresh = false;
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

// The post-collision species sets between a fused-mode Collide_Species and
// Streaming, gathered on the host: cell c's population l of a NB_PUSH cell
// sits in buffer 1 at its streamed slot wrap(c + e_l) (the value a low
// species keeps included), the side cells' sets in the side store.  Other
// cells keep what buffer 1 holds there.
int read_pushed_fpc(lbm_ctx* c, double* host)
{
    const Dims& d = c->d;
    const int64_t n = d.n;
    const size_t count = (size_t)n * Q * c->ns;
    std::vector<uint8_t> type(n), flag(n);
    std::vector<int> side_cell(c->nside);
    std::vector<double> B(count), S((size_t)c->nside * Q * c->ns);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(type.data(), c->type, n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(flag.data(), c->nbflag, n, hipMemcpyDeviceToHost));
    if (c->nside) HIP_TRY(c, hipMemcpy(side_cell.data(), c->side_cell, sizeof(int) * c->nside, hipMemcpyDeviceToHost));
    const size_t nsz = (size_t)c->nside * Q * c->ns;
    if (fp32(c)) {
        std::vector<float> tb(count), ts(nsz);
        HIP_TRY(c, hipMemcpy(tb.data(), c->sf[1], count * 4, hipMemcpyDeviceToHost));
        if (nsz) HIP_TRY(c, hipMemcpy(ts.data(), c->fside, nsz * 4, hipMemcpyDeviceToHost));
        for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl) {
            for (int64_t i = 0; i < n; ++i) B[pl * n + i] = (double)tb[pl * n + i] + c->mfref[pl];
            for (int64_t s = 0; s < c->nside; ++s) S[pl * c->nside + s] = (double)ts[pl * c->nside + s] + c->mfref[pl];
        }
    } else {
        HIP_TRY(c, hipMemcpy(B.data(), c->sf[1], count * 8, hipMemcpyDeviceToHost));
        if (nsz) HIP_TRY(c, hipMemcpy(S.data(), c->fside, nsz * 8, hipMemcpyDeviceToHost));
    }
    std::copy(B.begin(), B.end(), host);
    auto wrap = [](int i, int m) { return i == 0 ? m - 2 : (i == m - 1 ? 1 : i); };
    for (int i = 1; i < d.nx - 1; ++i)
        for (int j = 1; j < d.ny - 1; ++j)
            for (int k = 1; k < d.nz - 1; ++k) {
                const int64_t ci = ((int64_t)i * d.ny + j) * d.nz + k;
                if (type[ci] != TYPE_F || !(flag[ci] & NB_PUSH)) continue;
                for (int l = 0; l < Q; ++l) {
                    const int64_t dst =
                        ((int64_t)wrap(i + cx(l), d.nx) * d.ny + wrap(j + cy(l), d.ny)) * d.nz + wrap(k + cz(l), d.nz);
                    for (int a = 0; a < c->ns; ++a)
                        host[((size_t)a * Q + l) * n + ci] = B[((size_t)a * Q + l) * n + dst];
                }
            }
    for (int64_t s = 0; s < c->nside; ++s)
        for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl) host[pl * n + side_cell[s]] = S[pl * c->nside + s];
    return LBM_OK;
}

}  // namespace

namespace {

// In-place mode: host copies of the single species buffer (values, the FP32
// shifts added back) and of the side store.
int aa_download(lbm_ctx* c, std::vector<double>& B, std::vector<double>& S)
{
    const int64_t n = c->d.n;
    const size_t count = (size_t)n * Q * c->ns, nsz = (size_t)c->nside * Q * c->ns;
    B.resize(count);
    S.resize(nsz);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (fp32(c)) {
        std::vector<float> tb(count), ts(nsz);
        HIP_TRY(c, hipMemcpy(tb.data(), c->sf[0], count * 4, hipMemcpyDeviceToHost));
        if (nsz) HIP_TRY(c, hipMemcpy(ts.data(), c->fside, nsz * 4, hipMemcpyDeviceToHost));
        for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl) {
            for (int64_t i = 0; i < n; ++i) B[pl * n + i] = (double)tb[pl * n + i] + c->mfref[pl];
            for (int64_t s = 0; s < c->nside; ++s) S[pl * c->nside + s] = (double)ts[pl * c->nside + s] + c->mfref[pl];
        }
    } else {
        HIP_TRY(c, hipMemcpy(B.data(), c->sf[0], count * 8, hipMemcpyDeviceToHost));
        if (nsz) HIP_TRY(c, hipMemcpy(S.data(), c->fside, nsz * 8, hipMemcpyDeviceToHost));
    }
    return LBM_OK;
}

// In-place mode: the species populations in the natural order (post: the
// post-collision set, else the incoming one), gathered on the host from the
// buffer's current layout.  Incoming population l of a fluid cell x sits at
// (l, x) in layout 0 and at (opp(l), x - e_l) in layout 1 (wrapped through
// periodic ghosts).  The post-collision population l of a NB_PUSH cell is the
// incoming population l of wrap(x + e_l); side cells keep theirs in the side
// store.  Before the first in-place step the post-collision set is the
// uploaded one, or the rest state.  Other cells read as zero.
int aa_gather(lbm_ctx* c, bool post, double* host)
{
    const Dims& d = c->d;
    const int64_t n = d.n;
    const size_t count = (size_t)n * Q * c->ns;
    if (post && c->aa_first) {
        if (c->aa_pending) {
            if (fp32(c)) {
                std::vector<float> t(count);
                HIP_TRY(c, hipMemcpy(t.data(), c->aa_pending, count * 4, hipMemcpyDeviceToHost));
                for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl)
                    for (int64_t i = 0; i < n; ++i) host[pl * n + i] = (double)t[pl * n + i] + c->mfref[pl];
            } else {
                HIP_TRY(c, hipMemcpy(host, c->aa_pending, count * 8, hipMemcpyDeviceToHost));
            }
        } else {
            for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl)
                for (int64_t i = 0; i < n; ++i) host[pl * n + i] = fp32(c) ? c->mfref[pl] : 0.0;
        }
        return LBM_OK;
    }
    std::vector<double> B, S;
    if (int rc = aa_download(c, B, S)) return rc;
    std::vector<uint8_t> type(n), flag(n);
    std::vector<int> side_cell(c->nside);
    HIP_TRY(c, hipMemcpy(type.data(), c->type, n, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(flag.data(), c->nbflag, n, hipMemcpyDeviceToHost));
    if (c->nside) HIP_TRY(c, hipMemcpy(side_cell.data(), c->side_cell, sizeof(int) * c->nside, hipMemcpyDeviceToHost));
    std::fill(host, host + count, 0.0);
    auto at = [&](int i, int j, int k) { return ((int64_t)i * d.ny + j) * d.nz + k; };
    auto wrap = [](int i, int m) { return i == 0 ? m - 2 : (i == m - 1 ? 1 : i); };
    // slot of the incoming population l of cell (i, j, k) in the current layout
    auto in_slot = [&](int i, int j, int k, int l) -> int64_t {
        if (!c->aa_swapped) return (int64_t)l * n + at(i, j, k);
        int si = i - cx(l), sj = j - cy(l), sk = k - cz(l);
        if (type[at(si, sj, sk)] == TYPE_P) {
            si = wrap(si, d.nx);
            sj = wrap(sj, d.ny);
            sk = wrap(sk, d.nz);
        }
        return (int64_t)opp(l) * n + at(si, sj, sk);
    };
    for (int i = 1; i < d.nx - 1; ++i)
        for (int j = 1; j < d.ny - 1; ++j)
            for (int k = 1; k < d.nz - 1; ++k) {
                const int64_t ci = at(i, j, k);
                if (type[ci] != TYPE_F) continue;
                if (post && !(flag[ci] & NB_PUSH)) continue;  // side store below
                for (int l = 0; l < Q; ++l) {
                    const int64_t sl = post ? in_slot(wrap(i + cx(l), d.nx), wrap(j + cy(l), d.ny), wrap(k + cz(l), d.nz), l)
                                            : in_slot(i, j, k, l);
                    for (int a = 0; a < c->ns; ++a)
                        host[((size_t)a * Q + l) * n + ci] = B[(size_t)a * Q * n + sl];
                }
            }
    if (post)
        for (int64_t s = 0; s < c->nside; ++s)
            for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl) host[pl * n + side_cell[s]] = S[pl * c->nside + s];
    return LBM_OK;
}

// In-place mode uploads.  LBM_POP_F: the buffer in layout 0.  LBM_POP_FPC:
// kept for the first step (the post-collision values a low species keeps)
// in a transient buffer, and copied into the side store.
int aa_upload(lbm_ctx* c, bool post, const double* host)
{
    const int64_t n = c->d.n;
    const size_t count = (size_t)n * Q * c->ns;
    void* dev = c->sf[0];
    if (post) {
        if (!c->aa_pending) HIP_TRY(c, hipMalloc(&c->aa_pending, count * c->esize));
        dev = c->aa_pending;
    }
    if (fp32(c)) {
        std::vector<float> tmp(count);
        for (size_t pl = 0; pl < (size_t)Q * c->ns; ++pl)
            for (int64_t i = 0; i < n; ++i) tmp[pl * n + i] = (float)(host[pl * n + i] - c->mfref[pl]);
        HIP_TRY(c, hipMemcpyAsync(dev, tmp.data(), count * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(dev, host, count * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    c->aa_first = true;
    if (!post) {
        c->aa_swapped = false;
        return LBM_OK;
    }
    // side cells take their post-collision sets from the upload (k_mc_side_fill reads v.fo)
    return DISPATCH(c, launch(c, "mc_side_fill", [&] {
        MultiView<P> v = make_mview<P>(c, 0, 0);
        v.fo = static_cast<P*>(c->aa_pending);
        launch_mc<P>(McKernel::SideFill, v, c->stream);
    }));
}

}  // namespace

int lbm_set_populations(lbm_ctx* c, int which, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    c->omix_fresh = false;
    if (which < 0 || which > 3) return fail(c, LBM_E_ARG, "population set out of range");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (multi(c)) {
        if (c->pushed)
            return fail(c, LBM_E_STATE,
                        "populations cannot be replaced between lbm_collide_species and lbm_stream in the fused "
                        "and in-place modes (the species are already pushed to their streamed slots)");
        if (inplace(c) && (which == LBM_POP_F || which == LBM_POP_FPC)) return aa_upload(c, which == LBM_POP_FPC, host);
        const int slot = (which == LBM_POP_F || which == LBM_POP_G) ? 0 : 1;
        const bool sp = (which == LBM_POP_F || which == LBM_POP_FPC);
        void* dev = sp ? c->sf[slot] : c->pg[slot];
        const size_t count = (size_t)c->d.n * Q * (sp ? c->ns : 1);
        if (fp32(c)) {
            const int64_t n = c->d.n;
            const double* ref = sp ? c->mfref : c->mgref;
            std::vector<float> tmp(count);
            for (size_t pl = 0; pl < count / n; ++pl)
                for (int64_t i = 0; i < n; ++i) tmp[pl * n + i] = (float)(host[pl * n + i] - ref[pl]);
            HIP_TRY(c, hipMemcpyAsync(dev, tmp.data(), count * 4, hipMemcpyHostToDevice, c->stream));
        } else {
            HIP_TRY(c, hipMemcpyAsync(dev, host, count * 8, hipMemcpyHostToDevice, c->stream));
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->push_ready = false;
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
        if (inplace(c) && (which == LBM_POP_F || which == LBM_POP_FPC)) {
            if (which == LBM_POP_F && c->pushed)
                return fail(c, LBM_E_STATE,
                            "the in-place mode has overwritten the pre-collision species populations: LBM_POP_F "
                            "cannot be read between lbm_collide_species and lbm_stream");
            return aa_gather(c, which == LBM_POP_FPC, host);
        }
        if (which == LBM_POP_FPC && c->pushed) return read_pushed_fpc(c, host);
        // the fused mode keeps no post-collision species buffer: rebuild it
        // from the streamed populations and the side store
        if (which == LBM_POP_FPC && fused(c) && c->push_ready)
            if (int rc = run_mc(c, McKernel::Unstream, "mc_unstream", 0, 1)) return rc;
        const void* dev = sp ? c->sf[slot] : c->pg[slot];
        const size_t count = (size_t)c->d.n * Q * (sp ? c->ns : 1);
        if (fp32(c)) {
            const int64_t n = c->d.n;
            const double* ref = sp ? c->mfref : c->mgref;
            std::vector<float> tmp(count);
            HIP_TRY(c, hipMemcpyAsync(tmp.data(), dev, count * 4, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            for (size_t pl = 0; pl < count / n; ++pl)
                for (int64_t i = 0; i < n; ++i) host[pl * n + i] = (double)tmp[pl * n + i] + ref[pl];
        } else {
            HIP_TRY(c, hipMemcpyAsync(host, dev, count * 8, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
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
// Device address and element count of a species plane (elements of c->esize
// bytes: binary32 in FP32 contexts).
int species_field_ptr(lbm_ctx* c, int field, void** dev, size_t* count)
{
    if (!multi(c)) return fail(c, LBM_E_STATE, "no species in this context");
    const int64_t n = c->d.n;
    char* const base = static_cast<char*>(c->sp_block);
    int plane;
    switch (field) {
    case LBM_SP_X: plane = 0; break;
    case LBM_SP_RHO: plane = 1; break;
    case LBM_SP_U:
    case LBM_SP_V:
    case LBM_SP_W: plane = 2 + field - LBM_SP_U; break;
    default: return fail(c, LBM_E_ARG, "species field id out of range");
    }
    *count = (size_t)c->ns * n;
    *dev = base + (size_t)plane * *count * c->esize;
    return LBM_OK;
}
}  // namespace

int lbm_set_species_field(lbm_ctx* c, int field, const double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    c->omix_fresh = false;
    void* dev;
    size_t cnt;
    if (int rc = species_field_ptr(c, field, &dev, &cnt)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (fp32(c)) {
        std::vector<float> tmp(host, host + cnt);
        HIP_TRY(c, hipMemcpyAsync(dev, tmp.data(), cnt * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(dev, host, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

int lbm_get_species_field(lbm_ctx* c, int field, double* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    void* dev;
    size_t cnt;
    if (int rc = species_field_ptr(c, field, &dev, &cnt)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (fp32(c)) {
        std::vector<float> tmp(cnt);
        HIP_TRY(c, hipMemcpyAsync(tmp.data(), dev, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        std::copy(tmp.begin(), tmp.end(), host);
    } else {
        HIP_TRY(c, hipMemcpyAsync(host, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

int lbm_init(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    c->omix_fresh = false;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede lbm_init");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (multi(c)) {
        if (fp32(c))
            if (int rc = set_multi_reference_state(c)) return rc;
        if (int rc = run_mc(c, McKernel::Init, "mc_init", 0, 0)) return rc;
        if (c->smoothing) {  // Init_smooting (lbm.hpp:170-174): Init, Smoothing, calculate_moment
            if (int rc = run_smoothing(c)) return rc;
            if (int rc = run_mc(c, McKernel::Moments, "mc_moments", 0, 0)) return rc;
        }
        c->fields_valid = true;
        c->push_ready = false;
        c->pushed = false;
        if (inplace(c)) {
            // fresh-context semantics: low species keep the rest state in the first step
            c->aa_swapped = false;
            c->aa_first = true;
            if (c->aa_pending) HIP_TRY(c, hipFree(c->aa_pending));
            c->aa_pending = nullptr;
            if (c->nside) HIP_TRY(c, hipMemsetAsync(c->fside, 0, (size_t)c->nside * Q * c->ns * c->esize, c->stream));
        }
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

namespace {

// (Re)allocates the low species list with room for `cap` entries.
// Whether a list of `cap` pairs fits: within the LBM_AMD_LOW_MAX pairs the
// environment allows (a test hook) and the device memory free once the
// current list is released.
bool low_list_fits(lbm_ctx* c, int64_t cap)
{
    if (cap > c->low_max) return false;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return true;
    const size_t need = (sizeof(int) + 1 + c->esize * Q) * (size_t)cap;
    return need + (64u << 20) <= free_b + c->low_bytes;
}

int aa_grow(lbm_ctx* c, int64_t cap)
{
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (void** p : {(void**)&c->low_cell, (void**)&c->low_sp, &c->low_keep}) {
        if (*p) HIP_TRY(c, hipFree(*p));
        *p = nullptr;
    }
    c->low_cap = 0;
    c->bytes -= c->low_bytes;
    c->low_bytes = 0;
    if (!low_list_fits(c, cap))
        return fail(c, LBM_E_HIP,
                    "the in-place mode's list of species held at SPECIES_MIN needs " + std::to_string(cap) +
                        " (cell, species) pairs, more than the device memory free");
    HIP_TRY(c, hipMalloc((void**)&c->low_cell, sizeof(int) * cap));
    HIP_TRY(c, hipMalloc((void**)&c->low_sp, cap));
    HIP_TRY(c, hipMalloc(&c->low_keep, c->esize * Q * cap));
    c->low_cap = cap;
    c->low_bytes = (sizeof(int) + 1 + c->esize * Q) * cap;
    c->bytes += c->low_bytes;
    return LBM_OK;
}

// In-place mode, first half of Collide_Species: list the low species pairs
// and save the post-collision values they keep (before the collision
// overwrites them); the count is read back asynchronously.
int aa_save(lbm_ctx* c)
{
    if (!c->low_cell) {  // first use: room for 1/16 of all (cell, species) pairs
        const int64_t cap = std::min<int64_t>(c->low_max, std::max<int64_t>(4096, c->d.n * c->ns / 16));
        if (int rc = aa_grow(c, cap)) return rc;
    }
    int rc = DISPATCH(c, launch(c, "mc_aa_save", [&] {
        launch_mc_aa_save<P>(make_mview<P>(c, 0, 0), c->aa_first, 0, (int)c->ns, c->stream);
    }));
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->low_count_host, c->low_count, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_save, c->stream));
    return LBM_OK;
}

// Second half: wait for the count (the chemistry and transport passes are
// queued behind it), grow the list and save again if it overflowed, then
// collide.  The layout flips.
int aa_collide_species(lbm_ctx* c)
{
    HIP_TRY(c, hipEventSynchronize(c->ev_save));
    const int64_t found = *c->low_count_host;
    const int ns = (int)c->ns;
    auto save = [&](int a0, int na) {
        return DISPATCH(c, launch(c, "mc_aa_save", [&] {
            launch_mc_aa_save<P>(make_mview<P>(c, 0, 0), c->aa_first, a0, na, c->stream);
        }));
    };
    auto collide = [&](int a0, int na, bool side) {
        return DISPATCH(c, launch(c, "mc_collide_species_inplace", [&] {
            launch_mc_aa_collide<P>(make_mview<P>(c, 0, 0), a0, na, side, c->stream);
        }));
    };
    if (found <= c->low_cap || low_list_fits(c, found + found / 4 + 4096)) {
        if (found > c->low_cap) {
            if (int rc = aa_grow(c, found + found / 4 + 4096)) return rc;
            if (int rc = save(0, ns)) return rc;
        }
        if (int rc = collide(0, ns, true)) return rc;
    } else {
        // every species' floor pairs at once would not fit: save and collide
        // one species at a time (a species' kept values sit only in its own
        // slots, which only its own collision overwrites), then stream the
        // side cells of all
        for (int a = 0; a < ns; ++a) {
            if (int rc = save(a, 1)) return rc;
            HIP_TRY(c, hipMemcpyAsync(c->low_count_host, c->low_count, sizeof(unsigned), hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            const int64_t fa = *c->low_count_host;
            if (fa > c->low_cap) {
                if (int rc = aa_grow(c, fa + fa / 4 + 4096)) return rc;
                if (int rc = save(a, 1)) return rc;
            }
            if (int rc = collide(a, 1, false)) return rc;
        }
        if (int rc = collide(0, 0, true)) return rc;
    }
    c->aa_swapped = !c->aa_swapped;
    c->aa_first = false;
    c->pushed = true;
    if (c->aa_pending) {  // consumed by the save pass (hipFree waits for it)
        HIP_TRY(c, hipFree(c->aa_pending));
        c->aa_pending = nullptr;
    }
    return LBM_OK;
}

}  // namespace

int lbm_collide_species(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!multi(c)) return LBM_OK;  // Collide_Species exists only with species (lbm-main.cpp:156-158)
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (inplace(c)) {
        if (c->pushed)
            return fail(c, LBM_E_STATE,
                        "the in-place mode's species are already collided into the next layout: lbm_stream must "
                        "follow lbm_collide_species");
        if (int rc = aa_save(c)) return rc;
    }
    if (c->cfg.reaction) {
        int rc;
        if (c->sur_on) {
            rc = DISPATCH(c, [&] {
                const MultiView<P> v = make_mview<P>(c, 0, 1);
                int r = launch(c, "mc_chemistry_surrogate", [&] {
                    launch_mc_sur_chemistry<P>(v, c->net, c->sur_x, c->sur_y, c->sur_flag, c->sur_rows,
                                               c->chem_list, c->chem_count, c->stream);
                });
                if (!r)
                    r = launch(c, "mc_chemistry_fallback",
                               [&] { launch_mc_chem_list<P>(v, c->chem_list, c->chem_count, c->stream); });
                return r;
            }());
        } else {
            rc = run_mc(c, McKernel::Chemistry, "mc_chemistry_plain", 0, 1);
        }
        if (rc) return rc;
    }
    int rc = run_mc(c, McKernel::SpeciesTransport, "mc_species_transport", 0, 1);
    c->omix_fresh = !rc && fp32(c);  // the FP32 pass also wrote the mixture collide's rates
    if (inplace(c)) {
        if (!rc) rc = aa_collide_species(c);
        ++c->step;
        return rc;
    }
    const bool push = fused(c) && c->push_ready;
    if (!rc)
        rc = push ? run_mc(c, McKernel::CollideSpeciesPush, "mc_collide_species_push", 0, 1)
                  : run_mc(c, McKernel::CollideSpecies, "mc_collide_species", 0, 1);
    if (!rc && push) c->pushed = true;
    ++c->step;
    return rc;
}

int lbm_collide(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    if (multi(c)) return run_mc(c, McKernel::Collide, "mc_collide_mixture", 0, 1);
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
    if (multi(c)) {
        if (inplace(c)) {  // Collide_Species wrote the species into the next step's layout
            if (!c->pushed)
                return fail(c, LBM_E_STATE, "lbm_collide_species must run before lbm_stream in the in-place mode");
            c->pushed = false;
            return run_mc(c, McKernel::StreamEnergy, "mc_stream_energy", 1, 0);
        }
        if (fused(c) && c->push_ready) {  // Collide_Species wrote the streamed species slots of buffer 1
            if (!c->pushed)
                return fail(c, LBM_E_STATE, "lbm_collide_species must run before lbm_stream in the fused mode");
            std::swap(c->sf[0], c->sf[1]);
            c->pushed = false;
            return run_mc(c, McKernel::StreamEnergy, "mc_stream_energy", 1, 0);
        }
        if (int rc = run_mc(c, McKernel::Stream, "mc_stream", 1, 0)) return rc;
        if (!fused(c)) return LBM_OK;
        if (int rc = run_mc(c, McKernel::SideFill, "mc_side_fill", 0, 1)) return rc;
        c->push_ready = true;
        return LBM_OK;
    }
    if (fused(c)) return LBM_OK;  // streaming happens in the next pull
    return run_stream(c, 1, 0);
}

int lbm_tms_bc(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    c->omix_fresh = false;
    if (multi(c)) {
        if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
        if (!c->tms_unsupported.empty()) return fail(c, LBM_E_UNSUPPORTED, c->tms_unsupported);
        // outflow, inflow and isothermal-wall sweeps (lbm-bc.cpp:126, 463, 941)
        if (int rc = DISPATCH(c, launch(c, "mc_tms_outflow", [&] {
                launch_mc_tms<P>(make_mview<P>(c, 0, 0), c->tms_cell, c->tms_dir, c->tms_count, c->tms_rec,
                                 c->stream);
            })))
            return rc;
        if (int rc = DISPATCH(c, launch(c, "mc_tms_inflow", [&] {
                launch_mc_tms_in<P>(make_mview<P>(c, 0, 0), c->tin_cell, c->tin_dir, c->tin_count, c->stream);
            })))
            return rc;
        return DISPATCH(c, launch(c, "mc_tms_wall", [&] {
            launch_mc_tms_wall<P>(make_mview<P>(c, 0, 0), c->twall_cell, c->twall_count, c->stream);
        }));
    }
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    // fused mode rejects TYPE_S/I/O neighbours at lbm_set_types; periodic,
    // adiabatic and free-slip walls need no reconstruction
    if (fused(c)) return LBM_OK;
    // outflow, inflow and isothermal-wall sweeps (lbm-bc.cpp:341, 809, 1111) on
    // the streamed populations of buffer 0
    if (c->sbc_count)
        if (int rc = DISPATCH(c, launch(c, "tms", [&] {
                launch_tms<P>(make_view<P>(c, 0, 0), c->sbc_cell, c->sbc_count, c->stream);
            })))
            return rc;
    // then the conjugate variant's solid side (lbm-bc.cpp:1230-1346)
    if (c->conjugate && c->cjc_count)
        return DISPATCH(c, launch(c, "tms_conj", [&] {
            launch_tms_conj<P>(make_view<P>(c, 0, 0), c->cjc_cell, c->cjc_count, c->stream);
        }));
    return LBM_OK;
}

int lbm_set_smoothing(lbm_ctx* c, int on, double alpha)
{
    if (!c) return LBM_E_ARG;
    if (on && !multi(c)) return fail(c, LBM_E_UNSUPPORTED, "SMOOTHING filters species populations (MULTICOMP only)");
    if (on && !(c->cfg.mode == LBM_MODE_STRICT))
        return fail(c, LBM_E_UNSUPPORTED,
                    "SMOOTHING reads the populations of in/outflow ghosts, which only LBM_MODE_STRICT keeps as the "
                    "reference does");
    if (on && !(alpha >= 0.0 && std::isfinite(alpha))) return fail(c, LBM_E_ARG, "smoothing alpha must be >= 0");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (on && !c->smooth_tmp) {
        const size_t bytes = (size_t)c->d.n * Q * c->esize;
        HIP_TRY(c, hipMalloc(&c->smooth_tmp, bytes));
        c->bytes += bytes;
    }
    c->smoothing = on != 0;
    c->smooth_alpha = alpha;
    return LBM_OK;
}

int lbm_set_conjugate(lbm_ctx* c, int on)
{
    if (!c) return LBM_E_ARG;
    if (on && multi(c)) return fail(c, LBM_E_UNSUPPORTED, "the conjugate variant is single-component");
    if (on && fused(c)) return fail(c, LBM_E_UNSUPPORTED, "the conjugate variant needs LBM_MODE_STRICT");
    c->conjugate = on != 0;
    return LBM_OK;
}

int lbm_moments(lbm_ctx* c)
{
    if (!c) return LBM_E_ARG;
    c->omix_fresh = false;
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede stepping");
    if (multi(c)) {
        if (int rc = run_mc(c, McKernel::Moments, "mc_moments", 0, 0)) return rc;
        if (!c->smoothing) return LBM_OK;
        // calculate_moment_smoothing (lbm.hpp:154-157): moments, Smoothing, moments
        if (int rc = run_smoothing(c)) return rc;
        return run_mc(c, McKernel::Moments, "mc_moments", 0, 0);
    }
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

namespace {

// Collide_Species and Collide of one fused-mode reacting step with the
// chemistry on a second stream.  The chemistry reads the fields and species
// planes of the last calculate_moment and writes only its own outputs
// (sources, heat release, the fallback list, capture rows); the transport
// pass and the mixture collide neither read nor write those, so the whole
// chemistry (network pass and fallback list) runs beside them and the push
// collide, which needs its sources, waits for it.  Each kernel sees exactly
// the inputs of the serial calls: the state is bitwise that of
// lbm_collide_species followed by lbm_collide.  (Measured on the 256x128x128
// layer: 557 against 538 MLUPS serial; the chemistry beside the mixture
// collide alone, or the fallback list after the wait, gained less.)
int collide_overlapped(lbm_ctx* c)
{
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (!c->cstream) {
        HIP_TRY(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_cstart, hipEventDisableTiming));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_cdone, hipEventDisableTiming));
    }
    if (inplace(c) && c->pushed)
        return fail(c, LBM_E_STATE,
                    "the in-place mode's species are already collided into the next layout: lbm_stream must "
                    "follow lbm_collide_species");
    HIP_TRY(c, hipEventRecord(c->ev_cstart, c->stream));
    HIP_TRY(c, hipStreamWaitEvent(c->cstream, c->ev_cstart, 0));
    // the in-place mode's save pass (it reads the species populations and
    // densities, which the chemistry leaves alone)
    if (inplace(c))
        if (int rc = aa_save(c)) {
            (void)hipEventRecord(c->ev_cdone, c->cstream);
            return rc;
        }
    int rc = DISPATCH(c, [&] {
        const MultiView<P> v = make_mview<P>(c, 0, 1);
        if (!c->sur_on)
            return launch_on(c, c->cstream, "mc_chemistry_plain",
                             [&] { launch_mc<P>(McKernel::Chemistry, v, c->cstream); });
        int r = launch_on(c, c->cstream, "mc_chemistry_surrogate", [&] {
            launch_mc_sur_chemistry<P>(v, c->net, c->sur_x, c->sur_y, c->sur_flag, c->sur_rows, c->chem_list,
                                       c->chem_count, c->cstream);
        });
        if (!r)
            r = launch_on(c, c->cstream, "mc_chemistry_fallback",
                          [&] { launch_mc_chem_list<P>(v, c->chem_list, c->chem_count, c->cstream); });
        return r;
    }());
    // recorded whatever happened, so that the main stream's wait ends
    HIP_TRY(c, hipEventRecord(c->ev_cdone, c->cstream));
    if (!rc) rc = run_mc(c, McKernel::SpeciesTransport, "mc_species_transport", 0, 1);
    c->omix_fresh = !rc && fp32(c);
    if (!rc) rc = run_mc(c, McKernel::Collide, "mc_collide_mixture", 0, 1);
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_cdone, 0));
    if (!rc && inplace(c)) {
        rc = aa_collide_species(c);
    } else if (!rc) {
        rc = run_mc(c, McKernel::CollideSpeciesPush, "mc_collide_species_push", 0, 1);
        if (!rc) c->pushed = true;
    }
    ++c->step;
    return rc;
}

}  // namespace

int lbm_step(lbm_ctx* c, int nsteps)
{
    if (!c) return LBM_E_ARG;
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        // the overlapped schedule covers the fused mode's pushed steps and the
        // in-place mode, with chemistry
        const bool ov = c->overlap && multi(c) && c->types_set && c->cfg.reaction &&
                        ((fused(c) && c->push_ready) || inplace(c));
        if (ov) {
            if ((rc = collide_overlapped(c))) return rc;
        } else if ((rc = lbm_collide_species(c)) || (rc = lbm_collide(c))) {
            return rc;
        }
        if ((rc = lbm_stream(c)) || (rc = lbm_tms_bc(c)) || (rc = lbm_moments(c))) return rc;
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
            launch_diag_partials(c->d, c->type, c->fld, c->cfg.gamma > 1.0 ? c->cfg.gamma : 1.4, c->diag_part,
                                 c->stream);
            // with species: the Mach number with each cell's mixture gamma
            if (multi(c)) {
                if (fp32(c)) launch_mc_diag_mach<float>(make_mview<float>(c, 0, 0), c->diag_part, c->stream);
                else launch_mc_diag_mach<double>(make_mview<double>(c, 0, 0), c->diag_part, c->stream);
            }
            launch_diag_final(c->diag_part, c->diag_npart, c->diag_out, c->stream);
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
    if (c->sur_count) {
        unsigned long long cnt[SUR_NCOUNT];
        HIP_TRY(c, hipMemcpy(cnt, c->sur_count, sizeof(cnt), hipMemcpyDeviceToHost));
        out->fallback_count = (int64_t)(cnt[SUR_FALLBACK] + cnt[SUR_AUDIT_FAIL]);
    }
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

// ------------------------------------------------------ chemistry surrogate
namespace {

constexpr uint64_t SUR_DEFAULT_SEED = 20261202ull;

// The network's 16-wide input and output layers hold the NS + 3 features and
// NS + 1 targets of mechanisms of up to 11 species; a build for a larger
// mechanism refuses the surrogate's entry points.
int sur_supported(lbm_ctx* c)
{
    if (sur::ENABLED) return LBM_OK;
    return fail(c, LBM_E_UNSUPPORTED,
                std::string("the chemistry surrogate's 16-wide layers hold at most 11 species; mechanism ") +
                    mech::MECH_NAME + " has " + std::to_string(mech::NS));
}

int sur_common(lbm_ctx* c)
{
    if (int rc = sur_supported(c)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (!c->sur_stats) {
        HIP_TRY(c, hipMalloc((void**)&c->sur_stats, sizeof(float) * sur::SS_SIZE));
        HIP_TRY(c, hipMemsetAsync(c->sur_stats, 0, sizeof(float) * sur::SS_SIZE, c->stream));
        HIP_TRY(c, hipMalloc((void**)&c->stats_part, sizeof(double) * sur::STATS_BLOCKS * 32 * 4));
        HIP_TRY(c, hipMalloc((void**)&c->sur_count, sizeof(unsigned long long) * SUR_NCOUNT));
        HIP_TRY(c, hipMemsetAsync(c->sur_count, 0, sizeof(unsigned long long) * SUR_NCOUNT, c->stream));
    }
    if (!c->net_ready) {
        HIP_TRY(c, sur::mlp_alloc(c->net, SUR_DEFAULT_SEED, c->stream));
        c->net_ready = true;
        c->adam_step = 0;
    }
    return LBM_OK;
}

// Training scratch for batches of up to `rows` rows.
int sur_work(lbm_ctx* c, int rows)
{
    if (c->work.cap >= rows) return LBM_OK;
    if (c->work.cap) sur::work_free(c->work);
    HIP_TRY(c, sur::work_alloc(c->work, rows));
    return LBM_OK;
}

int sur_batch(lbm_ctx* c, int rows)
{
    if (c->bcap >= rows) return LBM_OK;
    if (c->bx) (void)hipFree(c->bx);
    if (c->bt) (void)hipFree(c->bt);
    c->bx = nullptr;
    c->bt = nullptr;
    HIP_TRY(c, hipMalloc((void**)&c->bx, sizeof(uint16_t) * 16 * (size_t)rows));
    HIP_TRY(c, hipMalloc((void**)&c->bt, sizeof(float) * 16 * (size_t)rows));
    c->bcap = rows;
    return LBM_OK;
}

int copy_params(lbm_ctx* c, float* host, bool to_host)
{
    size_t off = 0;
    for (int L = 0; L < sur::NL; ++L) {
        const size_t nw = (size_t)sur::DIMS[L] * sur::DIMS[L + 1], nb = sur::DIMS[L + 1];
        if (to_host) {
            HIP_TRY(c, hipMemcpyAsync(host + off, c->net.W[L], 4 * nw, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipMemcpyAsync(host + off + nw, c->net.b[L], 4 * nb, hipMemcpyDeviceToHost, c->stream));
        } else {
            HIP_TRY(c, hipMemcpyAsync(c->net.W[L], host + off, 4 * nw, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->net.b[L], host + off + nw, 4 * nb, hipMemcpyHostToDevice, c->stream));
        }
        off += nw + nb;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int need_reacting(lbm_ctx* c)
{
    if (!multi(c)) return fail(c, LBM_E_STATE, "the chemistry surrogate needs a multicomponent context");
    if (!c->cfg.reaction) return fail(c, LBM_E_STATE, "the chemistry surrogate needs reaction = 1");
    return LBM_OK;
}

}  // namespace

extern "C" {

int64_t lbm_surrogate_param_count(void) { return sur::NPARAM; }

int lbm_surrogate_init(lbm_ctx* c, uint64_t seed)
{
    if (!c) return LBM_E_ARG;
    if (int rc = sur_supported(c)) return rc;
    if (int rc = refresh_finish(c, true, nullptr)) return rc;  // no refresh in flight
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->net_ready) sur::mlp_free(c->net);
    c->net_ready = false;
    HIP_TRY(c, sur::mlp_alloc(c->net, seed, c->stream));
    c->net_ready = true;
    c->adam_step = 0;
    c->norm_ready = false;
    c->sur_on = false;
    return sur_common(c);
}

int lbm_surrogate_capture(lbm_ctx* c, int64_t max_rows, int stride)
{
    if (!c) return LBM_E_ARG;
    if (int rc = sur_supported(c)) return rc;
    if (int rc = refresh_finish(c, true, nullptr)) return rc;  // no refresh in flight
    if (int rc = need_reacting(c)) return rc;
    if (max_rows < 0 || max_rows > ((int64_t)1 << 31) || stride < 0)
        return fail(c, LBM_E_ARG, "capture needs 0 <= max_rows <= 2^31 and stride >= 0");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (stride == 0 && max_rows > 0) {  // pause: keep what was captured
        c->capturing = false;
        return LBM_OK;
    }
    if (max_rows == 0) {
        if (c->cap_buf) (void)hipFree(c->cap_buf);
        if (c->cap_count) (void)hipFree(c->cap_count);
        c->cap_buf = nullptr;
        c->cap_count = nullptr;
        c->cap_max = 0;
        c->capturing = false;
        return LBM_OK;
    }
    if (!c->cap_count) {
        HIP_TRY(c, hipMalloc((void**)&c->cap_count, sizeof(unsigned long long)));
        HIP_TRY(c, hipMemset(c->cap_count, 0, sizeof(unsigned long long)));
    }
    if (max_rows > c->cap_max) {
        float* nb = nullptr;
        HIP_TRY(c, hipMalloc((void**)&nb, sizeof(float) * sur::CAP_STRIDE * (size_t)max_rows));
        if (c->cap_buf) {
            unsigned long long have = 0;
            HIP_TRY(c, hipMemcpy(&have, c->cap_count, sizeof(have), hipMemcpyDeviceToHost));
            const size_t keep = (size_t)std::min<long long>((long long)have, c->cap_max);
            HIP_TRY(c, hipMemcpy(nb, c->cap_buf, sizeof(float) * sur::CAP_STRIDE * keep, hipMemcpyDeviceToDevice));
            (void)hipFree(c->cap_buf);
        }
        c->cap_buf = nb;
        c->cap_max = max_rows;
    }
    c->cap_stride = stride;
    c->capturing = true;
    return LBM_OK;
}

int lbm_surrogate_get_capture(lbm_ctx* c, float* host, int64_t max_rows, int64_t* rows_out)
{
    if (!c || !rows_out || (max_rows > 0 && !host)) return fail(c, LBM_E_ARG, "null argument");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    unsigned long long have = 0;
    if (c->cap_count) HIP_TRY(c, hipMemcpy(&have, c->cap_count, sizeof(have), hipMemcpyDeviceToHost));
    const int64_t rows = std::min<int64_t>((int64_t)have, c->cap_max);
    const int64_t m = std::min<int64_t>(rows, max_rows);
    if (m > 0) HIP_TRY(c, hipMemcpy(host, c->cap_buf, sizeof(float) * sur::CAP_STRIDE * m, hipMemcpyDeviceToHost));
    *rows_out = rows;
    return LBM_OK;
}

namespace {

// Buffers of a training call of `iters` iterations on `batch` rows.
int train_buffers(lbm_ctx* c, int iters, int batch)
{
    if (int rc = sur_work(c, batch)) return rc;
    if (int rc = sur_batch(c, batch)) return rc;
    if (iters > c->lcap || !c->adam_c12) {
        if (c->bloss) (void)hipFree(c->bloss);
        if (c->adam_c12) (void)hipFree(c->adam_c12);
        c->bloss = nullptr;
        c->adam_c12 = nullptr;
        HIP_TRY(c, hipMalloc((void**)&c->bloss, sizeof(float) * std::max(iters, 1)));
        HIP_TRY(c, hipMalloc((void**)&c->adam_c12, sizeof(float) * 2 * std::max(iters, 1)));
        c->lcap = std::max(iters, 1);
    }
    if (!c->train_it) HIP_TRY(c, hipMalloc((void**)&c->train_it, sizeof(int)));
    return LBM_OK;
}

void train_release(lbm_ctx* c)
{
    if (c->texec) (void)hipGraphExecDestroy(c->texec);
    if (c->tgraph) (void)hipGraphDestroy(c->tgraph);
    c->texec = nullptr;
    c->tgraph = nullptr;
}

// Enqueues `iters` training iterations of `net` on stream st, from the first
// `rows` capture rows normalised by `stats`, with Adam steps adam_step + 1 ...
// One iteration (batch draw, forward and backward passes, loss, Adam: 45
// kernels) is captured once as a HIP graph and replayed `iters` times; the
// iteration index and the step's Adam corrections are read on the device.
// The graph is kept in c->tgraph / c->texec until train_release, after st
// has finished.
int train_enqueue(lbm_ctx* c, sur::Mlp& net, const float* stats, int64_t rows, int iters, int batch, float lr,
                  uint64_t seed, hipStream_t st)
{
    train_release(c);
    c->c12_host.assign(2 * (size_t)iters, 0.f);
    sur::adam_corrections(c->adam_step + 1, iters, c->c12_host.data());
    static const int zero = 0;
    HIP_TRY(c, hipMemcpyAsync(c->adam_c12, c->c12_host.data(), sizeof(float) * c->c12_host.size(),
                              hipMemcpyHostToDevice, st));
    HIP_TRY(c, hipMemcpyAsync(c->train_it, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    auto iteration = [&] {
        sur::gather_batch(c->cap_buf, rows, stats, seed, c->train_it, batch, c->bx, c->bt, st);
        sur::mlp_grad(net, c->bx, c->bt, batch, c->work, st);
        sur::loss_sum(c->work.loss_rows, batch, 1.f / ((float)batch * sur::NOUT), c->bloss, st, c->train_it);
        sur::mlp_adam(net, c->work, lr, c->adam_c12, c->train_it, st);
        sur::next_iteration(c->train_it, st);
    };
    // LBM_AMD_TRAIN_GRAPH=0 launches the iterations one by one (the same
    // kernels and arguments): rocprofv3's kernel tracing of this ROCm faults
    // on the host inside the graph calls
    const char* env = getenv("LBM_AMD_TRAIN_GRAPH");
    bool graphed = !(env && env[0] == '0') && hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
    if (graphed) {
        iteration();
        graphed = hipStreamEndCapture(st, &c->tgraph) == hipSuccess && c->tgraph &&
                  hipGraphInstantiate(&c->texec, c->tgraph, nullptr, nullptr, 0) == hipSuccess;
    }
    if (graphed) {
        for (int it = 0; it < iters; ++it) HIP_TRY(c, hipGraphLaunch(c->texec, st));
    } else {
        (void)hipGetLastError();
        train_release(c);
        for (int it = 0; it < iters; ++it) iteration();
    }
    HIP_TRY(c, hipGetLastError());
    return LBM_OK;
}

int64_t captured_rows(lbm_ctx* c)
{
    unsigned long long have = 0;
    if (c->cap_count && hipMemcpy(&have, c->cap_count, sizeof(have), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return std::min<int64_t>((int64_t)have, c->cap_max);
}

// Applies a finished asynchronous refresh (the network and envelope it
// trained replace the serving ones, on the step stream); with wait, waits
// for it first.  *applied = 1 when this call applied one.
int refresh_finish(lbm_ctx* c, bool wait, int* applied)
{
    if (applied) *applied = 0;
    if (!c->refreshing) return LBM_OK;
    if (wait) {
        HIP_TRY(c, hipEventSynchronize(c->ev_tdone));
    } else {
        const hipError_t q = hipEventQuery(c->ev_tdone);
        if (q == hipErrorNotReady) return LBM_OK;
        if (q != hipSuccess) return fail(c, LBM_E_HIP, std::string("surrogate refresh: ") + hipGetErrorString(q));
    }
    train_release(c);
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_tdone, 0));
    HIP_TRY(c, sur::mlp_copy(c->net, c->tnet, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->sur_stats, c->stats_next, sizeof(float) * sur::SS_SIZE, hipMemcpyDeviceToDevice,
                              c->stream));
    c->adam_step += c->refresh_iters;
    c->refreshing = false;
    if (applied) *applied = 1;
    return LBM_OK;
}

}  // namespace

int lbm_surrogate_train(lbm_ctx* c, int iters, int batch, float lr, uint64_t seed, float* losses)
{
    if (!c) return LBM_E_ARG;
    if (iters < 0 || batch < 16 || batch % 16 || batch > (1 << 20) || !(lr > 0.f))
        return fail(c, LBM_E_ARG, "train needs iters >= 0, batch a multiple of 16 in [16, 2^20], lr > 0");
    if (int rc = sur_common(c)) return rc;
    if (int rc = refresh_finish(c, true, nullptr)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t rows = captured_rows(c);
    if (rows < 0) return fail(c, LBM_E_HIP, "reading the capture count");
    if (rows < 1) return fail(c, LBM_E_STATE, "no captured rows to train on (lbm_surrogate_capture, then step)");
    if (!c->norm_ready) {
        sur::capture_stats(c->cap_buf, rows, c->stats_part, c->sur_stats, c->stream);
        c->norm_ready = true;
    }
    if (int rc = train_buffers(c, iters, batch)) return rc;
    if (iters > 0) {
        if (int rc = train_enqueue(c, c->net, c->sur_stats, rows, iters, batch, lr, seed, c->stream)) return rc;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        train_release(c);
        c->adam_step += iters;
    }
    HIP_TRY(c, hipGetLastError());
    if (losses && iters > 0)
        HIP_TRY(c, hipMemcpyAsync(losses, c->bloss, sizeof(float) * iters, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_surrogate_refresh(lbm_ctx* c, int iters, int batch, float lr, uint64_t seed, float* losses)
{
    if (!c) return LBM_E_ARG;
    if (!c->norm_ready) return fail(c, LBM_E_STATE, "no trained network to refresh (lbm_surrogate_train first)");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    if (int rc = refresh_finish(c, true, nullptr)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t rows = captured_rows(c);
    if (rows < 1) return fail(c, LBM_E_STATE, "no captured rows");
    sur::capture_stats(c->cap_buf, rows, c->stats_part, c->sur_stats, c->stream, true);
    return lbm_surrogate_train(c, iters, batch, lr, seed, losses);
}

int lbm_surrogate_refresh_async(lbm_ctx* c, int iters, int batch, float lr, uint64_t seed)
{
    if (!c) return LBM_E_ARG;
    if (iters < 1 || batch < 16 || batch % 16 || batch > (1 << 20) || !(lr > 0.f))
        return fail(c, LBM_E_ARG, "refresh needs iters >= 1, batch a multiple of 16 in [16, 2^20], lr > 0");
    if (!c->norm_ready) return fail(c, LBM_E_STATE, "no trained network to refresh (lbm_surrogate_train first)");
    if (c->refreshing) return fail(c, LBM_E_STATE, "a refresh is in flight (lbm_surrogate_refresh_poll)");
    if (int rc = sur_common(c)) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int64_t rows = captured_rows(c);
    if (rows < 1) return fail(c, LBM_E_STATE, "no captured rows");
    if (int rc = train_buffers(c, iters, batch)) return rc;
    if (!c->tnet.Wf) {
        HIP_TRY(c, sur::mlp_alloc_empty(c->tnet));
        HIP_TRY(c, hipMalloc((void**)&c->stats_next, sizeof(float) * sur::SS_SIZE));
        HIP_TRY(c, hipStreamCreateWithFlags(&c->tstream, hipStreamNonBlocking));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_tstart, hipEventDisableTiming));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_tdone, hipEventDisableTiming));
    }
    // the training copy starts from the serving network and envelope; the
    // rows [0, rows) were written before this call, later steps append past them
    HIP_TRY(c, sur::mlp_copy(c->tnet, c->net, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->stats_next, c->sur_stats, sizeof(float) * sur::SS_SIZE, hipMemcpyDeviceToDevice,
                              c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_tstart, c->stream));
    HIP_TRY(c, hipStreamWaitEvent(c->tstream, c->ev_tstart, 0));
    sur::capture_stats(c->cap_buf, rows, c->stats_part, c->stats_next, c->tstream, true);
    if (int rc = train_enqueue(c, c->tnet, c->stats_next, rows, iters, batch, lr, seed, c->tstream)) return rc;
    HIP_TRY(c, hipEventRecord(c->ev_tdone, c->tstream));
    c->refresh_iters = iters;
    c->refreshing = true;
    return LBM_OK;
}

int lbm_surrogate_refresh_poll(lbm_ctx* c, int wait, int* applied)
{
    if (!c || !applied) return fail(c, LBM_E_ARG, "null argument");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    return refresh_finish(c, wait != 0, applied);
}

int lbm_surrogate_enable(lbm_ctx* c, int on, int audit_stride, double audit_tol)
{
    if (!c) return LBM_E_ARG;
    if (!on) {
        c->sur_on = false;
        return LBM_OK;
    }
    if (int rc = need_reacting(c)) return rc;
    if (audit_stride < 0 || !(audit_tol >= 0.0)) return fail(c, LBM_E_ARG, "audit_stride >= 0 and audit_tol >= 0");
    if (int rc = sur_common(c)) return rc;
    if (!c->norm_ready) return fail(c, LBM_E_STATE, "train the surrogate (or set its normalisation) first");
    const int64_t n = c->d.n;
    if (!c->chem_list) {
        c->sur_rows = (n + 255) / 256 * 256;
        HIP_TRY(c, hipMalloc((void**)&c->chem_list, sizeof(unsigned) * n));
        HIP_TRY(c, hipMalloc((void**)&c->chem_count, sizeof(unsigned)));
        HIP_TRY(c, hipMalloc((void**)&c->tchem, sizeof(double) * n));
        HIP_TRY(c, hipMalloc((void**)&c->sur_x, sizeof(uint16_t) * 16 * c->sur_rows));
        HIP_TRY(c, hipMalloc((void**)&c->sur_y, sizeof(float) * 16 * c->sur_rows));
        HIP_TRY(c, hipMalloc((void**)&c->sur_flag, n));
        HIP_TRY(c, hipMemsetAsync(c->sur_x, 0, sizeof(uint16_t) * 16 * c->sur_rows, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->bytes += (sizeof(unsigned) + sizeof(double) + 1) * n + (2 + 4) * 16 * c->sur_rows;
    }
    c->audit_stride = audit_stride;
    c->audit_tol = (float)audit_tol;
    c->sur_on = true;
    return LBM_OK;
}

int lbm_surrogate_read_stats(lbm_ctx* c, lbm_surrogate_stats* out)
{
    if (!c || !out) return fail(c, LBM_E_ARG, "null argument");
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *out = lbm_surrogate_stats{};
    unsigned long long have = 0;
    if (c->cap_count) HIP_TRY(c, hipMemcpy(&have, c->cap_count, sizeof(have), hipMemcpyDeviceToHost));
    out->captured = std::min<int64_t>((int64_t)have, c->cap_max);
    if (c->sur_count) {
        unsigned long long cnt[SUR_NCOUNT];
        HIP_TRY(c, hipMemcpy(cnt, c->sur_count, sizeof(cnt), hipMemcpyDeviceToHost));
        out->predicted = (int64_t)cnt[SUR_PREDICTED];
        out->fallback = (int64_t)cnt[SUR_FALLBACK];
        out->audited = (int64_t)cnt[SUR_AUDITED];
        out->audit_failures = (int64_t)cnt[SUR_AUDIT_FAIL];
        const unsigned bits = (unsigned)cnt[SUR_MAX_ERR];
        float e;
        std::memcpy(&e, &bits, 4);
        out->max_audit_error = e;
    }
    return LBM_OK;
}

int lbm_surrogate_get_params(lbm_ctx* c, float* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = sur_common(c)) return rc;
    return copy_params(c, host, true);
}

int lbm_surrogate_set_params(lbm_ctx* c, const float* host)
{
    if (!c || !host) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = refresh_finish(c, true, nullptr)) return rc;  // no refresh in flight
    if (int rc = sur_common(c)) return rc;
    if (int rc = copy_params(c, const_cast<float*>(host), false)) return rc;
    for (int L = 0; L < sur::NL; ++L) {
        const size_t nw = (size_t)sur::DIMS[L] * sur::DIMS[L + 1], nb = sur::DIMS[L + 1];
        HIP_TRY(c, hipMemsetAsync(c->net.mW[L], 0, 4 * nw, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->net.vW[L], 0, 4 * nw, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->net.mb[L], 0, 4 * nb, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->net.vb[L], 0, 4 * nb, c->stream));
    }
    c->adam_step = 0;
    sur::refresh_bf16(c->net, c->stream);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_surrogate_get_norm(lbm_ctx* c, float* host96)
{
    if (!c || !host96) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = sur_common(c)) return rc;
    if (!c->norm_ready) return fail(c, LBM_E_STATE, "no normalisation record yet");
    HIP_TRY(c, hipMemcpyAsync(host96, c->sur_stats, sizeof(float) * sur::SS_SIZE, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_surrogate_set_norm(lbm_ctx* c, const float* host96)
{
    if (!c || !host96) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = refresh_finish(c, true, nullptr)) return rc;  // no refresh in flight
    if (int rc = sur_common(c)) return rc;
    for (int q = 0; q < 16; ++q)
        if (!(host96[sur::SS_FSTD + q] > 0.f) || !(host96[sur::SS_TSTD + q] >= 0.f))
            return fail(c, LBM_E_ARG, "feature scales must be > 0 and target scales >= 0");
    HIP_TRY(c, hipMemcpyAsync(c->sur_stats, host96, sizeof(float) * sur::SS_SIZE, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->norm_ready = true;
    return LBM_OK;
}

int lbm_surrogate_forward(lbm_ctx* c, const float* X, int64_t m, float* Y)
{
    if (!c || !X || !Y) return fail(c, LBM_E_ARG, "null argument");
    if (m < 1 || m > (1 << 24)) return fail(c, LBM_E_ARG, "m out of range [1, 2^24]");
    if (int rc = sur_common(c)) return rc;
    float *dx = nullptr, *dy = nullptr;
    const size_t bytes = sizeof(float) * 16 * (size_t)m;
    HIP_TRY(c, hipMalloc((void**)&dx, bytes));
    hipError_t e = hipMalloc((void**)&dy, bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, X, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        sur::mlp_infer(c->net, dx, (int)m, dy, c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(Y, dy, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return fail(c, LBM_E_HIP, std::string("surrogate forward: ") + hipGetErrorString(e));
    return LBM_OK;
}

}  // extern "C"

namespace {

// Allocation-free network pass: the device staging of its host rows lives in
// `scratch`, SUR_NOALLOC_ROWS input rows then as many output rows.  Each
// device holds its own copy; its address is resolved once per device with the
// device current and kept in noalloc_ptr (host code never passes the symbol
// itself).  Concurrent calls on one device share the array and race.
constexpr int64_t SUR_NOALLOC_ROWS = 65536;
static __device__ __align__(256) float scratch[2 * 16 * SUR_NOALLOC_ROWS];
constexpr int MAX_DEVICES = 64;
std::mutex noalloc_mutex;
float* noalloc_ptr[MAX_DEVICES] = {};

int noalloc_scratch(lbm_ctx* c, float** out)
{
    const int dev = c->cfg.device;
    if (dev < 0 || dev >= MAX_DEVICES) return fail(c, LBM_E_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        HIP_TRY(c, hipSetDevice(dev));
        float* p = nullptr;
        HIP_TRY(c, hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_surrogate_forward_noalloc(lbm_ctx* c, const float* X, int64_t m, float* Y)
{
    if (!c || !X || !Y) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = sur_supported(c)) return rc;
    if (m < 1 || m > SUR_NOALLOC_ROWS)
        return fail(c, LBM_E_ARG, "m out of range [1, " + std::to_string(SUR_NOALLOC_ROWS) +
                                      "] (the allocation-free pass stages its rows in a fixed device array)");
    if (!c->net_ready)
        return fail(c, LBM_E_STATE, "no network yet: lbm_surrogate_init, _set_params or _train allocates it");
    float* base = nullptr;
    if (int rc = noalloc_scratch(c, &base)) return rc;
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    float* dx = base;
    float* dy = base + 16 * SUR_NOALLOC_ROWS;
    const size_t bytes = sizeof(float) * 16 * (size_t)m;
    HIP_TRY(c, hipMemcpyAsync(dx, X, bytes, hipMemcpyHostToDevice, c->stream));
    sur::mlp_infer(c->net, dx, (int)m, dy, c->stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(Y, dy, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_surrogate_gradient(lbm_ctx* c, const float* X, const float* T, int64_t m, float* grad, float* loss)
{
    if (!c || !X || !T || !grad || !loss) return fail(c, LBM_E_ARG, "null argument");
    if (int rc = refresh_finish(c, true, nullptr)) return rc;  // shares the training buffers
    if (m < 16 || m % 16 || m > (1 << 20)) return fail(c, LBM_E_ARG, "m must be a multiple of 16 in [16, 2^20]");
    if (int rc = sur_common(c)) return rc;
    if (int rc = sur_work(c, (int)m)) return rc;
    if (int rc = sur_batch(c, (int)m)) return rc;
    if (c->lcap < 1) {
        HIP_TRY(c, hipMalloc((void**)&c->bloss, sizeof(float)));
        c->lcap = 1;
    }
    float* xf = nullptr;
    const size_t bytes = sizeof(float) * 16 * (size_t)m;
    HIP_TRY(c, hipMalloc((void**)&xf, bytes));
    hipError_t e = hipMemcpyAsync(xf, X, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->bt, T, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        sur::to_bf16(xf, c->bx, 16 * m, c->stream);
        sur::mlp_grad(c->net, c->bx, c->bt, (int)m, c->work, c->stream);
        sur::loss_sum(c->work.loss_rows, (int)m, 1.f / ((float)m * sur::NOUT), c->bloss, c->stream);
        e = hipGetLastError();
    }
    size_t off = 0;
    for (int L = 0; L < sur::NL && e == hipSuccess; ++L) {
        const size_t nw = (size_t)sur::DIMS[L] * sur::DIMS[L + 1], nb = sur::DIMS[L + 1];
        e = hipMemcpyAsync(grad + off, c->work.gW[L], 4 * nw, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(grad + off + nw, c->work.gb[L], 4 * nb, hipMemcpyDeviceToHost, c->stream);
        off += nw + nb;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(loss, c->bloss, sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(xf);
    if (e != hipSuccess) return fail(c, LBM_E_HIP, std::string("surrogate gradient: ") + hipGetErrorString(e));
    return LBM_OK;
}

}  // extern "C"

// ------------------------------------------------------------ output files
// OutputVTK (output.cpp:9-210), write_restart (restart_file.cpp:274-350) and
// read_restart (restart_file.cpp:352-440) over the device state: the fields
// and populations come down through the same copies as lbm_get_* (FP32
// storage converted back to FP64 with its rest-state shifts), the files keep
// the reference's byte layout.
namespace {

const char* const* const kSpeciesNames = mech::SPECIES_NAMES;

// Units_Conv of the context; the single-component reference keeps the
// default m = kg = s = K = 1 (units.hpp:10).
chem::Units io_units(const lbm_ctx* c) { return multi(c) ? c->units : chem::Units{1.f, 1.f, 1.f, 1.f}; }

// si_u and si_HRR of units.hpp:119,129 (float arguments; sq/cb in double)
float si_u(const chem::Units& u, double x) { return (float)x * u.m / u.s; }
float si_hrr(const chem::Units& u, double x)
{
    const double m = u.m, s = u.s;
    return (float)((double)(float)x * ((double)u.kg * (m * m) / (s * s * s)));
}

struct Download {
    std::vector<uint8_t> type;
    std::vector<double> fld[LBM_FIELD_COUNT];
    std::vector<double> sp[5];  // X, rho, u, v, w: [NS][n]
};

int download(lbm_ctx* c, Download& h, bool species)
{
    const int64_t n = c->d.n;
    h.type.resize(n);
    HIP_TRY(c, hipSetDevice(c->cfg.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(h.type.data(), c->type, n, hipMemcpyDeviceToHost));
    for (int f = 0; f < LBM_FIELD_COUNT; ++f) {
        h.fld[f].resize(n);
        if (int rc = lbm_get_field(c, f, h.fld[f].data())) return rc;
    }
    if (species)
        for (int f = 0; f < 5; ++f) {
            h.sp[f].resize((size_t)n * c->ns);
            if (int rc = lbm_get_species_field(c, f, h.sp[f].data())) return rc;
        }
    return LBM_OK;
}

struct File {
    FILE* fp = nullptr;
    ~File()
    {
        if (fp) fclose(fp);
    }
};

template <typename T>
void put(std::vector<char>& b, const T& x)
{
    const char* p = reinterpret_cast<const char*>(&x);
    b.insert(b.end(), p, p + sizeof(T));
}

// MIXTURE / SPECIES records (lbm.hpp:56-102): byte offsets of the x86-64
// layout, one record per cell in (i, j, k) order, k fastest.
constexpr size_t MIX1_SIZE = 1024;  // single component: type, f, fpc, p_tensor, rho u v w p, g, gpc, rhoe, q, temp
constexpr size_t MIXM_SIZE = 600;   // multicomponent:   type, p_tensor, rho u v w p, g, gpc, rhoe, q, temp, HRR
constexpr size_t SPEC_SIZE = 472;   // X, rho, f, fpc, u, v, w

}  // namespace

int lbm_write_vtk(lbm_ctx* c, const char* path, int flags)
{
    if (!c || !path) return fail(c, LBM_E_ARG, "null argument");
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede output");
    const bool mc = multi(c);
    Download h;
    if (int rc = download(c, h, mc)) return rc;
    const chem::Units U = io_units(c);
    const bool si = flags & LBM_VTK_SI, massfrac = flags & LBM_VTK_MASS_FRACTIONS;
    const int Nx = c->d.nx, Ny = c->d.ny, Nz = c->d.nz, dx = 1;
    const int64_t n = c->d.n;
    File f;
    f.fp = fopen(path, "wb");
    if (!f.fp) return fail(c, LBM_E_ARG, std::string("cannot open ") + path);
    FILE* fp = f.fp;
    const long N = (long)Nx * Ny * Nz;
    long offset = 0;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
    fprintf(fp, "  <RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", Nx, Ny, Nz);
    fprintf(fp, "  <Piece Extent=\"0 %d 0 %d 0 %d\">\n", Nx, Ny, Nz);
    fprintf(fp, "    <PointData>\n");
    fprintf(fp, "    </PointData>\n");
    fprintf(fp, "    <CellData>\n");
    auto arr = [&](const char* name, int ncomp) {
        if (ncomp == 1)
            fprintf(fp, "      <DataArray type=\"Float32\" Name=\"%s\" format=\"appended\" offset=\"%ld\" />\n", name, offset);
        else
            fprintf(fp,
                    "      <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
                    "offset=\"%ld\" />\n",
                    name, ncomp, offset);
        offset += 4 + (long)ncomp * 4 * N;
    };
    arr("Density", 1);
    arr("Velocity", 3);
    arr("CellType", 1);
    arr("Temperature", 1);
    arr("Pressure", 1);
    if (mc) {
        arr("HRR", 1);
        for (int a = 0; a < c->ns; ++a) arr((std::string("Mole Fraction ") + kSpeciesNames[a]).c_str(), 1);
        for (int a = 0; a < c->ns; ++a) arr((std::string("Species Velocity ") + kSpeciesNames[a]).c_str(), 1);
    }
    fprintf(fp, "    </CellData>\n");
    fprintf(fp, "    <Coordinates>\n");
    fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateX\" format=\"appended\" offset=\"%ld\" />\n", offset);
    offset += 4 + 4L * (Nx + 1);
    fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateY\" format=\"appended\" offset=\"%ld\" />\n", offset);
    offset += 4 + 4L * (Ny + 1);
    fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateZ\" format=\"appended\" offset=\"%ld\" />\n", offset);
    offset += 4 + 4L * (Nz + 1);
    fprintf(fp, "    </Coordinates>\n");
    fprintf(fp, "  </Piece>\n");
    fprintf(fp, "  </RectilinearGrid>\n");
    fprintf(fp, "  <AppendedData encoding=\"raw\">");
    fprintf(fp, "_");
    // arrays in (k, j, i) order, i fastest
    std::vector<float> buf;
    auto block = [&](int ncomp, auto value) {
        const uint32_t size = (uint32_t)(ncomp * 4 * N);
        fwrite(&size, sizeof(uint32_t), 1, fp);
        buf.clear();
        for (int k = 0; k < Nz; ++k)
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    const int64_t cc = ((int64_t)i * Ny + j) * Nz + k;
                    for (int q = 0; q < ncomp; ++q) buf.push_back(value(cc, q));
                }
        fwrite(buf.data(), sizeof(float), buf.size(), fp);
    };
    const auto& F = h.fld;
    block(1, [&](int64_t cc, int) { return si ? U.si_rho(F[LBM_FIELD_RHO][cc]) : (float)F[LBM_FIELD_RHO][cc]; });
    block(3, [&](int64_t cc, int q) {
        const double x = F[LBM_FIELD_U + q][cc];
        return si ? si_u(U, x) : (float)x;
    });
    // the cell type as a float in a 4-byte slot (output.cpp:106-115)
    block(1, [&](int64_t cc, int) { return (float)(int)h.type[cc]; });
    block(1, [&](int64_t cc, int) { return si ? U.si_temp(F[LBM_FIELD_T][cc]) : (float)F[LBM_FIELD_T][cc]; });
    block(1, [&](int64_t cc, int) { return si ? U.si_p(F[LBM_FIELD_P][cc]) : (float)F[LBM_FIELD_P][cc]; });
    if (mc) {
        block(1, [&](int64_t cc, int) { return si ? si_hrr(U, F[LBM_FIELD_HRR][cc]) : (float)F[LBM_FIELD_HRR][cc]; });
        for (int a = 0; a < c->ns; ++a)
            block(1, [&](int64_t cc, int) {
                return massfrac ? (float)(h.sp[LBM_SP_RHO][a * n + cc] / F[LBM_FIELD_RHO][cc])
                                : (float)h.sp[LBM_SP_X][a * n + cc];
            });
        // x component of the species velocity relative to the mixture (output.cpp:174-189)
        for (int a = 0; a < c->ns; ++a)
            block(1, [&](int64_t cc, int) {
                const double ua = h.sp[LBM_SP_U][a * n + cc], um = F[LBM_FIELD_U][cc];
                return si ? si_u(U, (float)(ua - um)) : (float)((double)(float)ua - um);
            });
    }
    for (int ax = 0; ax < 3; ++ax) {
        const int L = ax == 0 ? Nx : (ax == 1 ? Ny : Nz);
        const uint32_t size = (uint32_t)(4 * (L + 1));
        fwrite(&size, sizeof(uint32_t), 1, fp);
        for (int i = 0; i < L + 1; ++i) {
            const float x = (float)(i * dx);
            fwrite(&x, sizeof(float), 1, fp);
        }
    }
    fprintf(fp, "  </AppendedData>\n");
    fprintf(fp, "</VTKFile>\n");
    if (ferror(fp)) return fail(c, LBM_E_ARG, std::string("write error on ") + path);
    return LBM_OK;
}

int lbm_write_restart(lbm_ctx* c, const char* path, int step)
{
    if (!c || !path) return fail(c, LBM_E_ARG, "null argument");
    if (!c->types_set) return fail(c, LBM_E_STATE, "lbm_set_types must precede output");
    const bool mc = multi(c);
    Download h;
    if (int rc = download(c, h, mc)) return rc;
    const int64_t n = c->d.n;
    const int ns = c->ns;
    std::vector<double> f((size_t)n * Q * (mc ? ns : 1)), fpc(f.size()), g((size_t)n * Q), gpc(g.size());
    if (int rc = lbm_get_populations(c, LBM_POP_F, f.data())) return rc;
    if (int rc = lbm_get_populations(c, LBM_POP_FPC, fpc.data())) return rc;
    if (int rc = lbm_get_populations(c, LBM_POP_G, g.data())) return rc;
    if (int rc = lbm_get_populations(c, LBM_POP_GPC, gpc.data())) return rc;
    File fl;
    fl.fp = fopen(path, "wb");
    if (!fl.fp) return fail(c, LBM_E_ARG, std::string("cannot open ") + path);
    std::vector<char> b;
    const int one = 1;
    put(b, one);  // dx
    put(b, one);  // dy
    put(b, one);  // dz
    put(b, c->d.nx);
    put(b, c->d.ny);
    put(b, c->d.nz);
    put(b, one);  // dt_sim
    put(b, step);
    if (!mc) {
        const double Ra = 1.0;  // LBM::Ra default; not used on this path
        put(b, c->cfg.nu);
        put(b, c->cfg.gas_const);
        put(b, c->cfg.gamma);
        put(b, Ra);
        put(b, c->cfg.prtl);
    } else {
        const size_t nsp = (size_t)ns;
        put(b, nsp);
        for (int a = 0; a < ns; ++a) {
            const size_t len = strlen(kSpeciesNames[a]);
            put(b, len);
            b.insert(b.end(), kSpeciesNames[a], kSpeciesNames[a] + len);
        }
        const double permeability = 999999;  // LBM::permeability default
        put(b, permeability);
    }
    fwrite(b.data(), 1, b.size(), fl.fp);
    const auto& F = h.fld;
    std::vector<char> rec(mc ? MIXM_SIZE + ns * SPEC_SIZE : MIX1_SIZE);
    auto wr = [&](size_t off, double x) { memcpy(rec.data() + off, &x, 8); };
    for (int64_t cc = 0; cc < n; ++cc) {
        std::fill(rec.begin(), rec.end(), 0);
        const short t = (short)h.type[cc];
        memcpy(rec.data(), &t, sizeof(short));
        size_t o = 8;
        if (!mc) {
            for (int l = 0; l < Q; ++l) wr(o + 8 * l, f[(size_t)l * n + cc]);
            o += 8 * Q;
            for (int l = 0; l < Q; ++l) wr(o + 8 * l, fpc[(size_t)l * n + cc]);
            o += 8 * Q;
        }
        // p_tensor[3][3], symmetric
        const int pidx[9] = {LBM_FIELD_PXX, LBM_FIELD_PXY, LBM_FIELD_PXZ, LBM_FIELD_PXY, LBM_FIELD_PYY,
                             LBM_FIELD_PYZ, LBM_FIELD_PXZ, LBM_FIELD_PYZ, LBM_FIELD_PZZ};
        for (int q = 0; q < 9; ++q) wr(o + 8 * q, F[pidx[q]][cc]);
        o += 72;
        const int mfld[5] = {LBM_FIELD_RHO, LBM_FIELD_U, LBM_FIELD_V, LBM_FIELD_W, LBM_FIELD_P};
        for (int q = 0; q < 5; ++q) wr(o + 8 * q, F[mfld[q]][cc]);
        o += 40;
        for (int l = 0; l < Q; ++l) wr(o + 8 * l, g[(size_t)l * n + cc]);
        o += 8 * Q;
        for (int l = 0; l < Q; ++l) wr(o + 8 * l, gpc[(size_t)l * n + cc]);
        o += 8 * Q;
        wr(o, F[LBM_FIELD_RHOE][cc]);
        wr(o + 8, F[LBM_FIELD_QX][cc]);
        wr(o + 16, F[LBM_FIELD_QY][cc]);
        wr(o + 24, F[LBM_FIELD_QZ][cc]);
        wr(o + 32, F[LBM_FIELD_T][cc]);
        o += 40;
        if (mc) {
            wr(o, F[LBM_FIELD_HRR][cc]);
            o += 8;
            for (int a = 0; a < ns; ++a) {
                const size_t an = (size_t)a * n + cc;
                wr(o, h.sp[LBM_SP_X][an]);
                wr(o + 8, h.sp[LBM_SP_RHO][an]);
                for (int l = 0; l < Q; ++l) wr(o + 16 + 8 * l, f[((size_t)a * Q + l) * n + cc]);
                for (int l = 0; l < Q; ++l) wr(o + 16 + 8 * Q + 8 * l, fpc[((size_t)a * Q + l) * n + cc]);
                wr(o + 16 + 16 * Q, h.sp[LBM_SP_U][an]);
                wr(o + 24 + 16 * Q, h.sp[LBM_SP_V][an]);
                wr(o + 32 + 16 * Q, h.sp[LBM_SP_W][an]);
                o += SPEC_SIZE;
            }
        }
        fwrite(rec.data(), 1, rec.size(), fl.fp);
    }
    if (ferror(fl.fp)) return fail(c, LBM_E_ARG, std::string("write error on ") + path);
    return LBM_OK;
}

int lbm_read_restart(lbm_ctx* c, const char* path, int* step_out)
{
    if (!c || !path) return fail(c, LBM_E_ARG, "null argument");
    c->omix_fresh = false;
    const bool mc = multi(c);
    File fl;
    fl.fp = fopen(path, "rb");
    if (!fl.fp) return fail(c, LBM_E_ARG, std::string("cannot open ") + path);
    auto rd = [&](void* p, size_t sz) { return fread(p, 1, sz, fl.fp) == sz; };
    int hdr[8];
    if (!rd(hdr, sizeof(hdr))) return fail(c, LBM_E_ARG, "truncated restart header");
    if (hdr[3] != c->d.nx || hdr[4] != c->d.ny || hdr[5] != c->d.nz)
        return fail(c, LBM_E_ARG, "restart grid " + std::to_string(hdr[3]) + "x" + std::to_string(hdr[4]) + "x" +
                                      std::to_string(hdr[5]) + " differs from the context's (ghosts included)");
    const int64_t n = c->d.n;
    const int ns = c->ns;
    if (!mc) {
        double k[5];
        if (!rd(k, sizeof(k))) return fail(c, LBM_E_ARG, "truncated restart header");
        if (k[0] != c->cfg.nu || k[1] != c->cfg.gas_const || k[2] != c->cfg.gamma || k[4] != c->cfg.prtl)
            return fail(c, LBM_E_ARG, "restart gas constants (nu, R, gamma, Pr) differ from the context's");
    } else {
        size_t nsp = 0;
        if (!rd(&nsp, sizeof(nsp)) || nsp != (size_t)ns) return fail(c, LBM_E_ARG, "restart species count differs");
        for (int a = 0; a < ns; ++a) {
            size_t len = 0;
            if (!rd(&len, sizeof(len)) || len > 64) return fail(c, LBM_E_ARG, "bad species name in restart");
            std::string name(len, '\0');
            if (!rd(&name[0], len) || name != kSpeciesNames[a])
                return fail(c, LBM_E_ARG, "restart species " + name + " differs from " + kSpeciesNames[a]);
        }
        double perm;
        if (!rd(&perm, sizeof(perm))) return fail(c, LBM_E_ARG, "truncated restart header");
    }
    std::vector<uint8_t> type(n);
    std::vector<double> fld[LBM_FIELD_COUNT];
    for (auto& v : fld) v.assign(n, 0.0);
    std::vector<double> f((size_t)n * Q * (mc ? ns : 1)), fpc(f.size()), g((size_t)n * Q), gpc(g.size());
    std::vector<double> sp[5];
    if (mc)
        for (auto& v : sp) v.assign((size_t)n * ns, 0.0);
    std::vector<char> rec(mc ? MIXM_SIZE + ns * SPEC_SIZE : MIX1_SIZE);
    auto get = [&](size_t off) {
        double x;
        memcpy(&x, rec.data() + off, 8);
        return x;
    };
    for (int64_t cc = 0; cc < n; ++cc) {
        if (!rd(rec.data(), rec.size())) return fail(c, LBM_E_ARG, "truncated restart records");
        short t;
        memcpy(&t, rec.data(), sizeof(short));
        if (t < 0 || t > TYPE_Q) return fail(c, LBM_E_ARG, "cell type out of range in restart");
        type[cc] = (uint8_t)t;
        size_t o = 8;
        if (!mc) {
            for (int l = 0; l < Q; ++l) f[(size_t)l * n + cc] = get(o + 8 * l);
            o += 8 * Q;
            for (int l = 0; l < Q; ++l) fpc[(size_t)l * n + cc] = get(o + 8 * l);
            o += 8 * Q;
        }
        fld[LBM_FIELD_PXX][cc] = get(o);
        fld[LBM_FIELD_PXY][cc] = get(o + 8);
        fld[LBM_FIELD_PXZ][cc] = get(o + 16);
        fld[LBM_FIELD_PYY][cc] = get(o + 32);
        fld[LBM_FIELD_PYZ][cc] = get(o + 40);
        fld[LBM_FIELD_PZZ][cc] = get(o + 64);
        o += 72;
        const int mfld[5] = {LBM_FIELD_RHO, LBM_FIELD_U, LBM_FIELD_V, LBM_FIELD_W, LBM_FIELD_P};
        for (int q = 0; q < 5; ++q) fld[mfld[q]][cc] = get(o + 8 * q);
        o += 40;
        for (int l = 0; l < Q; ++l) g[(size_t)l * n + cc] = get(o + 8 * l);
        o += 8 * Q;
        for (int l = 0; l < Q; ++l) gpc[(size_t)l * n + cc] = get(o + 8 * l);
        o += 8 * Q;
        fld[LBM_FIELD_RHOE][cc] = get(o);
        fld[LBM_FIELD_QX][cc] = get(o + 8);
        fld[LBM_FIELD_QY][cc] = get(o + 16);
        fld[LBM_FIELD_QZ][cc] = get(o + 24);
        fld[LBM_FIELD_T][cc] = get(o + 32);
        o += 40;
        if (mc) {
            fld[LBM_FIELD_HRR][cc] = get(o);
            o += 8;
            for (int a = 0; a < ns; ++a) {
                const size_t an = (size_t)a * n + cc;
                sp[LBM_SP_X][an] = get(o);
                sp[LBM_SP_RHO][an] = get(o + 8);
                for (int l = 0; l < Q; ++l) f[((size_t)a * Q + l) * n + cc] = get(o + 16 + 8 * l);
                for (int l = 0; l < Q; ++l) fpc[((size_t)a * Q + l) * n + cc] = get(o + 16 + 8 * Q + 8 * l);
                sp[LBM_SP_U][an] = get(o + 16 + 16 * Q);
                sp[LBM_SP_V][an] = get(o + 24 + 16 * Q);
                sp[LBM_SP_W][an] = get(o + 32 + 16 * Q);
                o += SPEC_SIZE;
            }
        }
    }
    if (int rc = lbm_set_types(c, type.data())) return rc;
    for (int q = 0; q < LBM_FIELD_COUNT; ++q)
        if (int rc = lbm_set_field(c, q, fld[q].data())) return rc;
    if (mc)
        for (int q = 0; q < 5; ++q)
            if (int rc = lbm_set_species_field(c, q, sp[q].data())) return rc;
    // the post-collision sets matter to the strict calls (Streaming reads
    // them); fused mode resumes from f, g and the moments
    if (!fused(c) || mc) {
        if (int rc = lbm_set_populations(c, LBM_POP_FPC, fpc.data())) return rc;
        if (int rc = lbm_set_populations(c, LBM_POP_GPC, gpc.data())) return rc;
    }
    if (int rc = lbm_set_populations(c, LBM_POP_F, f.data())) return rc;
    if (int rc = lbm_set_populations(c, LBM_POP_G, g.data())) return rc;
    if (step_out) *step_out = hdr[7];
    return LBM_OK;
}
