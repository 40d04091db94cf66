/*
 * lbm_amd: C ABI of the MI355X (gfx950) D3Q27 compressible thermal LBM hot path.
 *
 * The reference has no plugin or FFI layer; its seam is the member calls that
 * LBM::loop makes every time step (lbm-main.cpp:153-193):
 *
 *     Collide_Species(); Collide(); Streaming(); TMS_BC(); calculate_moment();
 *
 * plus Init() from LBM::run (lbm-main.cpp:126-147) and the host accesses to
 * lb.mixture[i][j][k] by set-ups and outputs.  Each entry point below replaces
 * one of those, with the state living in HBM inside an lbm_ctx.
 *
 * Conventions
 *   - Every call returns 0 on success, a negative LBM_E_* code otherwise;
 *     lbm_last_error(ctx) gives the text.  No C++ exception crosses the ABI.
 *   - Host arrays are borrowed for the duration of the call only.  They use
 *     the reference's ghost-inclusive grid (nx+2, ny+2, nz+2) in (i, j, k)
 *     order with k fastest (lbm-main.cpp:16-35); populations are [27][ncell].
 *   - Calls are stream-ordered on the context's HIP stream and asynchronous
 *     until lbm_sync or a download.  One host thread per context.
 */
#ifndef LBM_AMD_H
#define LBM_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

typedef struct lbm_ctx lbm_ctx;

enum lbm_status {
    LBM_OK = 0,
    LBM_E_ARG = -1,      /* bad argument */
    LBM_E_HIP = -2,      /* HIP runtime error (no device, allocation, launch) */
    LBM_E_STATE = -3,    /* call not valid in the context's current state */
    LBM_E_UNSUPPORTED = -4
};

/* Population storage and collision arithmetic. */
enum lbm_precision {
    LBM_FP64 = 0,          /* plain method: FP64 populations, FP64 collision */
    LBM_FP32_SHIFTED = 1   /* FP32 deviations f - f_ref from the rest state */
};

/* How the five step-calls map onto kernels.  Single component, fused: the
 * pull-stream runs inside the collision and lbm_stream is a no-op (periodic,
 * TYPE_A and TYPE_FS neighbours only).  With species, fused: Collide_Species
 * writes each species population to the slot Streaming would move it to,
 * lbm_stream swaps the species buffers and streams only the energy
 * populations; every state is bitwise that of the strict calls (TYPE_FS and
 * TYPE_Q neighbours need LBM_MODE_STRICT), and the post-collision species set
 * (LBM_POP_FPC) is rebuilt when it is read. */
enum lbm_mode {
    LBM_MODE_FUSED = 0,
    LBM_MODE_STRICT = 1,  /* one kernel per reference call, explicit f and fpc */
    /* With species only: one species population buffer streamed in place by
     * the AA pattern (Bailey et al. 2009), half the species memory of the
     * other modes (the 768x384x384 layer fits one 288 GB GPU).  Collide_Species
     * writes each post-collision population where the next step reads it;
     * every state is bitwise that of the strict calls in a fresh context (the
     * first step after lbm_init takes the post-collision value a low species
     * keeps as the rest state, as a fresh strict context holds it).  Between
     * lbm_collide_species and lbm_stream the pre-collision species set is gone
     * (reading LBM_POP_F or uploading populations there is LBM_E_STATE);
     * TYPE_FS and TYPE_Q neighbours need LBM_MODE_STRICT. */
    LBM_MODE_INPLACE = 2
};

/* Host-visible per-cell fields (MIXTURE members, lbm.hpp:56-88). */
enum lbm_field {
    LBM_FIELD_RHO = 0, LBM_FIELD_U, LBM_FIELD_V, LBM_FIELD_W, LBM_FIELD_P, LBM_FIELD_T, LBM_FIELD_RHOE,
    LBM_FIELD_PXX, LBM_FIELD_PYY, LBM_FIELD_PZZ, LBM_FIELD_PXY, LBM_FIELD_PXZ, LBM_FIELD_PYZ,
    LBM_FIELD_QX, LBM_FIELD_QY, LBM_FIELD_QZ,   /* energy_flux */
    LBM_FIELD_HRR,                              /* heat release rate (multicomponent) */
    LBM_FIELD_COUNT
};

/* Population sets (MIXTURE::f, fpc, g, gpc; with species, f and fpc are the
 * per-species SPECIES::f, fpc). */
enum lbm_pop { LBM_POP_F = 0, LBM_POP_FPC = 1, LBM_POP_G = 2, LBM_POP_GPC = 3 };

/* Per-species fields (SPECIES members, lbm.hpp:90-102). */
enum lbm_species_field { LBM_SP_X = 0, LBM_SP_RHO = 1, LBM_SP_U = 2, LBM_SP_V = 3, LBM_SP_W = 4 };

typedef struct {
    int nx, ny, nz;        /* interior cells; ghosts are added as in LBM::LBM (lbm-main.cpp:19-21) */
    int n_species;         /* 0: single-component path; else the species count of the build's
                              mechanism (lbm_mechanism): 9 for H2-air (H2 H O O2 OH H2O HO2
                              H2O2 N2), 6 for the two-step methane and propane mechanisms */
    int precision;         /* enum lbm_precision */
    int mode;              /* enum lbm_mode */
    int device;            /* HIP device ordinal */
    /* single-component gas (LBM setters, lbm.hpp:264-271) */
    double nu, gas_const, gamma, prtl;
    /* rest state the FP32 deviations are taken from (ignored for FP64) */
    double ref_rho, ref_T;
    /* multicomponent only: units.set_m_kg_s(dx, dt) (units.hpp:32-44), and
     * whether Collide_Species integrates the chemistry source (REACTION) */
    double dx_si, dt_si;
    int reaction;
} lbm_config;

typedef struct {
    double mass;         /* sum rho over TYPE_F cells */
    double energy;       /* sum rhoe */
    double kinetic;      /* sum rho |u|^2 (no 1/2, as output.cpp:232-233) */
    double enstrophy;    /* output.cpp:235-262 */
    double max_mach;     /* max |u| / sqrt(gamma p / rho): gamma of the config (1.4 if unset) for a
                            single component; with species each cell's frozen mixture gamma
                            c_p / (c_p - R / W) at its T and composition */
    int64_t nan_count;   /* TYPE_F cells with non-finite rho or T */
    int64_t tau_warnings;/* collisions with 1/omega <= 0.5 so far (lbm-collision.cpp:33-34 prints instead) */
    int64_t fallback_count; /* cell updates whose chemistry the plain integrator redid
                               (surrogate envelope fallbacks + failed audits) */
} lbm_diag;

/* LBM::LBM (lbm-main.cpp:16-35): allocates the device state. */
int lbm_create(const lbm_config* cfg, lbm_ctx** out);
/* LBM::~LBM (lbm-main.cpp:121-123). */
void lbm_destroy(lbm_ctx* ctx);
const char* lbm_last_error(const lbm_ctx* ctx);
/* Text of the last lbm_create failure in this thread. */
const char* lbm_create_error(void);

/* The chemical mechanism this build of the library is compiled for, in place
 * of the reference's compile-time Cantera::newSolution choice
 * (lbm-main.cpp:83-88): its name ("h2air", "ch4_2s" or "c3h8"), with the
 * species count in *n_species when that pointer is not null.  Each mechanism
 * is its own library (liblbm_amd.so, liblbm_amd_ch4_2s.so, liblbm_amd_c3h8.so);
 * the single-component path is the same in all of them. */
const char* lbm_mechanism(int* n_species);
/* Name of species k in the order of every per-species array of this API
 * (gas->speciesName(k), setup.cpp); null when k is out of range. */
const char* lbm_species_name(int k);

/* The conjugate heat-transfer variant (#define CONJUGATE, defines.hpp:62-63;
 * lbm-collision.cpp:101-180, lbm-streaming.cpp:120-165, lbm-moment.cpp:85-120,
 * lbm-bc.cpp:1230-1346): interior TYPE_S cells conduct heat with ten times the
 * gas conductivity, stream g among themselves and couple to the gas through
 * TMS_BC.  Single-component LBM_MODE_STRICT contexts only (LBM_E_UNSUPPORTED
 * otherwise); 0 turns it off.  TYPE_S cells on the ghost layer stay isothermal
 * walls. */
int lbm_set_conjugate(lbm_ctx* ctx, int on);
/* The SMOOTHING variant (defines.hpp; LBM::Smoothing, lbm-smoothing.cpp:95-173,
 * called by Init_smooting and calculate_moment_smoothing, lbm.hpp:154-174,
 * lbm-main.cpp:133-134,178-179) as a run-time switch: lbm_init runs Init,
 * Smoothing, calculate_moment and lbm_moments runs calculate_moment,
 * Smoothing, calculate_moment.  Smoothing moves every species population of
 * the TYPE_F and TYPE_I cells by alpha (the reference's 0.1) times its
 * discrete Laplacian (fd_laplace, FD.cpp:40-50).  The reference updates in
 * place while reading neighbours (order-dependent); this is the race-free
 * form, every Laplacian from the populations before the call.  Multicomponent
 * LBM_MODE_STRICT contexts only (LBM_E_UNSUPPORTED otherwise); holds one
 * extra [27][n] population buffer while on. */
int lbm_set_smoothing(lbm_ctx* ctx, int on, double alpha);

/* Cell types, uint8 per ghost-inclusive cell (MIXTURE::type, defines.hpp:136-145). */
int lbm_set_types(lbm_ctx* ctx, const uint8_t* types);
/* Set-up / output access to per-cell fields (lb.mixture[i][j][k].<field>). */
int lbm_set_field(lbm_ctx* ctx, int field, const double* host);
int lbm_get_field(lbm_ctx* ctx, int field, double* host);
/* Population access, [27][ncell] doubles (restart and parity checks).  With
 * species, LBM_POP_F / LBM_POP_FPC move all species at once, [n_species][27][ncell]. */
int lbm_set_populations(lbm_ctx* ctx, int which, const double* host);
int lbm_get_populations(lbm_ctx* ctx, int which, double* host);
/* Per-species fields, [n_species][ncell] doubles (lb.species[a][i][j][k].<field>). */
int lbm_set_species_field(lbm_ctx* ctx, int field, const double* host);
int lbm_get_species_field(lbm_ctx* ctx, int field, double* host);

/* LBM::Init (single component lbm-init.cpp:8-82, multicomponent :85-146). */
int lbm_init(lbm_ctx* ctx);

/* The step-calls of LBM::loop (lbm-main.cpp:157-181). */
int lbm_collide_species(lbm_ctx* ctx); /* LBM::Collide_Species (lbm-collision.cpp:334-526); no-op without species */
int lbm_collide(lbm_ctx* ctx);         /* LBM::Collide (lbm-collision.cpp:9-188) */
int lbm_stream(lbm_ctx* ctx);          /* LBM::Streaming (lbm-streaming.cpp:3-171) */
int lbm_tms_bc(lbm_ctx* ctx);          /* LBM::TMS_BC (lbm-bc.cpp:122-1350) */
int lbm_moments(lbm_ctx* ctx);         /* LBM::calculate_moment (lbm-moment.cpp:11-125) */
/* nsteps iterations of the loop body above. */
int lbm_step(lbm_ctx* ctx, int nsteps);

/* Conservation and stability diagnostics (OutputKeEns, output.cpp:212-276, plus
 * mass, max Mach and counters), reduced on the device in FP64. */
int lbm_diagnostics(lbm_ctx* ctx, lbm_diag* out);

int lbm_sync(lbm_ctx* ctx);
/* The hipStream_t the context enqueues on (as void*). */
void* lbm_get_stream(lbm_ctx* ctx);

/* Kernel timing with HIP events on the context's stream.  While enabled every
 * kernel launch is bracketed by events; lbm_profile_read returns the summed
 * milliseconds and launch count of kernels whose name contains `name`. */
int lbm_profile_enable(lbm_ctx* ctx, int on);
int lbm_profile_read(lbm_ctx* ctx, const char* name, double* total_ms, int64_t* launches);

/* ---------------------------------------------------------------- files
 * OutputVTK (output.cpp:9-210): the current state as a VTK RectilinearGrid
 * (.vtr, raw appended Float32 arrays in (k, j, i) order): Density,
 * Velocity, CellType, Temperature, Pressure, and with species HRR, one mole
 * (or mass) fraction and one relative x velocity per species.  flags:
 * LBM_VTK_SI converts through Units_Conv (OUTPUT_SI), LBM_VTK_MASS_FRACTIONS
 * writes mass fractions (OUTPUT_MASSFRAC); the reference selects both at
 * compile time. */
enum { LBM_VTK_SI = 1, LBM_VTK_MASS_FRACTIONS = 2 };
int lbm_write_vtk(lbm_ctx* ctx, const char* path, int flags);
/* write_restart (restart_file.cpp:274-350): the reference's restart file
 * (int header dx dy dz Nx Ny Nz dt_sim step; the gas constants, or the
 * species names and permeability; one MIXTURE record per cell, followed by
 * its SPECIES records, in the x86-64 layout of lbm.hpp:56-102). */
int lbm_write_restart(lbm_ctx* ctx, const char* path, int step);
/* read_restart (restart_file.cpp:352-440) into this context: the grid,
 * species and gas constants must match; restores cell types, fields and
 * populations and returns the step in *step. */
int lbm_read_restart(lbm_ctx* ctx, const char* path, int* step);

/* Bytes of device memory the context holds. */
int64_t lbm_device_bytes(const lbm_ctx* ctx);

/*
 * Learned chemistry surrogate (multicomponent contexts with reaction on).
 *
 * The reference has no counterpart: it integrates the per-cell chemistry
 * source of Collide_Species (lbm-collision.cpp:370-399) every step.  These
 * calls let the per-cell MLP (T, p, Y_k, dt) -> (dT, dY_k) stand in for that
 * integration: 12 inputs (zero-padded to 16) -> 128 -> 128 -> 128 -> 10
 * outputs (padded to 16), GELU (tanh form) hidden layers, BF16 weights with FP32
 * accumulation on the matrix cores, Adam on FP32 master weights.
 *
 * Workflow: lbm_surrogate_capture while stepping with the plain integrator
 * samples (state, result) rows; lbm_surrogate_train fits the network on them
 * on the GPU; lbm_surrogate_enable switches Collide_Species to the network.
 * Cells whose features leave the captured min/max envelope, or whose network
 * output is not finite, fall back to the plain integrator in the same step;
 * one 32-cell chunk in `audit_stride` is recomputed by the plain integrator
 * and the plain result replaces the network's where they differ by more than
 * `audit_tol` target standard deviations.
 *
 * Parameter layout (lbm_surrogate_get/set_params, lbm_surrogate_gradient):
 * for each layer L = 0..3, W_L [N_L][K_L] row-major then b_L [N_L], with
 * (K, N) = (16, 128), (128, 128), (128, 128), (128, 16).
 * Normalisation record (lbm_surrogate_get/set_norm), 96 floats: feature mean
 * [16], feature scale [16], target mean [16], target scale [16], envelope
 * min [16], envelope max [16].  Raw features: T [K], p [Pa], Y_k (9), dt [s];
 * targets: dT [K] (from the exact initial temperature) and dY_k (9) over one
 * step.  A target scale of 0 marks an output that is not learned (it decodes
 * to its mean); the predicted sources' net mass is removed in proportion to
 * the target scales.
 */
typedef struct {
    int64_t captured;        /* rows in the capture buffer */
    int64_t predicted;       /* cell updates whose chemistry came from the network */
    int64_t fallback;        /* cell updates outside the envelope (plain integrator) */
    int64_t audited;         /* network results recomputed by the plain integrator */
    int64_t audit_failures;  /* audits above the tolerance (plain result kept) */
    double max_audit_error;  /* largest audit error, in target standard deviations */
} lbm_surrogate_stats;

/* Number of FP32 parameters (weights and biases, padded layout). */
int64_t lbm_surrogate_param_count(void);
/* (Re)initialises the network (Glorot-uniform from `seed`, zero biases and
 * Adam moments) and forgets the normalisation record. */
int lbm_surrogate_init(lbm_ctx* ctx, uint64_t seed);
/* Samples one cell in `stride` (hashed per step) of every plain-integrator
 * chemistry call into a capture buffer of `max_rows` rows.  stride = 0 stops
 * sampling and keeps the rows; max_rows = 0 releases the buffer.  Rows
 * already captured are kept while the buffer grows. */
int lbm_surrogate_capture(lbm_ctx* ctx, int64_t max_rows, int stride);
/* Copies up to max_rows captured rows [rows][32] (16 raw features | 16 raw
 * targets) to the host; *rows_out receives the count. */
int lbm_surrogate_get_capture(lbm_ctx* ctx, float* host, int64_t max_rows, int64_t* rows_out);
/* `iters` Adam steps (learning rate `lr`) on mini-batches of `batch` rows
 * (multiple of 16) drawn from the capture with `seed`; losses[iters] (may be
 * null) receives each batch's mean squared error in normalised units.  The
 * first call after lbm_surrogate_init derives the normalisation record from
 * the capture. */
int lbm_surrogate_train(lbm_ctx* ctx, int iters, int batch, float lr, uint64_t seed, float* losses);
/* Switches Collide_Species between the network (on != 0) and the plain integrator. */
int lbm_surrogate_enable(lbm_ctx* ctx, int on, int audit_stride, double audit_tol);
/* No counterpart in the reference.  Widens the training envelope to every
 * captured row (keeping the normalisation the network was trained with) and
 * trains `iters` more steps on all of them.  While lbm_surrogate_capture is on
 * with the surrogate in the step, the cells the plain integrator computes
 * (envelope fallbacks and audits) are captured, so a refresh teaches the
 * network the states it fell back on. */
int lbm_surrogate_refresh(lbm_ctx* ctx, int iters, int batch, float lr, uint64_t seed, float* losses);
/* The same refresh without stopping the steps: a copy of the network and
 * envelope is trained on a second stream from the rows captured so far, while
 * the steps keep using the serving network; lbm_surrogate_refresh_poll applies
 * the result (network and envelope together, ordered before the next step) once
 * the training has finished, *applied = 1 (with wait != 0 it waits for it).
 * One refresh at a time (LBM_E_STATE otherwise); calls that change the network,
 * the envelope or the capture buffer wait for and apply an unfinished one. */
int lbm_surrogate_refresh_async(lbm_ctx* ctx, int iters, int batch, float lr, uint64_t seed);
int lbm_surrogate_refresh_poll(lbm_ctx* ctx, int wait, int* applied);
int lbm_surrogate_read_stats(lbm_ctx* ctx, lbm_surrogate_stats* out);
int lbm_surrogate_get_params(lbm_ctx* ctx, float* host);
int lbm_surrogate_set_params(lbm_ctx* ctx, const float* host);
int lbm_surrogate_get_norm(lbm_ctx* ctx, float* host96);
int lbm_surrogate_set_norm(lbm_ctx* ctx, const float* host96);
/* Network outputs Y [m][16] for normalised inputs X [m][16] with the in-step
 * inference kernel. */
int lbm_surrogate_forward(lbm_ctx* ctx, const float* X, int64_t m, float* Y);
/* No counterpart in the reference.  The same network pass for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc): the rows are staged in a fixed per-device array of 65536 input
 * and 65536 output rows (8 MiB, held until the process ends) instead of
 * buffers allocated per call; m > 65536 is LBM_E_ARG, and a context without a
 * network yet (lbm_surrogate_init / _set_params / _train) is LBM_E_STATE.
 * Calls that overlap on one GPU share that array and race. */
int lbm_surrogate_forward_noalloc(lbm_ctx* ctx, const float* X, int64_t m, float* Y);
/* Mean squared error over the 10 outputs of m rows (m multiple of 16) and its
 * gradient in the parameter layout, from the training kernels. */
int lbm_surrogate_gradient(lbm_ctx* ctx, const float* X, const float* T, int64_t m, float* grad, float* loss);

#ifdef __cplusplus
}
#endif
#endif
