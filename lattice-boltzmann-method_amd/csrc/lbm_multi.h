// Device view of the multicomponent state (internal).
#pragma once
#include "chem_device.h"
#include "lbm_internal.h"
#include "surrogate.h"

namespace lbm {

template <typename P>
struct MultiView {
    Dims d;
    const uint8_t* type;
    const uint8_t* nbflag;  // NB_SIMPLE: every neighbour is fluid or a periodic ghost
    P* f;   // species populations of the source buffer, [NS][27][n]
    P* fo;  // destination buffer
    P* g;   // mixture energy populations, [27][n]
    P* go;
    double* fld[LBM_FIELD_COUNT];
    double* hrr;   // [n] heat release rate
    // species planes, stored in P (binary32 in FP32 contexts: every kernel
    // computes in binary64 from the loaded values)
    P* X;     // [NS][n] mole fractions
    P* rhoa;  // [NS][n] species densities
    P* ua;    // [3][NS][n] species velocities
    chem::Units u;
    int reaction;
    unsigned long long* tau_count;
    const double* fref;  // FP32 storage: rest-state shifts [NS][27] of the species populations
    const double* gref;  // and [27] of the energy populations (unused for FP64)
    P* src;         // [NS][n] species mass added by the chemistry over the step (lattice units)
    P* tsp;         // [4][NS][n] Collide_Species transport of the step: omega_a, then the
                    // Maxwell-Stefan friction impulse per axis
    double* omix;   // FP32 storage: [7][n] omega and omega_1 of the mixture collide
                    // (written by the transport pass), then theta, U - 3/2 theta and
                    // the energy-flux correction dq[3] (written by the collide's
                    // first kernel); used when use_omix is set, else the collide
                    // evaluates the mixture transport itself in one kernel
    int use_omix;
    double* tchem;  // [n] temperature after the chemistry [K] (surrogate audit only)
    // chemistry surrogate (all null when unused)
    unsigned step;                 // step counter for the sampling hashes
    float* cap_buf;                // capture rows [cap_max][sur::CAP_STRIDE]
    unsigned long long* cap_count;
    long long cap_max;
    int cap_stride;
    const float* sur_stats;        // normalisation and envelope record [sur::SS_SIZE]
    unsigned long long* sur_count; // SUR_* counters
    int audit_stride;              // one 32-cell chunk in audit_stride is audited
    float audit_tol;               // audit tolerance in target standard deviations
    // fused mode: the cells without NB_PUSH keep their post-collision species
    // sets here, side slot s of cell side_cell[s] (side_slot[c] = s, else -1)
    const int* side_slot;  // [n]
    const int* side_cell;  // [nside]
    P* fside;              // [NS][27][nside]
    int64_t nside;
    // in-place mode: one species buffer (f == fo) streamed by the AA pattern.
    // swapped = 0: the incoming population l of cell x sits at slot (l, x);
    // swapped = 1: at slot (opp(l), x - e_l), the coordinates wrapped when
    // x - e_l is a periodic ghost.  A step reads one layout and writes the other.
    int swapped;
    // the step's list of NB_PUSH (cell, species) pairs at the SPECIES_MIN floor
    // and, per entry, the 27 post-collision values such a species keeps
    // (lbm-collision.cpp:512-513), saved before the collision overwrites them
    int* low_cell;           // [low_cap]
    uint8_t* low_sp;         // [low_cap]
    P* low_keep;             // [27][low_cap]
    unsigned* low_count;     // entries found (the host grows the list when it exceeds low_cap)
    int64_t low_cap;
    const P* pending;        // post-collision set [NS][27][n] uploaded for the first in-place step, or null
};

// counters of the surrogate path (cumulative)
enum { SUR_PREDICTED = 0, SUR_FALLBACK = 1, SUR_AUDITED = 2, SUR_AUDIT_FAIL = 3, SUR_MAX_ERR = 4, SUR_NCOUNT = 8 };

// bit 31 of a chemistry-list entry marks an audit (else a fallback)
constexpr unsigned CHEM_AUDIT = 0x80000000u;

// CollideSpeciesPush: the fused mode's Collide_Species, each post-collision
// species population written to the slot the next Streaming would move it to
// (buffer fo), so that only the energy populations are streamed
// (StreamEnergy); SideFill copies the post-collision species sets of the side
// cells out of fo after a strict step; Unstream rebuilds the post-collision
// species sets in fo from the streamed f.
// SpeciesTransport runs first in Collide_Species (per cell: omega_a and the
// Maxwell-Stefan friction), then CollideSpecies or CollideSpeciesPush.
enum class McKernel { Init, Chemistry, SpeciesTransport, CollideSpecies, Collide, Stream, Moments, CollideSpeciesPush,
                      StreamEnergy, SideFill, Unstream };

template <typename P>
void launch_mc(McKernel k, const MultiView<P>& v, hipStream_t s);

// In-place mode, Collide_Species in the layout v.swapped.  launch_mc_aa_save
// lists the NB_PUSH (cell, species) pairs at the SPECIES_MIN floor and saves
// the post-collision values they keep: from their streamed slots, or (first
// step after Init or an upload) from v.pending, else the rest state (stored
// 0).  *v.low_count is zeroed first on the stream.  launch_mc_aa_collide then
// collides every pair (the listed ones from their saved values, side cells
// into the side store) and writes each result where the other layout keeps it.
// Both cover the species a0 .. a0 + na - 1 (all NS, or one at a time when
// the list of every species' floor pairs would not fit in device memory);
// with `side`, launch_mc_aa_collide then streams the side cells.
template <typename P>
void launch_mc_aa_save(const MultiView<P>& v, bool first, int a0, int na, hipStream_t s);
template <typename P>
void launch_mc_aa_collide(const MultiView<P>& v, int a0, int na, bool side, hipStream_t s);
// The moment pass and TMS_BC read the species populations in the layout
// v.swapped; the energy populations stream as in the fused mode.

// SMOOTHING (lbm-smoothing.cpp:95-173) of every species, race-free, in the
// strict mode's layout; tmp holds [27][n] populations of the storage type.
template <typename P>
void launch_mc_smoothing(const MultiView<P>& v, void* tmp, double alpha, hipStream_t s);

// Max Mach number with each cell's mixture gamma into slot 5 of the
// diagnostics partials of launch_diag_partials (same block decomposition).
template <typename P>
void launch_mc_diag_mach(const MultiView<P>& v, double* partials, hipStream_t s);

// NSCBC sigma of the partially non-reflecting outflow (NSCBC_SIGMA_OUT, defines.hpp:100-102)
constexpr double NSCBC_SIGMA_OUT = 0.25;

// TMS_BC outflow reconstruction over `count` interface cells (cell index and
// the axis-aligned direction whose upstream neighbour is the TYPE_O_C ghost);
// `rec` holds mc_tms_record_fields() x count doubles of per-cell values.
template <typename P>
void launch_mc_tms(const MultiView<P>& v, const int* cells, const int* dirs, int count, double* rec, hipStream_t s);
int mc_tms_record_fields();

// TMS_BC inflow reconstruction for TYPE_I_C ghosts (same list layout; runs
// after the outflow sweep).
template <typename P>
void launch_mc_tms_in(const MultiView<P>& v, const int* cells, const int* dirs, int count, hipStream_t s);
// TMS_BC isothermal-wall sweep (TYPE_S) on the interface cells `cells`
template <typename P>
void launch_mc_tms_wall(const MultiView<P>& v, const int* cells, int count, hipStream_t s);

// Surrogate chemistry of every fluid cell: features (BF16 [rows][16] in X,
// envelope flags), network (BF16 MFMA, FP32 outputs [rows][16] in Y; rows >=
// n, a multiple of 256, rows past n zero), decoded sources into v.src /
// v.hrr / v.tchem; cells outside the training envelope and the audited
// chunks are appended to `list` (count in *count, zeroed first on the stream).
template <typename P>
void launch_mc_sur_chemistry(const MultiView<P>& v, const sur::Mlp& net, uint16_t* X, float* Y, uint8_t* flag,
                             int64_t rows, unsigned* list, unsigned* count, hipStream_t s);
// Plain chemistry of the listed cells: fallbacks are overwritten, audits are
// compared with the network's result and overwritten above the tolerance.
template <typename P>
void launch_mc_chem_list(const MultiView<P>& v, const unsigned* list, const unsigned* count, hipStream_t s);

// Per-block statistics (MC_SUM_BLOCKS x MC_STATS doubles into `part`): the
// sums of 1, T, p and X_a over TYPE_F cells (MC_SUMS values), then the minima
// of X_a over TYPE_F cells and inflow ghosts (NS values).
constexpr int MC_SUMS = 3 + chem::NS, MC_STATS = MC_SUMS + chem::NS, MC_SUM_BLOCKS = 512;
template <typename S>
void launch_mc_state_sums(const uint8_t* type, const double* T, const double* p, const S* X, int64_t n,
                          double* part, hipStream_t s);

}  // namespace lbm
