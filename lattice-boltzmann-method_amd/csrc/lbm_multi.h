// Device view of the multicomponent state (internal).
#pragma once
#include "chem_device.h"
#include "lbm_internal.h"

namespace lbm {

template <typename P>
struct MultiView {
    Dims d;
    const uint8_t* type;
    P* f;   // species populations of the source buffer, [NS][27][n]
    P* fo;  // destination buffer
    P* g;   // mixture energy populations, [27][n]
    P* go;
    double* fld[LBM_FIELD_COUNT];
    double* hrr;   // [n] heat release rate
    double* X;     // [NS][n] mole fractions
    double* rhoa;  // [NS][n] species densities
    double* ua;    // [3][NS][n] species velocities
    chem::Units u;
    int reaction;
    unsigned long long* tau_count;
    // chemistry surrogate (all null when unused)
    unsigned step;                 // step counter for sampling hashes
    float* cap_buf;                // capture rows [cap_max][32]: features 16 | targets 16
    unsigned long long* cap_count;
    long long cap_max;
    int cap_stride;
    const float* sur_pred;         // [n][16] normalised network outputs
    const uint8_t* sur_flag;       // [n] 1: outside the training envelope
    const float* sur_stats;        // feature mean/std [16]x2, target mean/std [16]x2, envelope lo/hi [16]x2
    unsigned long long* sur_count; // fallback, audited, audit failures, max audit error (float bits)
    int audit_stride;
    float audit_tol;
};

// layout of sur_stats
enum { SS_FMEAN = 0, SS_FSTD = 16, SS_TMEAN = 32, SS_TSTD = 48, SS_LO = 64, SS_HI = 80, SS_SIZE = 96 };

enum class McKernel { Init, CollideSpecies, Collide, Stream, Moments };

template <typename P>
void launch_mc(McKernel k, const MultiView<P>& v, hipStream_t s);

// NSCBC sigma of the partially non-reflecting outflow (NSCBC_SIGMA_OUT, defines.hpp:100-102)
constexpr double NSCBC_SIGMA_OUT = 0.25;

// TMS_BC outflow reconstruction over `count` interface cells (cell index and
// the axis-aligned direction whose upstream neighbour is the TYPE_O_C ghost).
template <typename P>
void launch_mc_tms(const MultiView<P>& v, const int* cells, const int* dirs, int count, hipStream_t s);

// Normalised surrogate features [n][16] and envelope flags of every fluid cell.
template <typename P>
void launch_mc_sur_features(const MultiView<P>& v, float* X, uint8_t* flag, hipStream_t s);

}  // namespace lbm
