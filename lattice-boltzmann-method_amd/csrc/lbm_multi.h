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
};

enum class McKernel { Init, CollideSpecies, Collide, Stream, Moments };

template <typename P>
void launch_mc(McKernel k, const MultiView<P>& v, hipStream_t s);

// NSCBC sigma of the partially non-reflecting outflow (NSCBC_SIGMA_OUT, defines.hpp:100-102)
constexpr double NSCBC_SIGMA_OUT = 0.25;

// TMS_BC outflow reconstruction over `count` interface cells (cell index and
// the axis-aligned direction whose upstream neighbour is the TYPE_O_C ghost).
template <typename P>
void launch_mc_tms(const MultiView<P>& v, const int* cells, const int* dirs, int count, hipStream_t s);

}  // namespace lbm
