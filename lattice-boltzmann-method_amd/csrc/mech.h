// Selects the mechanism tables a build of the library is compiled for: one
// build per mechanism, as the reference binds its mechanism at compile time
// (lbm-main.cpp:83-88).  LBM_MECH_CH4_2S and LBM_MECH_C3H8 pick the two-step
// methane and propane mechanisms, LBM_MECH_NH3 the whole 31-species ammonia
// mechanism; the H2-air tables are the default.
#pragma once
#if defined(LBM_MECH_CH4_2S)
#include "mech_ch4_2s.h"
#elif defined(LBM_MECH_C3H8)
#include "mech_c3h8.h"
#elif defined(LBM_MECH_NH3)
#include "mech_nh3.h"
#else
#include "mech_h2air.h"
#endif
