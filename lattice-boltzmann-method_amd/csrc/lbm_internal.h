// Internal device-state layout shared by the kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_amd.h"
#include "lbm_lattice.h"

namespace lbm {

// Ghost-inclusive grid; cell index c = (i*ny + j)*nz + k (k fastest).
struct Dims {
    int nx, ny, nz;
    int64_t n;
};

// XCD-aware workgroup order.  The dispatcher hands consecutive workgroups to
// the 8 XCDs round-robin, and each XCD has its own L2; remapping gives every
// XCD one contiguous range of tiles, so the rows that neighbouring tiles pull
// from meet in the same L2.  Launch with xcd_grid (a multiple of 8 workgroups).
constexpr unsigned N_XCD = 8;
__device__ inline unsigned xcd_block()
{
    const unsigned per = gridDim.x / N_XCD;
    return (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
}
inline unsigned xcd_grid(int64_t n, int block)
{
    const int64_t nb = (n + block - 1) / block;
    return (unsigned)((nb + N_XCD - 1) / N_XCD * N_XCD);
}
// Per-(cell, species) kernels: one workgroup per (cell block, species) pair,
// the species fastest in xcd_block() order, so that the `ns` workgroups of a
// cell block run on one XCD back to back and read the block's per-cell planes
// (temperature, velocity, types, transport terms) from that XCD's L2 after
// the first of them.  Workgroups past the last cell block get cb >= nb.
inline unsigned xcd_species_grid(int64_t n, int block, int ns)
{
    const int64_t nb = (n + block - 1) / block;
    return (unsigned)((nb * ns + N_XCD - 1) / N_XCD * N_XCD);
}
__device__ inline void xcd_species_block(int ns, int& cb, int& a)
{
    const unsigned L = xcd_block();
    cb = (int)(L / (unsigned)ns);
    a = (int)(L % (unsigned)ns);
}

// nbflag bits, computed once per lbm_set_types.
enum : uint8_t {
    NB_SIMPLE = 1,  // TYPE_F cell whose 26 neighbours are all TYPE_F or TYPE_P
    NB_KEEP = 2,    // TYPE_F cell with an S/O/O_C/I/I_C neighbour (stream keeps stale values)
    NB_PUSH = 4     // multicomponent fused mode: every population of the cell streams to a
                    // fluid cell that TMS_BC leaves alone, and it pulls only from such moves
};

// Device view of the single-component state; P is the population storage and
// collision type (double: plain FP64, float: FP32 deviations).
template <typename P>
struct SingleView {
    Dims d;
    const uint8_t* type;
    const uint8_t* nbflag;
    P* f;    // [27][n] populations of the source buffer
    P* g;
    P* fo;   // [27][n] destination buffer
    P* go;
    double* fld[LBM_FIELD_COUNT];
    P* qd[3];               // Q_d = rho u_d (1 - 3 R T) - rho u_d^3 of the last moment pass
    double R, cv, cp, nu, prtl, gamma;
    // Rest-state populations subtracted in storage (all 0 for FP64) and their
    // sums {sum f, sum cx^2 f, sum cy^2 f, sum cz^2 f, sum g}; the odd and
    // mixed moments of the rest state vanish.
    double fref[27], gref[27];
    double fref_sum[5];
    __host__ __device__ double fref_d(int l) const { return fref[l]; }
    __host__ __device__ double gref_d(int l) const { return gref[l]; }
    unsigned long long* tau_count;
    int conjugate;  // CONJUGATE: TYPE_S cells conduct heat (strict mode)
};

// Kernel launchers (lbm_single.hip).  All enqueue on `s`.
template <typename P> void launch_init(const SingleView<P>& v, hipStream_t s);
template <typename P> void launch_collide_explicit(const SingleView<P>& v, hipStream_t s);
template <typename P> void launch_stream(const SingleView<P>& v, hipStream_t s);
template <typename P> void launch_moments(const SingleView<P>& v, hipStream_t s);
template <typename P> void launch_fused_collide(const SingleView<P>& v, hipStream_t s);
template <typename P> void launch_fused_moments(const SingleView<P>& v, bool materialize, hipStream_t s);
// TMS_BC on the interface cells `cells` (fluid cells with a TYPE_O, TYPE_I or
// TYPE_S upstream neighbour whose downstream neighbour is fluid), in place on v.f/v.g.
template <typename P> void launch_tms(const SingleView<P>& v, const int* cells, int count, hipStream_t s);
// CONJUGATE part of TMS_BC on the solid interface cells `cells` (TYPE_S cells
// with a gas cell upstream and a solid cell downstream of some direction).
template <typename P> void launch_tms_conj(const SingleView<P>& v, const int* cells, int count, hipStream_t s);
// gamma: ratio of specific heats used for the Mach number |u| / sqrt(gamma p / rho)
void launch_diag(const Dims& d, const uint8_t* type, double* const* fld, double gamma, double* partials, int npart,
                 double* out8, hipStream_t s);
int diag_partials(const Dims& d);
// The two halves of launch_diag: per-block partials [npart][6] (sums of
// rho, rho e, rho |u|^2, rho |omega|^2, non-finite cells; max Mach), then the
// fixed-order final reduction into out8[0..5].  One block per 256 cells.
void launch_diag_partials(const Dims& d, const uint8_t* type, double* const* fld, double gamma, double* partials,
                          hipStream_t s);
void launch_diag_final(const double* partials, int npart, double* out8, hipStream_t s);

}  // namespace lbm
