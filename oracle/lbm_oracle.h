/*
 * CPU oracle of the D3Q27 compressible thermal LBM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * of the HIP path.  Nothing in the product links it or falls back to it.
 *
 * Plain C99, FP64, one function per reference routine, each citing the
 * reference file:line it restates.  Grid layout: ghost-inclusive dimensions
 * (nx, ny, nz) = interior + 2, cell index c = (i*ny + j)*nz + k (k fastest,
 * as the reference's contiguous storage, lbm-main.cpp:25-33); every field is
 * one array of ncell doubles; populations are population-major [27][ncell].
 *
 * Pinning (see tests/test_oracle_*.py):
 *   orc_feq, orc_geq, orc_gstr, orc_geq_moment  - vectors from the reference's
 *       own compiled calculate_feq / calculate_geq / calculate_gstr
 *       (oracle/_ref/ref_driver fn), plus moment identities
 *   orc_init/collide/stream/moments (single)    - whole-state snapshots of the
 *       compiled reference on a periodic Taylor-Green box (ref_driver tgv)
 *   orc_stream wall branches (S, A, FS)         - permutation/mirror identities
 *       on hand-built cases (parity unpinned against the reference binary)
 *   the CONJUGATE branches (solid cells)        - whole-state snapshots of the
 *       reference compiled with -DCONJUGATE (oracle/_ref/ref_driver_conj) on a
 *       fluid layer between two solid slabs
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F = 0, ORC_S = 1, ORC_P = 2, ORC_A = 3, ORC_FS = 4, ORC_I = 5, ORC_O = 6, ORC_IC = 7, ORC_OC = 8, ORC_Q = 9 };

typedef struct {
    int nx, ny, nz;            /* ghost-inclusive */
    double R, gamma, prtl, nu; /* gas constant, Cp/Cv, Prandtl, kinematic viscosity */
    uint8_t* type;             /* [ncell] */
    double *rho, *u, *v, *w, *p, *T, *rhoe; /* [ncell] */
    double* ptens;             /* [9][ncell], row-major 3x3 */
    double* eflux;             /* [3][ncell] */
    double *f, *fpc, *g, *gpc; /* [27][ncell] */
    int64_t tau_warnings;      /* cells with 1/omega <= 0.5 seen by collide */
    int conjugate;             /* CONJUGATE: TYPE_S cells conduct heat (defines.hpp:62-63) */
} orc_single;

double orc_feq(int l, double rho, const double vel[3], double theta, const double corr[3]);
double orc_geq(int l, double rho, double U, double theta, const double vel[3]);
double orc_geq_moment(int l, double rhoe, const double q[3], const double Rt[9], double theta);
double orc_gstr(int l, double geq, const double dq[3]);

void orc_init_single(orc_single* s);
void orc_collide_single(orc_single* s);
void orc_stream_single(orc_single* s);
void orc_moments_single(orc_single* s);
/* TMS_BC: outflow (TYPE_O), inflow (TYPE_I), isothermal wall (TYPE_S) */
void orc_tms_single(orc_single* s);
/* out = {sum rho|u|^2, enstrophy, sum rho, sum rhoe} over TYPE_F cells */
void orc_diag_single(const orc_single* s, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
