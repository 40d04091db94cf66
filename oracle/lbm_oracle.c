/*
 * CPU oracle, single-component path.  TEST INFRASTRUCTURE ONLY: the checker
 * of the HIP kernels, never part of the product (see lbm_oracle.h).
 */
#include "lbm_oracle.h"
#include <math.h>
#include <stddef.h>

/* D3Q27 velocity set, order and weights of lbm.hpp:49-53. */
static const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
static const int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
static const int OPP[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

static int cdir(int l, int d) { return d == 0 ? CX[l] : (d == 1 ? CY[l] : CZ[l]); }

/* One-axis factor of the product-form equilibrium, lbm.hpp:282-301:
 * Psi(0) = 1 - P, Psi(+-1) = (+-u + P)/2 with P = theta + u^2 + corr. */
static double psi1(int c, double u, double P)
{
    if (c == 0) return 1.0 - P;
    if (c == 1) return 0.5 * (u + P);
    return 0.5 * (P - u);
}

/* calculate_feq, lbm.hpp:276-305. */
double orc_feq(int l, double rho, const double vel[3], double theta, const double corr[3])
{
    double r = rho;
    for (int d = 0; d < 3; ++d) {
        double P = theta + vel[d] * vel[d] + corr[d];
        r *= psi1(cdir(l, d), vel[d], P);
    }
    return r;
}

/* calculate_geq(l, rho, U, theta, v), lbm-edf.cpp:38-98.
 *
 * The reference spells out 27 expanded polynomials.  They are exactly
 *   g/rho = (U - 3 theta/2) Px Py Pz + 1/2 sum_d (O_d^2 P_d) prod_{e!=d} P_e
 * with P_d = psi1(c_d, v_d, theta + v_d^2) and the operator
 * O_d = v_d + theta d/dv_d applied to P_d as a function of v_d.  Carried out:
 *   O^2 Psi(0)   = theta - 3 theta^2 + v^2 - 6 theta v^2 - v^4
 *   O^2 Psi(+-1) = (+-(v^3 + 3 theta v) + 3 theta^2 + 6 theta v^2 + v^4)/2
 * (the oracle tests check this against the reference's compiled polynomials
 * and against sum_l g = rho (U + |v|^2/2)). */
double orc_geq(int l, double rho, double U, double theta, const double vel[3])
{
    double Pf[3], Kf[3];
    for (int d = 0; d < 3; ++d) {
        int c = cdir(l, d);
        double v = vel[d], v2 = v * v;
        Pf[d] = psi1(c, v, theta + v2);
        if (c == 0)
            Kf[d] = theta - 3.0 * theta * theta + v2 - 6.0 * theta * v2 - v2 * v2;
        else
            Kf[d] = 0.5 * (c * (v2 * v + 3.0 * theta * v) + 3.0 * theta * theta + 6.0 * theta * v2 + v2 * v2);
    }
    double prod = Pf[0] * Pf[1] * Pf[2];
    double kin = Kf[0] * Pf[1] * Pf[2] + Pf[0] * Kf[1] * Pf[2] + Pf[0] * Pf[1] * Kf[2];
    return rho * ((U - 1.5 * theta) * prod + 0.5 * kin);
}

/* calculate_geq(l, rhoe, q, R, theta), lbm-edf.cpp:100-138 (Init only). */
double orc_geq_moment(int l, double rhoe, const double q[3], const double Rt[9], double theta)
{
    double c[3] = {CX[l], CY[l], CZ[l]};
    double g = rhoe + (q[0] * c[0] + q[1] * c[1] + q[2] * c[2]) / theta;
    double tr = 0.0;
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
            double A = Rt[3 * m + n] - (m == n ? rhoe * theta : 0.0);
            double B = c[m] * c[n] - (m == n ? theta : 0.0);
            tr += A * B;
        }
    g += tr / (2.0 * theta * theta);
    double wgt = 1.0;
    for (int m = 0; m < 3; ++m) wgt *= (c[m] == 0.0) ? (1.0 - theta) : 0.5 * theta;
    g *= wgt;
    double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    for (int m = 0; m < 3; ++m) {
        double B = (c2 == 0.0) ? 1.0 : (c2 == 1.0 ? -0.5 * fabs(c[m]) : 0.0);
        double Z = (1.0 - 3.0 * theta) / (2.0 * theta) * (Rt[4 * m] - theta * rhoe);
        g += B * Z;
    }
    return g;
}

/* calculate_gstr, lbm-edf.cpp:140-150: only the six axis directions shift. */
double orc_gstr(int l, double geq, const double dq[3])
{
    int c2 = CX[l] * CX[l] + CY[l] * CY[l] + CZ[l] * CZ[l];
    if (c2 != 1) return geq;
    return geq + 0.5 * (CX[l] * dq[0] + CY[l] * dq[1] + CZ[l] * dq[2]);
}

/* fd_central, FD.hpp:14-33: periodic ghosts count as walls (one-sided). */
static int is_wallish(int t) { return t == ORC_S || t == ORC_A || t == ORC_FS || t == ORC_P || t == ORC_O || t == ORC_OC; }
static double fd_central(double l, double c, double r, int tl, int tr)
{
    if (is_wallish(tl)) return tr == ORC_P ? 0.0 : (r - c);
    if (is_wallish(tr)) return tl == ORC_P ? 0.0 : (c - l);
    return 0.5 * (r - l);
}

static size_t cell(const orc_single* s, int i, int j, int k) { return ((size_t)i * s->ny + j) * s->nz + k; }
static size_t ncell(const orc_single* s) { return (size_t)s->nx * s->ny * s->nz; }

/* LBM::Init single-component, lbm-init.cpp:8-82.  f and g are built at ZERO
 * velocity while the heat flux uses the set-up velocity (reference quirk). */
void orc_init_single(orc_single* s)
{
    size_t N = ncell(s);
    double cv = s->R / (s->gamma - 1.0), cp = cv + s->R;
    for (size_t c = 0; c < N; ++c) {
        int t = s->type[c];
        if (!(t == ORC_F || t == ORC_S || t == ORC_I || t == ORC_O || t == ORC_Q)) continue;
        double vel0[3] = {0.0, 0.0, 0.0}, corr[3] = {0.0, 0.0, 0.0};
        double rho = s->p[c] / (s->R * s->T[c]);
        s->rho[c] = rho;
        double theta = s->R * s->T[c];
        for (int a = 0; a < 9; ++a) s->ptens[a * N + c] = (a % 4 == 0) ? s->p[c] : 0.0;
        double rhoe = rho * (cv * s->T[c]);
        s->rhoe[c] = rhoe;
        double H = cp * s->T[c];
        double Rt[9];
        for (int a = 0; a < 9; ++a) Rt[a] = (a % 4 == 0) ? H * s->p[c] : 0.0;
        double q[3] = {H * rho * s->u[c], H * rho * s->v[c], H * rho * s->w[c]};
        for (int l = 0; l < 27; ++l) {
            s->f[l * N + c] = orc_feq(l, rho, vel0, theta, corr);
            s->g[l * N + c] = orc_geq_moment(l, rhoe, q, Rt, theta);
        }
    }
}

/* Q_d = rho u_d (1 - 3 R T) - rho u_d^3, the flux whose divergence the
 * single-component Collide corrects for (lbm-collision.cpp:65-67). */
static double qdev(const orc_single* s, size_t c, int d)
{
    double ud = d == 0 ? s->u[c] : (d == 1 ? s->v[c] : s->w[c]);
    return s->rho[c] * ud * (1.0 - 3.0 * s->R * s->T[c]) - s->rho[c] * ud * ud * ud;
}

/* LBM::Collide single-component, lbm-collision.cpp:9-99. */
void orc_collide_single(orc_single* s)
{
    size_t N = ncell(s);
    double cv = s->R / (s->gamma - 1.0), cp = cv + s->R;
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                size_t c = cell(s, i, j, k);
                /* CONJUGATE solid cells collide with the same formula and a
                 * conductivity ten times the gas's (lbm-collision.cpp:101-180) */
                int solid = s->conjugate && s->type[c] == ORC_S;
                if (s->type[c] != ORC_F && !solid) continue;
                double rho = s->rho[c], p = s->p[c];
                double theta = s->R * s->T[c];
                double mu = s->nu * rho;
                double kappa = solid ? 10.0 * mu * cp / s->prtl : mu * cp / s->prtl;
                double om1 = 2.0 * p * cp / (p * cp + 2.0 * kappa);
                double om = 2.0 * p / (p + 2.0 * mu);
                if (1.0 / om <= 0.5 || 1.0 / om1 <= 0.5) s->tau_warnings++;
                double vel[3] = {s->u[c], s->v[c], s->w[c]};
                double Uint = s->rhoe[c] / rho - 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
                size_t nbm[3] = {cell(s, i - 1, j, k), cell(s, i, j - 1, k), cell(s, i, j, k - 1)};
                size_t nbp[3] = {cell(s, i + 1, j, k), cell(s, i, j + 1, k), cell(s, i, j, k + 1)};
                double dQ[3];
                for (int d = 0; d < 3; ++d)
                    dQ[d] = fd_central(qdev(s, nbm[d], d), qdev(s, c, d), qdev(s, nbp[d], d), s->type[nbm[d]],
                                       s->type[nbp[d]]);
                double dq[3], corr[3];
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        double eqP = (a == b ? p : 0.0) + rho * vel[a] * vel[b];
                        acc += vel[a] * (s->ptens[(3 * a + b) * N + c] - eqP);
                    }
                    dq[b] = acc + 0.5 * vel[b] * dQ[b];
                    corr[b] = (2.0 - om) / (2.0 * rho * om) * dQ[b];
                }
                for (int l = 0; l < 27; ++l) {
                    double feq = orc_feq(l, rho, vel, theta, corr);
                    s->fpc[l * N + c] = (1.0 - om) * s->f[l * N + c] + om * feq;
                    double geq = orc_geq(l, rho, Uint, theta, vel);
                    double gs = orc_gstr(l, geq, dq);
                    double g = s->g[l * N + c];
                    s->gpc[l * N + c] = g + om1 * (geq - g) + (om - om1) * (geq - gs);
                }
            }
}

static int dir_index(int cx, int cy, int cz)
{
    for (int a = 0; a < 27; ++a)
        if (CX[a] == cx && CY[a] == cy && CZ[a] == cz) return a;
    return -1;
}

/* LBM::dirSlip, lbm-bc.cpp:1352-1399. */
static void dir_slip(const orc_single* s, int l, int i, int j, int k, int* lp, int* ip, int* jp, int* kp)
{
    int in = i - CX[l], jn = j - CY[l], kn = k - CZ[l];
    int cx = CX[l], cy = CY[l], cz = CZ[l];
    *ip = in; *jp = jn; *kp = kn;
    if (s->type[cell(s, in, j, k)] == ORC_FS) { cx = -cx; *ip = i; }
    else if (s->type[cell(s, i, jn, k)] == ORC_FS) { cy = -cy; *jp = j; }
    else if (s->type[cell(s, i, j, kn)] == ORC_FS) { cz = -cz; *kp = k; }
    *ip = ((*ip - 1 + (s->nx - 2)) % (s->nx - 2)) + 1;
    *jp = ((*jp - 1 + (s->ny - 2)) % (s->ny - 2)) + 1;
    *kp = ((*kp - 1 + (s->nz - 2)) % (s->nz - 2)) + 1;
    *lp = dir_index(cx, cy, cz);
}

/* LBM::Streaming, lbm-streaming.cpp:3-119 (pull; walls, free slip, skipped
 * inflow/outflow sources, periodic wrap through the ghost index). */
void orc_stream_single(orc_single* s)
{
    size_t N = ncell(s);
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                size_t c = cell(s, i, j, k);
                if (s->conjugate && s->type[c] == ORC_S) {
                    /* CONJUGATE solid cells stream g only: nothing from gas
                     * cells (TMS_BC supplies it), bounce-back from adiabatic
                     * walls, a pull from solid and periodic cells through the
                     * ghost wrap (lbm-streaming.cpp:120-165) */
                    for (int l = 0; l < 27; ++l) {
                        int in = i - CX[l], jn = j - CY[l], kn = k - CZ[l];
                        int tn = s->type[cell(s, in, jn, kn)];
                        if (tn == ORC_F || tn == ORC_Q) {
                        } else if (tn == ORC_A) {
                            s->g[l * N + c] = s->gpc[OPP[l] * N + c];
                        } else if (tn == ORC_S || tn == ORC_P) {
                            if (in == 0) in = s->nx - 2; else if (in == s->nx - 1) in = 1;
                            if (jn == 0) jn = s->ny - 2; else if (jn == s->ny - 1) jn = 1;
                            if (kn == 0) kn = s->nz - 2; else if (kn == s->nz - 1) kn = 1;
                            s->g[l * N + c] = s->gpc[l * N + cell(s, in, jn, kn)];
                        }
                    }
                    continue;
                }
                if (s->type[c] != ORC_F) continue;
                for (int l = 0; l < 27; ++l) {
                    int in = i - CX[l], jn = j - CY[l], kn = k - CZ[l];
                    int tn = s->type[cell(s, in, jn, kn)];
                    if (tn == ORC_S) {
                        s->f[l * N + c] = s->fpc[OPP[l] * N + c];
                    } else if (tn == ORC_FS) {
                        int lp, ip, jp, kp;
                        dir_slip(s, l, i, j, k, &lp, &ip, &jp, &kp);
                        size_t src = cell(s, ip, jp, kp);
                        s->f[l * N + c] = s->fpc[lp * N + src];
                        s->g[l * N + c] = s->gpc[lp * N + src];
                    } else if (tn == ORC_A) {
                        s->f[l * N + c] = s->fpc[OPP[l] * N + c];
                        s->g[l * N + c] = s->gpc[OPP[l] * N + c];
                    } else if (tn == ORC_O || tn == ORC_OC || tn == ORC_I || tn == ORC_IC) {
                        /* source is an inflow/outflow ghost: left for TMS_BC */
                    } else {
                        if (in == 0) in = s->nx - 2; else if (in == s->nx - 1) in = 1;
                        if (jn == 0) jn = s->ny - 2; else if (jn == s->ny - 1) jn = 1;
                        if (kn == 0) kn = s->nz - 2; else if (kn == s->nz - 1) kn = 1;
                        size_t src = cell(s, in, jn, kn);
                        s->f[l * N + c] = s->fpc[l * N + src];
                        s->g[l * N + c] = s->gpc[l * N + src];
                    }
                }
            }
}

/* LBM::calculate_moment single-component, lbm-moment.cpp:11-85. */
void orc_moments_single(orc_single* s)
{
    size_t N = ncell(s);
    double cv = s->R / (s->gamma - 1.0);
    for (size_t c = 0; c < N; ++c) {
        if (s->conjugate && s->type[c] == ORC_S) {
            /* CONJUGATE solid cells: energy and its flux from g, then
             * T = (rho e / rho) / c_v and p = rho R T at the stored density
             * (lbm-moment.cpp:85-120) */
            double re = 0, qx = 0, qy = 0, qz = 0;
            for (int l = 0; l < 27; ++l) {
                double gl = s->g[l * N + c];
                re += gl;
                qx += gl * CX[l];
                qy += gl * CY[l];
                qz += gl * CZ[l];
            }
            s->rhoe[c] = re;
            s->eflux[c] = qx;
            s->eflux[N + c] = qy;
            s->eflux[2 * N + c] = qz;
            s->T[c] = s->rhoe[c] / s->rho[c] / cv;
            s->p[c] = s->rho[c] * s->R * s->T[c];
            continue;
        }
        if (s->type[c] != ORC_F) continue;
        double rho = 0, mx = 0, my = 0, mz = 0, re = 0, qx = 0, qy = 0, qz = 0;
        double P[9] = {0};
        for (int l = 0; l < 27; ++l) {
            double fl = s->f[l * N + c], gl = s->g[l * N + c];
            double e[3] = {CX[l], CY[l], CZ[l]};
            rho += fl;
            mx += fl * e[0];
            my += fl * e[1];
            mz += fl * e[2];
            for (int a = 0; a < 3; ++a)
                for (int b = a; b < 3; ++b) P[3 * a + b] += fl * e[a] * e[b];
            re += gl;
            qx += gl * e[0];
            qy += gl * e[1];
            qz += gl * e[2];
        }
        P[3] = P[1]; P[6] = P[2]; P[7] = P[5];
        for (int a = 0; a < 9; ++a) s->ptens[a * N + c] = P[a];
        s->rho[c] = rho;
        s->u[c] = mx / rho;
        s->v[c] = my / rho;
        s->w[c] = mz / rho;
        s->rhoe[c] = re;
        s->eflux[c] = qx;
        s->eflux[N + c] = qy;
        s->eflux[2 * N + c] = qz;
        double e_int = re / rho - 0.5 * (s->u[c] * s->u[c] + s->v[c] * s->v[c] + s->w[c] * s->w[c]);
        s->T[c] = e_int / cv;
        s->p[c] = rho * s->R * s->T[c];
    }
}

/* One TMS_BC reconstruction of cell c (lbm-bc.cpp:392-455, 869-935,
 * 1164-1225): the unknown directions (upstream cell of type `kind`) are
 * replaced by the boundary set {fb, gb} in the local moments, and every
 * direction is corrected by the equilibrium difference
 *   f <- f_tgt + f - f_loc (known),   f <- f_tgt + f_b - f_loc (unknown).
 * The outflow and solid sweeps use f_tgt = f_b; the inflow sweep uses the
 * equilibrium of the local density at the inflow velocity and temperature. */
static void tms_correct(orc_single* s, int i, int j, int k, int kind, const double fb[27], const double gb[27],
                        const double* ftgt, const double* gtgt)
{
    size_t N = ncell(s), c = cell(s, i, j, k);
    double cv = s->R / (s->gamma - 1.0);
    double rho = 0, mx = 0, my = 0, mz = 0, re = 0;
    for (int l = 0; l < 27; ++l) {
        int up = s->type[cell(s, i - CX[l], j - CY[l], k - CZ[l])] == kind;
        double a = up ? fb[l] : s->f[l * N + c];
        double b = up ? gb[l] : s->g[l * N + c];
        rho += a;
        mx += a * CX[l];
        my += a * CY[l];
        mz += a * CZ[l];
        re += b;
    }
    double vel[3] = {mx / rho, my / rho, mz / rho}, corr[3] = {0, 0, 0};
    double U = re / rho - 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    double T = U / cv, theta = s->R * T, Ucv = cv * T;
    for (int l = 0; l < 27; ++l) {
        double floc = orc_feq(l, rho, vel, theta, corr);
        double gloc = orc_geq(l, rho, Ucv, theta, vel);
        int up = s->type[cell(s, i - CX[l], j - CY[l], k - CZ[l])] == kind;
        double ft = ftgt ? ftgt[l] : fb[l], gt = gtgt ? gtgt[l] : gb[l];
        double fk = up ? fb[l] : s->f[l * N + c];
        double gk = up ? gb[l] : s->g[l * N + c];
        s->f[l * N + c] = ft + fk - floc;
        s->g[l * N + c] = gt + gk - gloc;
    }
}

/* calculate_feq_geq, lbm-edf.cpp:153-164. */
static void feq_geq(const orc_single* s, double rho, const double vel[3], double T, double f[27], double g[27])
{
    double cv = s->R / (s->gamma - 1.0), corr[3] = {0, 0, 0};
    for (int l = 0; l < 27; ++l) {
        f[l] = orc_feq(l, rho, vel, s->R * T, corr);
        g[l] = orc_geq(l, rho, cv * T, s->R * T, vel);
    }
}

/* The first direction whose upstream cell has type `kind` and whose
 * downstream cell is fluid (lbm-bc.cpp:357-368, 827-838), or -1. */
static int tms_first(const orc_single* s, int i, int j, int k, int kind)
{
    for (int l = 0; l < 27; ++l)
        if (s->type[cell(s, i - CX[l], j - CY[l], k - CZ[l])] == kind &&
            s->type[cell(s, i + CX[l], j + CY[l], k + CZ[l])] == ORC_F)
            return l;
    return -1;
}

/* LBM::TMS_BC single-component (lbm-bc.cpp:341-460 outflow, 809-938 inflow,
 * 1111-1230 isothermal wall), in the reference's sweep order. */
void orc_tms_single(orc_single* s)
{
    double fb[27], gb[27], ft[27], gt[27];
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                if (s->type[cell(s, i, j, k)] != ORC_F) continue;
                int l = tms_first(s, i, j, k, ORC_O);
                if (l < 0) continue;
                /* velocity and temperature downstream, pressure of the ghost */
                size_t dn = cell(s, i + CX[l], j + CY[l], k + CZ[l]), up = cell(s, i - CX[l], j - CY[l], k - CZ[l]);
                double vel[3] = {s->u[dn], s->v[dn], s->w[dn]}, T = s->T[dn];
                double rho = s->p[up] / (s->R * T);
                feq_geq(s, rho, vel, T, fb, gb);
                tms_correct(s, i, j, k, ORC_O, fb, gb, NULL, NULL);
            }
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                if (s->type[cell(s, i, j, k)] != ORC_F) continue;
                int l = tms_first(s, i, j, k, ORC_I);
                if (l < 0) continue;
                /* velocity and temperature of the ghost, pressure downstream */
                size_t dn = cell(s, i + CX[l], j + CY[l], k + CZ[l]), up = cell(s, i - CX[l], j - CY[l], k - CZ[l]);
                double vel[3] = {s->u[up], s->v[up], s->w[up]}, T = s->T[up];
                double rho = s->p[dn] / (s->R * T);
                feq_geq(s, rho, vel, T, fb, gb);
                /* target: the local density (the moments with f_in in the
                 * unknown directions) at the inflow velocity and temperature */
                size_t N = ncell(s), c = cell(s, i, j, k);
                double rl = 0;
                for (int m = 0; m < 27; ++m)
                    rl += s->type[cell(s, i - CX[m], j - CY[m], k - CZ[m])] == ORC_I ? fb[m] : s->f[m * N + c];
                feq_geq(s, rl, vel, T, ft, gt);
                tms_correct(s, i, j, k, ORC_I, fb, gb, ft, gt);
            }
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                size_t N = ncell(s), c = cell(s, i, j, k);
                if (s->type[c] != ORC_F) continue;
                int nd = 0;
                double vel[3] = {0, 0, 0}, T = 0, q = 0.5;
                for (int l = 0; l < 27; ++l) {
                    size_t up = cell(s, i - CX[l], j - CY[l], k - CZ[l]), dn = cell(s, i + CX[l], j + CY[l], k + CZ[l]);
                    if (s->type[up] != ORC_S || s->type[dn] != ORC_F) continue;
                    ++nd;
                    /* wall value interpolated at q = 1/2 between the cell
                     * downstream and the wall cell (lbm-bc.cpp:1152-1158) */
                    vel[0] += (q * s->u[dn] + s->u[up]) / (1.0 + q);
                    vel[1] += (q * s->v[dn] + s->v[up]) / (1.0 + q);
                    vel[2] += (q * s->w[dn] + s->w[up]) / (1.0 + q);
                    T += (q * s->T[dn] + s->T[up]) / (1.0 + q);
                }
                if (nd == 0) continue;
                double rbb = 0;
                for (int l = 0; l < 27; ++l) rbb += s->f[l * N + c];
                for (int d = 0; d < 3; ++d) vel[d] = 1.0 / nd * vel[d];
                T = 1.0 / nd * T;
                feq_geq(s, rbb, vel, T, fb, gb);
                tms_correct(s, i, j, k, ORC_S, fb, gb, NULL, NULL);
            }
    if (!s->conjugate) return;
    /* CONJUGATE (lbm-bc.cpp:1230-1346): solid cells with a gas neighbour
     * upstream and solid downstream rebuild g (f is left as it is) from the
     * target state at rest with the interface temperature at q = 1/2, the
     * gas directions replaced by the target set in the local moments. */
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                size_t N = ncell(s), c = cell(s, i, j, k);
                if (s->type[c] != ORC_S) continue;
                int nd = 0;
                double T = 0, q = 0.5, zero[3] = {0, 0, 0};
                for (int l = 0; l < 27; ++l) {
                    size_t up = cell(s, i - CX[l], j - CY[l], k - CZ[l]), dn = cell(s, i + CX[l], j + CY[l], k + CZ[l]);
                    if (!((s->type[up] == ORC_F || s->type[up] == ORC_Q) && s->type[dn] == ORC_S)) continue;
                    ++nd;
                    T += (q * s->T[dn] + s->T[up]) / (1.0 + q);
                }
                if (nd == 0) continue;
                double rbb = 0;
                for (int l = 0; l < 27; ++l) rbb += s->f[l * N + c];
                T = 1.0 / nd * T;
                feq_geq(s, rbb, zero, T, fb, gb);
                double cv = s->R / (s->gamma - 1.0);
                double rho = 0, mx = 0, my = 0, mz = 0, re = 0;
                for (int l = 0; l < 27; ++l) {
                    int ut = s->type[cell(s, i - CX[l], j - CY[l], k - CZ[l])];
                    int gas = ut == ORC_F || ut == ORC_Q;
                    double a = gas ? fb[l] : s->f[l * N + c];
                    rho += a;
                    mx += a * CX[l];
                    my += a * CY[l];
                    mz += a * CZ[l];
                    re += gas ? gb[l] : s->g[l * N + c];
                }
                double vl[3] = {mx / rho, my / rho, mz / rho};
                double Tl = (re / rho - 0.5 * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2])) / cv;
                feq_geq(s, rho, vl, Tl, ft, gt);
                for (int l = 0; l < 27; ++l) {
                    int ut = s->type[cell(s, i - CX[l], j - CY[l], k - CZ[l])];
                    if (ut == ORC_F || ut == ORC_Q)
                        s->g[l * N + c] = 2 * gb[l] - gt[l];
                    else
                        s->g[l * N + c] = gb[l] + s->g[l * N + c] - gt[l];
                }
            }
}

/* Kinetic energy and enstrophy as output.cpp:212-276 defines them, plus the
 * conserved sums. */
void orc_diag_single(const orc_single* s, double out[4])
{
    double ke = 0, ens = 0, mass = 0, en = 0;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = cell(s, i, j, k);
                if (s->type[c] != ORC_F) continue;
                ke += s->rho[c] * (s->u[c] * s->u[c] + s->v[c] * s->v[c] + s->w[c] * s->w[c]);
                int ia = i + 1, ib = i - 1, ja = j + 1, jb = j - 1, ka = k + 1, kb = k - 1;
                if (ib < 1) ib = nx - 2; else if (ia > nx - 2) ia = 1;
                if (jb < 1) jb = ny - 2; else if (ja > ny - 2) ja = 1;
                if (kb < 1) kb = nz - 2; else if (ka > nz - 2) ka = 1;
                double uy = 0.5 * (s->u[cell(s, i, ja, k)] - s->u[cell(s, i, jb, k)]);
                double uz = 0.5 * (s->u[cell(s, i, j, ka)] - s->u[cell(s, i, j, kb)]);
                double vx = 0.5 * (s->v[cell(s, ia, j, k)] - s->v[cell(s, ib, j, k)]);
                double vz = 0.5 * (s->v[cell(s, i, j, ka)] - s->v[cell(s, i, j, kb)]);
                double wx = 0.5 * (s->w[cell(s, ia, j, k)] - s->w[cell(s, ib, j, k)]);
                double wy = 0.5 * (s->w[cell(s, i, ja, k)] - s->w[cell(s, i, jb, k)]);
                ens += s->rho[c] * ((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
                mass += s->rho[c];
                en += s->rhoe[c];
            }
    out[0] = ke; out[1] = ens; out[2] = mass; out[3] = en;
}
