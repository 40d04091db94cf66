// Compiled CPU restatement of the multicomponent step of oracle/multicomp.py
// (MultiState.init, collide_species with the plain chemistry, collide,
// stream on a periodic box, moments), FP64, one OpenMP loop over the cells
// per call.
//
// TEST INFRASTRUCTURE ONLY: the compiled all-cores CPU baseline of the
// reacting mixing layer (the cpu_baseline leg of bench.py) and a second CPU
// statement of the oracle, checked against the numpy oracle by
// tests/test_oracle_mc_cpu.py.  The product never runs it.
//
// Thermo, kinetics, transport and the binary32 unit crossings are the
// host-callable functions of chem_device.h over the generated mechanism
// tables (the numpy oracle's chem.py reads the same tables from mech/*.json);
// the D3Q27 order and the fd_* stencil rules come from lbm_lattice.h.  Each
// function names the reference lines the numpy oracle restates.
//
// Usage: mc_cpu <case.bin> <steps> <threads> [out.bin]
//   case.bin: "MCCPU01\0", int32 nx ny nz (ghost-inclusive), int32 reaction,
//             int32 has_state, float64 dx dt, uint8 type[n], float64 rho u v w
//             p T [n] each, float64 X[NS][n]; with has_state, then float64
//             rhoe [n], ptens [6][n], rho_a [NS][n], u_a [3][NS][n],
//             f [NS][27][n], g [27][n]: the run starts from that state in
//             place of Init and calculate_moment
//   out.bin:  float64 rho u v w p T rhoe [n] each, rho_a [NS][n], u_a [3][NS][n],
//             f [NS][27][n], g [27][n]
// Prints one JSON line: cells, steps, threads, seconds, MLUPS.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "chem_device.h"
#include "lbm_lattice.h"

using namespace lbm;
using namespace lbm::chem;

namespace {

constexpr uint8_t T_F = TYPE_F, T_P = TYPE_P;

struct State {
    int nx, ny, nz;
    int64_t n;
    Units un;
    int reaction;
    std::vector<uint8_t> type;
    std::vector<double> rho, u, v, w, p, T, rhoe, hrr;
    std::vector<double> ptens;  // [6][n]: xx yy zz xy xz yz
    std::vector<double> g, gpc;        // [27][n]
    std::vector<double> X, rhoa;       // [NS][n]
    std::vector<double> ua;            // [3][NS][n]
    std::vector<double> f, fpc;        // [NS][27][n]
    std::vector<double> src;           // [NS][n] chemistry source of the step
    std::vector<double> om_a, Fa;      // [NS][n], [3][NS][n] species transport of the step
    std::vector<int> fluid;            // TYPE_F cells

    int64_t idx(int i, int j, int k) const { return ((int64_t)i * ny + j) * nz + k; }
    double* vel(int d) { return d == 0 ? u.data() : (d == 1 ? v.data() : w.data()); }
    int64_t nb(int64_t c, int d, int s) const
    {
        const int64_t st = d == 0 ? (int64_t)ny * nz : (d == 1 ? nz : 1);
        return c + s * st;
    }
};

// calculate_feq (lbm.hpp:276-305), product form
inline double feq(int l, double rho, const double vel[3], double theta, const double corr[3])
{
    double r = rho;
    const int cs[3] = {cx(l), cy(l), cz(l)};
    for (int d = 0; d < 3; ++d) {
        const double P = theta + vel[d] * vel[d] + corr[d];
        r = r * (cs[d] == 0 ? (1.0 - P) : (cs[d] == 1 ? (vel[d] + P) / 2 : (-vel[d] + P) / 2));
    }
    return r;
}

// calculate_geq(l, rho, U, theta, v) (lbm-edf.cpp:38-98), the operator form
// of the numpy oracle's geq
inline double geq(int l, double rho, double U, double theta, const double vel[3])
{
    double P[3], K[3];
    const int cs[3] = {cx(l), cy(l), cz(l)};
    for (int d = 0; d < 3; ++d) {
        const double vv = vel[d], v2 = vv * vv;
        const double even = 3 * theta * theta + 6 * theta * v2 + v2 * v2, odd = v2 * vv + 3 * theta * vv;
        if (cs[d] == 0) {
            P[d] = 1.0 - theta - v2;
            K[d] = theta - 3 * theta * theta + v2 - 6 * theta * v2 - v2 * v2;
        } else if (cs[d] == 1) {
            P[d] = 0.5 * (vv + theta + v2);
            K[d] = 0.5 * (even + odd);
        } else {
            P[d] = 0.5 * (theta + v2 - vv);
            K[d] = 0.5 * (even - odd);
        }
    }
    return rho * ((U - 1.5 * theta) * P[0] * P[1] * P[2] +
                  0.5 * (K[0] * P[1] * P[2] + P[0] * K[1] * P[2] + P[0] * P[1] * K[2]));
}

inline bool fluidlike(uint8_t t) { return t == TYPE_F || t == TYPE_I || t == TYPE_O || t == TYPE_I_C || t == TYPE_O_C; }

void normalized_Y(const State& s, int64_t c, double Y[NS])
{
    for (int a = 0; a < NS; ++a) Y[a] = s.rhoa[a * s.n + c] / s.rho[c];
    normalize(Y);
}

// Init (lbm-init.cpp:85-143)
void init(State& s)
{
    const int64_t n = s.n;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        if (!fluidlike(s.type[c])) continue;
        double Xc[NS], Y[NS], xs = 0;
        for (int a = 0; a < NS; ++a) xs += s.X[a * n + c];
        for (int a = 0; a < NS; ++a) Xc[a] = s.X[a * n + c] / xs;
        Y_from_X(Xc, Y);
        const double Tsi = s.un.si_temp(s.T[c]), psi = s.un.si_p(s.p[c]);
        const double rho_si = psi * mean_W(Y) / (R_UNIV * Tsi);
        const double rho = s.un.rho(rho_si);
        s.rho[c] = rho;
        const double vel[3] = {s.u[c], s.v[c], s.w[c]};
        const double Uint = s.un.energy_mass(u_mass(Tsi, Y));
        s.rhoe[c] = rho * (Uint + 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]));
        const double theta = s.un.energy_mass(R_UNIV * Tsi / mean_W(Y));
        for (int l = 0; l < Q; ++l) s.g[l * n + c] = geq(l, rho, Uint, theta, vel);
        const double zero[3] = {0, 0, 0};
        for (int a = 0; a < NS; ++a) {
            const double ra = std::max(Y[a] * rho, SPECIES_MIN);
            s.rhoa[a * n + c] = ra;
            for (int d = 0; d < 3; ++d) s.ua[(d * NS + a) * n + c] = vel[d];
            const double th_a = s.un.energy_mass(R_UNIV * Tsi / W[a]);
            for (int l = 0; l < Q; ++l) s.f[((int64_t)a * Q + l) * n + c] = feq(l, ra, vel, th_a, zero);
        }
    }
}

// Collide_Species (lbm-collision.cpp:334-526): the chemistry source (two
// explicit sub-steps with a UV re-solve, :370-399), omega_a from the species
// viscosities and Wilke weights (:405-429), the Maxwell-Stefan friction
// (:433-470), then the species collision with the delta-Q correction (:452-520).
void collide_species(State& s)
{
    const int64_t n = s.n;
    const int nf = (int)s.fluid.size();
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nf; ++q) {
        const int64_t c = s.fluid[q];
        const double rho = s.rho[c];
        double Y0[NS];
        normalized_Y(s, c, Y0);
        const double Tsi = s.un.si_temp(s.T[c]);
        const double rho_si = s.un.si_rho(rho);
        double ra[NS], src[NS] = {};
        for (int a = 0; a < NS; ++a) ra[a] = s.rhoa[a * n + c];
        if (s.reaction) {
            double Uint = s.rhoe[c] / rho;
            for (int a = 0; a < NS; ++a) {
                double k2 = 0;
                for (int d = 0; d < 3; ++d) k2 += s.ua[(d * NS + a) * n + c] * s.ua[(d * NS + a) * n + c];
                Uint = Uint - 0.5 * ra[a] / rho * k2;
            }
            const double usi = s.un.si_energy_mass(Uint);
            double Tcur = Tsi, Ycur[NS];
            for (int a = 0; a < NS; ++a) Ycur[a] = Y0[a];
            for (int sub = 0; sub < 2; ++sub) {
                double wdot[NS];
                net_production(Tcur, rho_si, Ycur, wdot);
                for (int a = 0; a < NS; ++a) {
                    src[a] += 1.0 / 2.0 * (double)s.un.rho_dot(wdot[a] * W[a]);
                    Ycur[a] = (ra[a] + src[a]) / rho;
                }
                normalize(Ycur);
                Tcur = T_from_UV(usi, Ycur, Tcur);
            }
            double hrr = 0;
            for (int a = 0; a < NS; ++a)
                hrr += -1.0 * (double)s.un.energy_mass(h_RT(a, Tcur) * R_UNIV * Tcur / W[a]) * src[a];
            s.hrr[c] = hrr;
        }
        for (int a = 0; a < NS; ++a) s.src[a * n + c] = src[a];
        // omega_a
        const double lnT = log(Tsi);
        double mu[NS];
        for (int a = 0; a < NS; ++a) mu[a] = species_viscosity(a, lnT);
        double phi[NS][NS];
        for (int a = 0; a < NS; ++a)
            for (int b = a; b < NS; ++b) {
                const double f1 = 1.0 + sqrt(mu[a] / mu[b]) * sqrt(sqrt(W[b] / W[a]));
                phi[a][b] = f1 * f1 / sqrt(8.0 * (1.0 + W[a] / W[b]));
                phi[b][a] = mu[b] / mu[a] * W[a] / W[b] * phi[a][b];
            }
        const double p = s.p[c];
        double Xa[NS];
        for (int a = 0; a < NS; ++a) Xa[a] = s.X[a * n + c];
        for (int a = 0; a < NS; ++a) {
            double den = 0;
            for (int b = 0; b < NS; ++b) den = den + Xa[b] * phi[a][b];
            const double visc = Xa[a] * (double)s.un.mu(mu[a]) / den;
            s.om_a[a * n + c] = 2 * Xa[a] * p / (Xa[a] * p + 2 * visc);
        }
        // Maxwell-Stefan friction at the state's pressure
        const double p_si = pressure(Tsi, rho_si, Y0);
        for (int a = 0; a < NS; ++a) {
            double acc[3] = {0, 0, 0};
            for (int b = 0; b < NS; ++b) {
                const double D = b == a ? 1.0 : (double)s.un.nu(binary_diff(a, b, lnT, p_si));
                const double k = b == a ? 0.0 : -1.0 * p * Xa[a] * Xa[b] / D;
                for (int d = 0; d < 3; ++d)
                    acc[d] = acc[d] + k * (s.ua[(d * NS + a) * n + c] - s.ua[(d * NS + b) * n + c]);
            }
            for (int d = 0; d < 3; ++d) s.Fa[(d * NS + a) * n + c] = acc[d] * 1.0 / ra[a];
        }
    }
    // the species collision reads omega_a of its own cell only; the delta-Q
    // stencil reads rho_a, u_a and T of the neighbours (fields of the state)
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nf; ++q) {
        const int64_t c = s.fluid[q];
        const double T = s.T[c];
        const double vmix[3] = {s.u[c], s.v[c], s.w[c]};
        for (int a = 0; a < NS; ++a) {
            const double ra = s.rhoa[a * n + c];
            const double* fa = &s.f[(int64_t)a * Q * n];
            double mom[3] = {0, 0, 0};
            for (int l = 0; l < Q; ++l) {
                const double fl = fa[l * n + c];
                mom[0] += cx(l) * fl;
                mom[1] += cy(l) * fl;
                mom[2] += cz(l) * fl;
            }
            const double uf[3] = {mom[0] / ra, mom[1] / ra, mom[2] / ra};
            const double Fa[3] = {s.Fa[(0 * NS + a) * n + c], s.Fa[(1 * NS + a) * n + c], s.Fa[(2 * NS + a) * n + c]};
            const double udu[3] = {uf[0] + Fa[0], uf[1] + Fa[1], uf[2] + Fa[2]};
            const double Ra = (double)s.un.cp(R_UNIV / W[a]);
            const double theta = Ra * T;
            const double om = s.om_a[a * n + c];
            double corr[3];
            for (int d = 0; d < 3; ++d) {
                auto Qf = [&](int64_t m) {
                    const double ud = s.ua[(d * NS + a) * n + m], r = s.rhoa[a * n + m];
                    return r * ud * (1.0 - 3.0 * Ra * s.T[m]) - r * (ud * ud * ud);
                };
                const int64_t m0 = s.nb(c, d, -1), m1 = s.nb(c, d, 1);
                const double dQ = fd_central<double>(Qf(m0), Qf(c), Qf(m1), s.type[m0], s.type[m1]);
                corr[d] = 1 * (2.0 - om) / (2.0 * ra * om) * dQ;
            }
            const double zero[3] = {0, 0, 0};
            const double sa = s.src[a * n + c];
            for (int l = 0; l < Q; ++l) {
                const double fe = feq(l, ra, uf, theta, corr);
                const double fdu = feq(l, ra, udu, theta, corr);
                const double fr = feq(l, sa, vmix, theta, zero);
                if (ra <= SPECIES_MIN && fr <= 0.0) continue;  // post-collision value left as it was
                const double fl = fa[l * n + c];
                s.fpc[((int64_t)a * Q + l) * n + c] = fl + om * (fe - fl) + (fdu - fe) + fr;
            }
        }
    }
}

// Collide (lbm-collision.cpp:191-332)
void collide(State& s)
{
    const int64_t n = s.n;
    const int nf = (int)s.fluid.size();
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nf; ++q) {
        const int64_t c = s.fluid[q];
        const double rho = s.rho[c];
        double Y0[NS], Xs[NS];
        normalized_Y(s, c, Y0);
        const double Tsi = s.un.si_temp(s.T[c]);
        X_from_Y(Y0, Xs);
        const double lnT = log(Tsi);
        const double cp = s.un.cp(cp_mass(Tsi, Y0));
        const double Rmix = s.un.cp(R_UNIV / mean_W(Y0));
        const double mu = s.un.mu(mix_viscosity(lnT, Xs));
        const double kap = s.un.thermal_conductivity(mix_conductivity(lnT, Xs));
        const double theta = s.un.energy_mass(R_UNIV * Tsi / mean_W(Y0));
        const double vel[3] = {s.u[c], s.v[c], s.w[c]};
        const double p = s.p[c];
        const double om1 = 2 * p * cp / (p * cp + 2 * kap);
        const double om = 2 * p / (p + 2 * mu);
        double Uint = s.rhoe[c] / rho;
        for (int a = 0; a < NS; ++a) {
            double k2 = 0;
            for (int d = 0; d < 3; ++d) k2 += s.ua[(d * NS + a) * n + c] * s.ua[(d * NS + a) * n + c];
            Uint = Uint - 0.5 * s.rhoa[a * n + c] / rho * k2;
        }
        double qdiff[3] = {0, 0, 0}, qcorr[3] = {0, 0, 0};
        for (int a = 0; a < NS; ++a) {
            const double ha = s.un.energy_mass(h_RT(a, Tsi) * R_UNIV * Tsi / W[a]);
            const double ra = s.rhoa[a * n + c];
            for (int d = 0; d < 3; ++d) {
                const double uad = s.ua[(d * NS + a) * n + c];
                qdiff[d] += om1 / (-om + om1) * ha * ra * (uad - vel[d]);
                const int64_t m0 = s.nb(c, d, -1), m1 = s.nb(c, d, 1);
                const double Yn = s.rhoa[a * n + m0] / s.rho[m0], Yp = s.rhoa[a * n + m1] / s.rho[m1];
                const double dY = fd_fuw<double>(Yn, ra / rho, Yp, uad - vel[d], s.type[m0], s.type[m1]);
                qcorr[d] += 0.5 * (2.0 - om1) / (-om + om1) * 1 * p * ha * dY;
            }
        }
        auto Rn = [&](int64_t m) -> double {
            if (!fluidlike(s.type[m])) return 0.0;
            double sYW = 0, sY = 0;
            for (int a = 0; a < NS; ++a) {
                const double Ym = s.rhoa[a * n + m] / s.rho[m];
                sYW += Ym / W[a];
                sY += Ym;
            }
            return s.un.cp(R_UNIV * sYW / sY);
        };
        double dQ[3];
        for (int d = 0; d < 3; ++d) {
            const double* vv = s.vel(d);
            auto Qf = [&](int64_t m, double R) {
                return s.rho[m] * vv[m] * (1.0 - 3.0 * R * s.T[m]) - s.rho[m] * (vv[m] * vv[m] * vv[m]);
            };
            const int64_t m0 = s.nb(c, d, -1), m1 = s.nb(c, d, 1);
            dQ[d] = fd_central<double>(Qf(m0, Rn(m0)), Qf(c, Rmix), Qf(m1, Rn(m1)), s.type[m0], s.type[m1]);
        }
        double dq[3];
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
            for (int a = 0; a < 3; ++a) {
                const int id = (a == b) ? a : (3 + a + b - 1);
                const double eqP = (a == b ? p : 0.0) + rho * vel[a] * vel[b];
                acc = acc + vel[a] * (s.ptens[id * n + c] - eqP);
            }
            dq[b] = acc + qdiff[b] + qcorr[b] + 0.5 * 1 * vel[b] * dQ[b];
        }
        for (int l = 0; l < Q; ++l) {
            const double ge = geq(l, rho, Uint, theta, vel);
            const double gs = c2(l) == 1 ? ge + 0.5 * (cx(l) * dq[0] + cy(l) * dq[1] + cz(l) * dq[2]) : ge;
            const double g = s.g[l * n + c];
            s.gpc[l * n + c] = g + om1 * (ge - g) + (om - om1) * (ge - gs);
        }
    }
}

// Streaming (lbm-streaming.cpp:3-119) on a box whose ghosts are all periodic
void stream(State& s)
{
    const int64_t n = s.n;
    const int nf = (int)s.fluid.size();
    auto wrap = [](int i, int m) { return i == 0 ? m - 2 : (i == m - 1 ? 1 : i); };
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nf; ++q) {
        const int64_t c = s.fluid[q];
        const int i = (int)(c / ((int64_t)s.ny * s.nz)), j = (int)((c / s.nz) % s.ny), k = (int)(c % s.nz);
        for (int l = 0; l < Q; ++l) {
            const int64_t src = s.idx(wrap(i - cx(l), s.nx), wrap(j - cy(l), s.ny), wrap(k - cz(l), s.nz));
            for (int a = 0; a < NS; ++a) s.f[((int64_t)a * Q + l) * n + c] = s.fpc[((int64_t)a * Q + l) * n + src];
            s.g[l * n + c] = s.gpc[l * n + src];
        }
    }
}

// calculate_moment (lbm-moment.cpp:193-377): moments, the SPECIES_MIN clamp,
// the Maxwell-Stefan velocities (Cholesky) with the previous state's D_ab, p
// and X, and the temperature by UV at the new composition.
void moments(State& s)
{
    const int64_t n = s.n;
    const int nf = (int)s.fluid.size();
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nf; ++q) {
        const int64_t c = s.fluid[q];
        double Yold[NS];
        normalized_Y(s, c, Yold);
        const double Tsi_old = s.un.si_temp(s.T[c]);
        const double rho_old_si = s.un.si_rho(s.rho[c]);
        double re = 0;
        for (int l = 0; l < Q; ++l) re += s.g[l * n + c];
        s.rhoe[c] = re;
        double ra[NS], mom[3][NS], Pt[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NS; ++a) {
            double r = 0, m0 = 0, m1 = 0, m2 = 0;
            for (int l = 0; l < Q; ++l) {
                const double f = s.f[((int64_t)a * Q + l) * n + c];
                r += f;
                m0 += cx(l) * f;
                m1 += cy(l) * f;
                m2 += cz(l) * f;
                Pt[0] += cx(l) * cx(l) * f;
                Pt[1] += cy(l) * cy(l) * f;
                Pt[2] += cz(l) * cz(l) * f;
                Pt[3] += cx(l) * cy(l) * f;
                Pt[4] += cx(l) * cz(l) * f;
                Pt[5] += cy(l) * cz(l) * f;
            }
            ra[a] = r;
            mom[0][a] = m0;
            mom[1][a] = m1;
            mom[2][a] = m2;
        }
        for (int id = 0; id < 6; ++id) s.ptens[id * n + c] = Pt[id];
        double rho = 0, mv[3] = {0, 0, 0};
        for (int a = 0; a < NS; ++a) {
            const bool keep = ra[a] > SPECIES_MIN;
            s.rhoa[a * n + c] = keep ? std::max(ra[a], SPECIES_MIN) : SPECIES_MIN;
            for (int d = 0; d < 3; ++d) s.ua[(d * NS + a) * n + c] = keep ? mom[d][a] / ra[a] : 0.0;
            rho += ra[a];
            for (int d = 0; d < 3; ++d) mv[d] += mom[d][a];
        }
        s.rho[c] = rho;
        s.u[c] = mv[0] / rho;
        s.v[c] = mv[1] / rho;
        s.w[c] = mv[2] / rho;
        // Maxwell-Stefan: A u = (rho u)_a with the old state's D_ab, p and X
        const double p_old_si = pressure(Tsi_old, rho_old_si, Yold);
        const double lnT = log(Tsi_old);
        const double p = s.p[c];
        double A[NS][NS];
        for (int a = 0; a < NS; ++a)
            for (int b = 0; b < NS; ++b) {
                if (a != b) {
                    const double D = s.un.nu(binary_diff(a, b, lnT, p_old_si));
                    A[a][b] = -1.0 * 1 / 2.0 * p * s.X[a * n + c] * s.X[b * n + c] / D;
                } else {
                    const double r = s.rhoa[a * n + c];
                    A[a][a] = r == 0.0 ? 1.0 : r;
                }
            }
        for (int a = 0; a < NS; ++a)
            for (int b = 0; b < NS; ++b)
                if (b != a) A[a][a] -= A[a][b];
        // Cholesky A = L L^T, then L y = b, L^T x = y per axis
        double L[NS][NS] = {};
        for (int i = 0; i < NS; ++i) {
            for (int j = 0; j <= i; ++j) {
                double sum = A[i][j];
                for (int k = 0; k < j; ++k) sum -= L[i][k] * L[j][k];
                L[i][j] = (i == j) ? sqrt(sum) : sum / L[j][j];
            }
        }
        for (int d = 0; d < 3; ++d) {
            double y[NS], x[NS];
            for (int i = 0; i < NS; ++i) {
                double sum = mom[d][i];
                for (int k = 0; k < i; ++k) sum -= L[i][k] * y[k];
                y[i] = sum / L[i][i];
            }
            for (int i = NS - 1; i >= 0; --i) {
                double sum = y[i];
                for (int k = i + 1; k < NS; ++k) sum -= L[k][i] * x[k];
                x[i] = sum / L[i][i];
            }
            for (int a = 0; a < NS; ++a) s.ua[(d * NS + a) * n + c] = x[a];
        }
        // temperature from the internal energy at the new composition
        double Y[NS];
        normalized_Y(s, c, Y);
        double Uint = re / rho;
        for (int a = 0; a < NS; ++a) {
            double k2 = 0;
            for (int d = 0; d < 3; ++d) k2 += s.ua[(d * NS + a) * n + c] * s.ua[(d * NS + a) * n + c];
            Uint = Uint - 0.5 * s.rhoa[a * n + c] / rho * k2;
        }
        const double usi = s.un.si_energy_mass(Uint);
        const double rho_si = s.un.si_rho(rho);
        const double Tsi = T_from_UV(usi, Y, Tsi_old);
        double X[NS];
        X_from_Y(Y, X);
        for (int a = 0; a < NS; ++a) s.X[a * n + c] = X[a];
        s.T[c] = s.un.temp(Tsi);
        s.p[c] = s.un.p(pressure(Tsi, rho_si, Y));
    }
}

bool read_all(FILE* fp, void* p, size_t bytes) { return fread(p, 1, bytes, fp) == bytes; }

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: mc_cpu <case.bin> <steps> <threads> [out.bin]\n");
        return 2;
    }
    FILE* fp = fopen(argv[1], "rb");
    if (!fp) return 3;
    char magic[8];
    int32_t hdr[5];
    double dxdt[2];
    State s;
    if (!read_all(fp, magic, 8) || memcmp(magic, "MCCPU01", 8) || !read_all(fp, hdr, sizeof hdr) ||
        !read_all(fp, dxdt, sizeof dxdt)) {
        fprintf(stderr, "bad case header\n");
        return 3;
    }
    s.nx = hdr[0];
    s.ny = hdr[1];
    s.nz = hdr[2];
    s.reaction = hdr[3];
    s.n = (int64_t)s.nx * s.ny * s.nz;
    s.un = make_units(dxdt[0], dxdt[1]);
    const int64_t n = s.n;
    s.type.resize(n);
    for (auto* vec : {&s.rho, &s.u, &s.v, &s.w, &s.p, &s.T, &s.rhoe, &s.hrr}) vec->assign(n, 0.0);
    s.ptens.assign(6 * n, 0.0);
    s.g.assign(Q * n, 0.0);
    s.gpc.assign(Q * n, 0.0);
    s.X.assign(NS * n, 0.0);
    s.rhoa.assign(NS * n, 0.0);
    s.ua.assign(3 * NS * n, 0.0);
    s.f.assign((int64_t)NS * Q * n, 0.0);
    s.fpc.assign((int64_t)NS * Q * n, 0.0);
    s.src.assign(NS * n, 0.0);
    s.om_a.assign(NS * n, 0.0);
    s.Fa.assign(3 * NS * n, 0.0);
    bool ok = read_all(fp, s.type.data(), n);
    for (auto* vec : {&s.rho, &s.u, &s.v, &s.w, &s.p, &s.T}) ok = ok && read_all(fp, vec->data(), 8 * n);
    ok = ok && read_all(fp, s.X.data(), 8 * NS * n);
    const bool has_state = hdr[4] != 0;
    if (has_state)
        for (auto* vec : {&s.rhoe, &s.ptens, &s.rhoa, &s.ua, &s.f, &s.g})
            ok = ok && read_all(fp, vec->data(), 8 * vec->size());
    fclose(fp);
    if (!ok) {
        fprintf(stderr, "short case file\n");
        return 3;
    }
    for (int64_t c = 0; c < n; ++c) {
        const uint8_t t = s.type[c];
        if (t != T_F && t != T_P) {
            fprintf(stderr, "mc_cpu: cell %lld has type %d; only a periodic box (TYPE_F, TYPE_P) is restated\n",
                    (long long)c, (int)t);
            return 4;
        }
        if (t == T_F) s.fluid.push_back((int)c);
    }
    const int steps = atoi(argv[2]), threads = atoi(argv[3]);
    omp_set_num_threads(threads);
    if (!has_state) {
        init(s);
        moments(s);
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < steps; ++it) {
        collide_species(s);
        collide(s);
        stream(s);
        moments(s);
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (argc > 4) {
        FILE* fo = fopen(argv[4], "wb");
        if (!fo) return 5;
        for (auto* vec : {&s.rho, &s.u, &s.v, &s.w, &s.p, &s.T, &s.rhoe, &s.rhoa, &s.ua, &s.f, &s.g})
            fwrite(vec->data(), 8, vec->size(), fo);
        fclose(fo);
    }
    const double cells = (double)s.fluid.size();
    printf("{\"cells\": %.0f, \"steps\": %d, \"threads\": %d, \"seconds\": %.6f, \"MLUPS\": %.6f}\n", cells, steps,
           threads, sec, steps > 0 ? cells * steps / sec / 1e6 : 0.0);
    return 0;
}
