// Thermo, kinetics, transport and unit conversions of the multicomponent
// path, as __host__ __device__ functions over the generated mechanism tables
// (mech.h).
//
// These replace the Cantera calls the reference makes per cell (ideal-gas
// NASA7 thermo, setState_UV / TD / TP, getNetProductionRates with
// Arrhenius, third-body, Troe and Lindemann falloff and pressure-dependent
// Arrhenius (PLOG) reactions and reverse rates from equilibrium constants,
// mixture-averaged transport) and the binary32
// Units_Conv crossings between lattice and SI units (units.hpp:75-131).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "mech.h"

namespace lbm {
namespace chem {

using namespace mech;
constexpr double SPECIES_MIN = 1e-8;  // defines.hpp:96

// 1 / W_a, for products in place of divisions by the molar masses
struct InvW {
    double v[NS];
};
constexpr InvW make_inv_w()
{
    InvW r{};
    for (int a = 0; a < NS; ++a) r.v[a] = 1.0 / W[a];
    return r;
}
constexpr InvW INV_W_TABLE = make_inv_w();
__host__ __device__ inline double inv_w(int a) { return INV_W_TABLE.v[a]; }

// Reciprocal in arithmetic type R: exact for double; for float (the FP32
// path's binary32 transport) the hardware reciprocal, within 1 ulp.
__host__ __device__ inline double rcp_r(double x) { return 1.0 / x; }
__host__ __device__ inline float rcp_r(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// ---------------------------------------------------------------- units
// Units_Conv after set_m_kg_s(dx, dt): m, kg, s, K stored as float; every
// conversion rounds its argument to float and computes in float, or in
// double where the reference's double cb()/sq() enters, then returns float.
struct Units {
    float m, kg, s, K;

    __host__ __device__ float temp(double x) const { return (float)x / K; }
    __host__ __device__ float si_temp(double x) const { return (float)x * K; }
    __host__ __device__ float energy_mass(double x) const { return (float)x * (s * s) / (m * m); }
    __host__ __device__ float si_energy_mass(double x) const { return (float)x * (m * m) / (s * s); }
    __host__ __device__ float u(double x) const { return (float)x * s / m; }  // velocity
    __host__ __device__ float rho(double x) const
    {
        const double md = m;
        return (float)((double)(float)x * (md * md * md) / (double)kg);
    }
    __host__ __device__ float si_rho(double x) const
    {
        const double md = m;
        return (float)((double)((float)x * kg) / (md * md * md));
    }
    __host__ __device__ float rho_dot(double x) const
    {
        const double md = m;
        return (float)((double)(float)x * (md * md * md) * (double)s / (double)kg);
    }
    __host__ __device__ float nu(double x) const
    {
        const double md = m;
        return (float)((double)((float)x * s) / (md * md));
    }
    __host__ __device__ float mu(double x) const { return (float)x * s * m / kg; }
    __host__ __device__ float thermal_conductivity(double x) const { return (float)x * s * s * s * K / (kg * m); }
    __host__ __device__ float cp(double x) const { return (float)x * s * s * K / (m * m); }
    __host__ __device__ float p(double x) const
    {
        const double sd = s;
        return (float)((double)(float)x * ((double)m * (sd * sd)) / (double)kg);
    }
    __host__ __device__ float si_p(double x) const
    {
        const double sd = s;
        return (float)((double)((float)x * kg) / ((double)m * (sd * sd)));
    }
};

// set_m_kg_s(dx, dt) of units.hpp:32-44.
inline Units make_units(double dx, double dt)
{
    Units u;
    u.m = (float)((double)(float)dx / 1.0);
    const double md = u.m;
    u.kg = (float)(1.0 / 1.0 * (md * md * md));
    u.s = (float)((double)(float)dt / 1.0);
    u.K = 12000.0f;
    return u;
}

// --------------------------------------------------------------- thermo
__host__ __device__ inline const double* nasa(int k, double T) { return T < T_MID[k] ? NASA_LO[k] : NASA_HI[k]; }

__host__ __device__ inline double cp_R(int k, double T)
{
    const double* a = nasa(k, T);
    return a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
}

__host__ __device__ inline double h_RT(int k, double T)
{
    const double* a = nasa(k, T);
    return a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
}

__host__ __device__ inline double s_R(int k, double T, double lnT)
{
    const double* a = nasa(k, T);
    return a[0] * lnT + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
}

// Cantera's setMassFractions: negatives clipped, then normalised.
__host__ __device__ inline void normalize(double Y[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        Y[k] = Y[k] > 0.0 ? Y[k] : 0.0;
        s += Y[k];
    }
    const double inv = 1.0 / s;
#pragma unroll
    for (int k = 0; k < NS; ++k) Y[k] *= inv;
}

// FAST = true (the FP32 path): products with 1/W and one reciprocal in place
// of a division per species (each result within a few ulp of the exact
// form, which the plain path keeps).
template <bool FAST = false>
__host__ __device__ inline double mean_W(const double Y[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) s += FAST ? Y[k] * inv_w(k) : Y[k] / W[k];
    return 1.0 / s;
}

template <bool FAST = false>
__host__ __device__ inline void X_from_Y(const double Y[NS], double X[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        X[k] = FAST ? Y[k] * inv_w(k) : Y[k] / W[k];
        s += X[k];
    }
    if constexpr (FAST) {
        const double is = 1.0 / s;
#pragma unroll
        for (int k = 0; k < NS; ++k) X[k] *= is;
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k) X[k] /= s;
    }
}

__host__ __device__ inline void Y_from_X(const double X[NS], double Y[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        Y[k] = (X[k] > 0.0 ? X[k] : 0.0) * W[k];
        s += Y[k];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) Y[k] /= s;
}

template <bool FAST = false>
__host__ __device__ inline double cp_mass(double T, const double Y[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) s += FAST ? Y[k] * cp_R(k, T) * inv_w(k) : Y[k] * cp_R(k, T) / W[k];
    return s * R_UNIV;
}

// h / RT with the reciprocal temperature given (one division per state
// instead of one per species)
__host__ __device__ inline double h_RT_inv(int k, double T, double invT)
{
    const double* a = nasa(k, T);
    return a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] * invT;
}

__host__ __device__ inline double u_mass(double T, const double Y[NS])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) s += Y[k] * (h_RT(k, T) - 1.0) / W[k];
    return s * R_UNIV * T;
}

// setState_UV at fixed composition (tolerance 1e-15 as the reference asks):
// Newton on u(T) = u_target from the current temperature T0.
// (FAST: Y/W hoisted as products with 1/W, one reciprocal of T per
// iteration; the same Newton and stopping rule.)
template <bool FAST = false>
__host__ __device__ inline double T_from_UV(double u_target, const double Y[NS], double T0)
{
    double T = T0;
    if constexpr (FAST) {
        double yw[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) yw[k] = Y[k] * inv_w(k);
        for (int it = 0; it < 60; ++it) {
            const double invT = 1.0 / T;
            double u = 0, cv = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                u += yw[k] * (h_RT_inv(k, T, invT) - 1.0);
                cv += yw[k] * (cp_R(k, T) - 1.0);
            }
            const double dT = -(u * R_UNIV * T - u_target) / (cv * R_UNIV);
            T += dT;
            if (fabs(dT) <= 1e-15 * fabs(T)) break;
        }
        return T;
    }
    for (int it = 0; it < 60; ++it) {
        double u = 0, cv = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const double yw = Y[k] / W[k];
            u += yw * (h_RT(k, T) - 1.0);
            cv += yw * (cp_R(k, T) - 1.0);
        }
        const double dT = -(u * R_UNIV * T - u_target) / (cv * R_UNIV);
        T += dT;
        if (fabs(dT) <= 1e-15 * fabs(T)) break;
    }
    return T;
}

template <bool FAST = false>
__host__ __device__ inline double pressure(double T, double rho, const double Y[NS])
{
    if constexpr (FAST) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) s += Y[k] * inv_w(k);
        return rho * R_UNIV * T * s;
    }
    return rho * R_UNIV * T / mean_W(Y);
}

// ------------------------------------------------------------ kinetics
// Concentration factor C^o of a rate of progress: plain products for the
// exponents 0-3 (any sign of C), otherwise (fractional reaction orders) C^o
// when C > 0, else 0 (Cantera's rule for non-integer exponents).  The
// exponents are table constants, so the branches fold at compile time.
__host__ __device__ inline double conc_power(double c, double o)
{
    if (o == 0.0) return 1.0;
    if (o == 1.0) return c;
    if (o == 2.0) return c * c;
    if (o == 3.0) return c * c * c;
    return c > 0.0 ? pow(c, o) : 0.0;
}

// Pressure-dependent Arrhenius (PLOG) rate of reaction r at ln P [Pa]:
// ln k interpolated linearly in ln P between the two bracketing levels, the
// first (last) level's rate below (above) the table (Cantera's PlogRate).
__host__ __device__ inline double plog_rate(int r, double lnT, double invT, double lnP)
{
    const int n = PLOG_N[r];
    auto lnk = [&](int i) { return PLOG_LNA[r][i] + PLOG_B[r][i] * lnT - PLOG_EA_R[r][i] * invT; };
    if (lnP <= PLOG_LNP[r][0]) return exp(lnk(0));
    for (int i = 0; i + 1 < n; ++i) {
        if (lnP < PLOG_LNP[r][i + 1]) {
            const double w = (lnP - PLOG_LNP[r][i]) * (1.0 / (PLOG_LNP[r][i + 1] - PLOG_LNP[r][i]));
            const double l1 = lnk(i);
            return exp(l1 + (lnk(i + 1) - l1) * w);
        }
    }
    return exp(lnk(n - 1));
}

constexpr bool any_plog()
{
    for (int r = 0; r < NR; ++r)
        if (PLOG_N[r] > 0) return true;
    return false;
}

// Net molar production rates [kmol/m^3/s] at (T, rho, Y), SI units.
__host__ __device__ inline void net_production(double T, double rho, const double Y[NS], double wdot[NS])
{
    double C[NS], gRT[NS];
    const double lnT = log(T), invT = 1.0 / T;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        C[k] = rho * Y[k] / W[k];
        gRT[k] = h_RT(k, T) - s_R(k, T, lnT);
        wdot[k] = 0.0;
    }
    const double ln_pref_RT = log(P_REF / (R_UNIV * T));
    // ln p of the PLOG rates, p = R T sum_k C_k
    double lnP = 0.0;
    if constexpr (any_plog()) {
        double Ct = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) Ct += C[k];
        lnP = log(R_UNIV * T * Ct);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double kf = KIND[r] == 3 ? plog_rate(r, lnT, invT, lnP) : ARR_A[r] * exp(ARR_B[r] * lnT - ARR_EA_R[r] * invT);
        if (KIND[r] == 1 || KIND[r] == 2) {
            double M = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k) M += EFF[r][k] * C[k];
            if (KIND[r] == 1) {
                kf *= M;
            } else {
                const double k0 = LOW_A[r] * exp(LOW_B[r] * lnT - LOW_EA_R[r] * invT);
                const double Pr = k0 * M / kf;
                if (TROE_ON[r]) {
                    double Fc = (1.0 - TROE[r][0]) * exp(-T / TROE[r][1]) + TROE[r][0] * exp(-T / TROE[r][2]);
                    if (TROE_HAS_T2[r]) Fc += exp(-TROE[r][3] * invT);
                    const double lFc = log10(Fc);
                    const double lPr = log10(Pr > 1e-300 ? Pr : 1e-300);
                    const double cc = -0.4 - 0.67 * lFc;
                    const double nn = 0.75 - 1.27 * lFc;
                    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
                    kf = kf * Pr / (1.0 + Pr) * exp10(lFc / (1.0 + f1 * f1));
                } else {  // Lindemann: broadening factor 1
                    kf = kf * Pr / (1.0 + Pr);
                }
            }
        }
        double dG = 0, dn = 0, cf = 1, cr = 1;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            dG += (NU_R[r][k] - NU_F[r][k]) * gRT[k];
            dn += NU_R[r][k] - NU_F[r][k];
            cf *= conc_power(C[k], ORD_F[r][k]);
            cr *= conc_power(C[k], NU_R[r][k]);
        }
        // k_r = k_f / K_c, K_c = exp(-dG) (p_ref / RT)^dn; 0 for irreversible steps
        const double kr = REVERSIBLE[r] ? kf * exp(dG - dn * ln_pref_RT) : 0.0;
        const double q = kf * cf - kr * cr;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if (NU_R[r][k] - NU_F[r][k] != 0.0) wdot[k] += (NU_R[r][k] - NU_F[r][k]) * q;
    }
}

// ------------------------------------------------------------ transport
__host__ __device__ inline double poly4(const double c[5], double x)
{
    return c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4])));
}

__host__ __device__ inline double species_viscosity(int k, double lnT) { return exp(poly4(VISC_FIT[k], lnT)); }
__host__ __device__ inline double species_conductivity(int k, double lnT) { return exp(poly4(COND_FIT[k], lnT)); }
// D_ab [m^2/s] at (T, p).
__host__ __device__ inline double binary_diff(int a, int b, double lnT, double p)
{
    return exp(poly4(DIFF_FIT[a][b], lnT)) / p;
}

// The same fits evaluated in arithmetic type R (double reproduces the
// functions above bit for bit; float is the FP32 path's binary32 transport).
template <typename R>
__host__ __device__ inline R poly4r(const double c[5], R x)
{
    return (R)c[0] + x * ((R)c[1] + x * ((R)c[2] + x * ((R)c[3] + x * (R)c[4])));
}
template <typename R>
__host__ __device__ inline R species_viscosity_r(int k, R lnT)
{
    return exp(poly4r<R>(VISC_FIT[k], lnT));
}
template <typename R>
__host__ __device__ inline R binary_diff_r(int a, int b, R lnT, R p)
{
    if constexpr (sizeof(R) == 4) return exp(poly4r<R>(DIFF_FIT[a][b], lnT)) * rcp_r(p);
    return exp(poly4r<R>(DIFF_FIT[a][b], lnT)) / p;
}

// Wilke weight Phi_ab with sqrt(mu_a / mu_b) given: (1 + sqrt(mu_a/mu_b)
// (W_b/W_a)^(1/4))^2 / sqrt(8 (1 + W_a/W_b)), the W factors folded at compile time.
__host__ __device__ inline float wilke_phi(int a, int b, float sq_ratio)
{
    const float r = 1.0f + sq_ratio * (float)sqrt(sqrt(W[b] / W[a]));
    return r * r * (float)(1.0 / sqrt(8.0 * (1.0 + W[a] / W[b])));
}

// Wilke mixture viscosity with the direct Phi_ab formula.
__host__ __device__ inline double mix_viscosity(double lnT, const double X[NS])
{
    double mu[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) mu[k] = species_viscosity(k, lnT);
    double tot = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        double den = 0;
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            const double r = 1.0 + sqrt(mu[a] / mu[b]) * sqrt(sqrt(W[b] / W[a]));
            den += X[b] * (r * r / sqrt(8.0 * (1.0 + W[a] / W[b])));
        }
        tot += X[a] * mu[a] / den;
    }
    return tot;
}

template <typename R>
__host__ __device__ inline R mix_viscosity_r(R lnT, const double X[NS])
{
    if constexpr (sizeof(R) == 4) {
        // sqrt(mu_a / mu_b) = sqrt(mu_a) / sqrt(mu_b): 9 square roots and
        // reciprocals in place of 81 divisions and square roots
        float mu[NS], sm[NS], rsm[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            mu[k] = species_viscosity_r<float>(k, lnT);
            sm[k] = sqrt(mu[k]);
            rsm[k] = rcp_r(sm[k]);
        }
        float tot = 0;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            float den = 0;
#pragma unroll
            for (int b = 0; b < NS; ++b) den += (float)X[b] * wilke_phi(a, b, sm[a] * rsm[b]);
            tot += (float)X[a] * mu[a] * rcp_r(den);
        }
        return tot;
    }
    R mu[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) mu[k] = species_viscosity_r<R>(k, lnT);
    R tot = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        R den = 0;
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            const R r = R(1) + sqrt(mu[a] / mu[b]) * (R)sqrt(sqrt(W[b] / W[a]));
            den += (R)X[b] * (r * r / (R)sqrt(8.0 * (1.0 + W[a] / W[b])));
        }
        tot += (R)X[a] * mu[a] / den;
    }
    return tot;
}

template <typename R, typename XT = double>
__host__ __device__ inline R mix_conductivity_r(R lnT, const XT X[NS])
{
    R s1 = 0, s2 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const R l = exp(poly4r<R>(COND_FIT[k], lnT));
        s1 += (R)X[k] * l;
        if constexpr (sizeof(R) == 4)
            s2 += (R)X[k] * rcp_r(l);
        else
            s2 += (R)X[k] / l;
    }
    if constexpr (sizeof(R) == 4) return R(0.5) * (s1 + rcp_r(s2));
    return R(0.5) * (s1 + R(1) / s2);
}

// Combination-averaged mixture conductivity 1/2 (sum X l + 1 / sum X / l).
__host__ __device__ inline double mix_conductivity(double lnT, const double X[NS])
{
    double s1 = 0, s2 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const double l = species_conductivity(k, lnT);
        s1 += X[k] * l;
        s2 += X[k] / l;
    }
    return 0.5 * (s1 + 1.0 / s2);
}

}  // namespace chem
}  // namespace lbm
