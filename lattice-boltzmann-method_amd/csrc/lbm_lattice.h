// D3Q27 lattice, product-form equilibria and finite-difference rules of the
// reference, as device (and host) inline functions for gfx950.
//
// Velocity order, weights and opposites: lbm.hpp:46-53.  Cell types:
// defines.hpp:136-145.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

enum CellType : uint8_t {
    TYPE_F = 0, TYPE_S = 1, TYPE_P = 2, TYPE_A = 3, TYPE_FS = 4,
    TYPE_I = 5, TYPE_O = 6, TYPE_I_C = 7, TYPE_O_C = 8, TYPE_Q = 9
};

constexpr int Q = 27;

__host__ __device__ constexpr int cx(int l)
{
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return t[l];
}
__host__ __device__ constexpr int cy(int l)
{
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    return t[l];
}
__host__ __device__ constexpr int cz(int l)
{
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    return t[l];
}
__host__ __device__ constexpr int opp(int l) { return l == 0 ? 0 : (l & 1 ? l + 1 : l - 1); }

// Index of the lattice direction (a, b, c), each in {-1, 0, 1}.
__host__ __device__ inline int dir_index(int a, int b, int c)
{
    int r = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l)
        if (cx(l) == a && cy(l) == b && cz(l) == c) r = l;
    return r;
}
__host__ __device__ constexpr int c2(int l) { return cx(l) * cx(l) + cy(l) * cy(l) + cz(l) * cz(l); }

// Per-axis factors of the product-form equilibrium (lbm.hpp:276-305):
// Psi(0) = 1 - P, Psi(+-1) = (+-u + P)/2, P = theta + u^2 + corr.
// psi[0] is c = -1, psi[1] is c = 0, psi[2] is c = +1.
template <typename T>
struct Axis3 { T m, z, p; };

template <typename T>
__host__ __device__ inline Axis3<T> psi_axis(T u, T P)
{
    Axis3<T> a;
    a.z = T(1) - P;
    a.p = T(0.5) * (u + P);
    a.m = T(0.5) * (P - u);
    return a;
}

template <typename T>
__host__ __device__ inline T pick(const Axis3<T>& a, int c) { return c == 0 ? a.z : (c > 0 ? a.p : a.m); }

// The energy equilibrium of lbm-edf.cpp:38-98 (27 expanded polynomials) is
//   g/rho = (U - 3 theta/2) Px Py Pz + 1/2 sum_d K_d prod_{e != d} P_e
// where P_d = Psi(c_d; v_d, theta + v_d^2) and K_d = (v_d + theta d/dv_d)^2 P_d:
//   K(0)   = theta - 3 theta^2 + v^2 - 6 theta v^2 - v^4
//   K(+-1) = (+-(v^3 + 3 theta v) + 3 theta^2 + 6 theta v^2 + v^4) / 2
// Both axis tables are built once per cell; each direction then costs a few
// multiplies instead of the ~550 flops of a verbatim polynomial.
template <typename T>
struct GeqAxes { Axis3<T> P[3], K[3]; };

template <typename T>
__host__ __device__ inline GeqAxes<T> geq_axes(T theta, const T v[3])
{
    GeqAxes<T> G;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T u = v[d], u2 = u * u;
        G.P[d] = psi_axis<T>(u, theta + u2);
        T even = T(3) * theta * theta + T(6) * theta * u2 + u2 * u2;
        T odd = u2 * u + T(3) * theta * u;
        G.K[d].z = theta - T(3) * theta * theta + u2 - T(6) * theta * u2 - u2 * u2;
        G.K[d].p = T(0.5) * (even + odd);
        G.K[d].m = T(0.5) * (even - odd);
    }
    return G;
}

template <typename T>
__host__ __device__ inline T geq_dir(const GeqAxes<T>& G, int l, T rho, T Ushift)
{
    const int a = cx(l), b = cy(l), c = cz(l);
    T px = pick(G.P[0], a), py = pick(G.P[1], b), pz = pick(G.P[2], c);
    T kx = pick(G.K[0], a), ky = pick(G.K[1], b), kz = pick(G.K[2], c);
    T pyz = py * pz;
    return rho * (Ushift * px * pyz + T(0.5) * (kx * pyz + px * (ky * pz + py * kz)));
}

// fd_central of FD.hpp:14-33 (dx = 1): one-sided next to S/A/FS/P/O/O_C,
// zero when both sides are such and the right one is periodic.
__host__ __device__ inline bool fd_wallish(uint8_t t)
{
    return t == TYPE_S || t == TYPE_A || t == TYPE_FS || t == TYPE_P || t == TYPE_O || t == TYPE_O_C;
}

template <typename T>
__host__ __device__ inline T fd_central(T l, T c, T r, uint8_t tl, uint8_t tr)
{
    if (fd_wallish(tl)) return tr == TYPE_P ? T(0) : (r - c);
    if (fd_wallish(tr)) return tl == TYPE_P ? T(0) : (c - l);
    return T(0.5) * (r - l);
}

// fd_fuw of FD.cpp:5-23 (dx = 1): first-order upwind, zero next to any
// non-fluid neighbour on the upwind side.
__host__ __device__ inline bool fd_blocked(uint8_t t)
{
    return t == TYPE_S || t == TYPE_A || t == TYPE_FS || t == TYPE_P || t == TYPE_O || t == TYPE_O_C ||
           t == TYPE_I || t == TYPE_I_C;
}

template <typename T>
__host__ __device__ inline T fd_fuw(T l, T c, T r, T vel, uint8_t tl, uint8_t tr)
{
    if (vel > T(0)) return fd_blocked(tl) ? T(0) : (c - l);
    return fd_blocked(tr) ? T(0) : (r - c);
}

}  // namespace lbm
