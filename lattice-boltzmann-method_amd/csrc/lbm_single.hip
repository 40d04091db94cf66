// Single-component D3Q27 thermal kernels for gfx950.
//
// Storage: structure of arrays, population-major [27][ncell], ghost-inclusive
// cell index c = (i*ny + j)*nz + k.  One thread per cell; consecutive lanes
// touch consecutive cells of every population plane, so each of the 54
// population loads and stores of a wavefront is one contiguous 512 B (FP64)
// or 256 B (FP32) segment, shifted by the streaming offset for pulls.
//
// Reference routines restated here:
//   Init           lbm-init.cpp:8-82
//   Collide        lbm-collision.cpp:9-99
//   Streaming      lbm-streaming.cpp:3-119 (+ dirSlip, lbm-bc.cpp:1352-1399)
//   calculate_moment lbm-moment.cpp:11-85
//   TMS_BC         lbm-bc.cpp:341-460, 809-938, 1111-1230 (+ calculate_feq_geq,
//                  lbm-edf.cpp:153-164)
//   OutputKeEns    output.cpp:212-276 (diagnostic reduction)
#include <type_traits>

#include "lbm_internal.h"

namespace lbm {

constexpr int BLOCK = 256;

template <typename P>
constexpr bool kShift = std::is_same<P, float>::value;

struct Ijk {
    int i, j, k;
};

__device__ inline Ijk ijk_of(const Dims& d, int c)
{
    int k = c % d.nz;
    int t = c / d.nz;
    return {t / d.ny, t % d.ny, k};
}

__device__ inline int cell_of(const Dims& d, int i, int j, int k) { return (i * d.ny + j) * d.nz + k; }

// Pull sources along one axis after the periodic wrap of lbm-streaming.cpp:103-105:
// s[c + 1] is the index x - c, with ghost index 0 -> n-2 and n-1 -> 1.
struct Src3 {
    int s[3];
};
__device__ inline Src3 src_axis(int x, int n)
{
    Src3 r;
    r.s[0] = (x + 1 == n - 1) ? 1 : x + 1;
    r.s[1] = x;
    r.s[2] = (x - 1 == 0) ? n - 2 : x - 1;
    return r;
}


template <typename P>
__device__ inline P ldp(const P* base, int l, int64_t n, int c)
{
    return base[(int64_t)l * n + c];
}
template <typename P>
__device__ inline void stp(P* base, int l, int64_t n, int c, P x)
{
    base[(int64_t)l * n + c] = x;
}

// Pull for a cell whose neighbours are all fluid or periodic ghosts.
template <typename P>
__device__ inline void pull_wrap(const SingleView<P>& v, int i, int j, int k, P fl[Q], P gl[Q])
{
    const Src3 si = src_axis(i, v.d.nx), sj = src_axis(j, v.d.ny), sk = src_axis(k, v.d.nz);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        int src = (si.s[cx(l) + 1] * v.d.ny + sj.s[cy(l) + 1]) * v.d.nz + sk.s[cz(l) + 1];
        fl[l] = ldp(v.f, l, v.d.n, src);
        gl[l] = ldp(v.g, l, v.d.n, src);
    }
}

// General pull with the wall rules of lbm-streaming.cpp:24-117.  Directions
// whose value the reference leaves untouched are flagged in keepf / keepg.
template <typename P>
__device__ inline void pull_general(const SingleView<P>& v, int i, int j, int k, int c, P fl[Q], P gl[Q],
                                    uint32_t& keepf, uint32_t& keepg)
{
    const Src3 si = src_axis(i, v.d.nx), sj = src_axis(j, v.d.ny), sk = src_axis(k, v.d.nz);
    keepf = 0;
    keepg = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const uint8_t tn = v.type[cell_of(v.d, i - cx(l), j - cy(l), k - cz(l))];
        if (tn == TYPE_S) {
            fl[l] = ldp(v.f, opp(l), v.d.n, c);
            gl[l] = P(0);
            keepg |= 1u << l;
        } else if (tn == TYPE_FS) {
            int a = cx(l), b = cy(l), e = cz(l);
            int ip = i - a, jp = j - b, kp = k - e;
            if (v.type[cell_of(v.d, i - a, j, k)] == TYPE_FS) {
                a = -a;
                ip = i;
            } else if (v.type[cell_of(v.d, i, j - b, k)] == TYPE_FS) {
                b = -b;
                jp = j;
            } else if (v.type[cell_of(v.d, i, j, k - e)] == TYPE_FS) {
                e = -e;
                kp = k;
            }
            ip = ((ip - 1 + (v.d.nx - 2)) % (v.d.nx - 2)) + 1;
            jp = ((jp - 1 + (v.d.ny - 2)) % (v.d.ny - 2)) + 1;
            kp = ((kp - 1 + (v.d.nz - 2)) % (v.d.nz - 2)) + 1;
            const int lp = dir_index(a, b, e);
            const int src = cell_of(v.d, ip, jp, kp);
            fl[l] = ldp(v.f, lp, v.d.n, src);
            gl[l] = ldp(v.g, lp, v.d.n, src);
        } else if (tn == TYPE_A) {
            fl[l] = ldp(v.f, opp(l), v.d.n, c);
            gl[l] = ldp(v.g, opp(l), v.d.n, c);
        } else if (tn == TYPE_O || tn == TYPE_O_C || tn == TYPE_I || tn == TYPE_I_C) {
            fl[l] = P(0);
            gl[l] = P(0);
            keepf |= 1u << l;
            keepg |= 1u << l;
        } else {
            int src = (si.s[cx(l) + 1] * v.d.ny + sj.s[cy(l) + 1]) * v.d.nz + sk.s[cz(l) + 1];
            fl[l] = ldp(v.f, l, v.d.n, src);
            gl[l] = ldp(v.g, l, v.d.n, src);
        }
    }
}

// Moments of calculate_moment (lbm-moment.cpp:35-77), in the collision type.
// With shifted storage the rest-state sums are added back.
template <typename P>
struct Mom {
    P rho, m[3], Pt[6], rhoe, q[3];  // Pt: xx yy zz xy xz yz
};

template <typename P>
__device__ inline Mom<P> moments_of(const SingleView<P>& v, const P fl[Q], const P gl[Q])
{
    Mom<P> M;
    P r = 0, mx = 0, my = 0, mz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
    P e = 0, qx = 0, qy = 0, qz = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const P a = fl[l], b = gl[l];
        r += a;
        if (cx(l)) mx += cx(l) * a;
        if (cy(l)) my += cy(l) * a;
        if (cz(l)) mz += cz(l) * a;
        if (cx(l)) pxx += a;
        if (cy(l)) pyy += a;
        if (cz(l)) pzz += a;
        if (cx(l) * cy(l)) pxy += (cx(l) * cy(l)) * a;
        if (cx(l) * cz(l)) pxz += (cx(l) * cz(l)) * a;
        if (cy(l) * cz(l)) pyz += (cy(l) * cz(l)) * a;
        e += b;
        if (cx(l)) qx += cx(l) * b;
        if (cy(l)) qy += cy(l) * b;
        if (cz(l)) qz += cz(l) * b;
    }
    if constexpr (kShift<P>) {
        const double* s = v.fref_sum;  // rho0, rho0 theta0 (x3), rhoe0
        r += P(s[0]);
        pxx += P(s[1]);
        pyy += P(s[2]);
        pzz += P(s[3]);
        e += P(s[4]);
    }
    M.rho = r;
    M.m[0] = mx; M.m[1] = my; M.m[2] = mz;
    M.Pt[0] = pxx; M.Pt[1] = pyy; M.Pt[2] = pzz; M.Pt[3] = pxy; M.Pt[4] = pxz; M.Pt[5] = pyz;
    M.rhoe = e;
    M.q[0] = qx; M.q[1] = qy; M.q[2] = qz;
    return M;
}

// Primitive state of one cell.
template <typename P>
struct Prim {
    P rho, u[3], p, T, rhoe, Pt[6];
};

// calculate_moment's derived quantities (lbm-moment.cpp:58-77).  T is formed
// in FP64 for both storage types.
template <typename P>
__device__ inline Prim<P> prim_of(const SingleView<P>& v, const Mom<P>& M)
{
    Prim<P> s;
    s.rho = M.rho;
    s.u[0] = M.m[0] / M.rho;
    s.u[1] = M.m[1] / M.rho;
    s.u[2] = M.m[2] / M.rho;
    s.rhoe = M.rhoe;
    const double ud = s.u[0], vd = s.u[1], wd = s.u[2];
    const double eint = (double)M.rhoe / (double)M.rho - 0.5 * (ud * ud + vd * vd + wd * wd);
    const double T = eint / v.cv;
    s.T = P(T);
    s.p = P((double)M.rho * v.R * T);
#pragma unroll
    for (int a = 0; a < 6; ++a) s.Pt[a] = M.Pt[a];
    return s;
}

// Q_d of lbm-collision.cpp:65-67.
template <typename P>
__device__ inline P qflux(P rho, P ud, P T, P R)
{
    return rho * ud * (P(1) - P(3) * R * T) - rho * (ud * ud * ud);
}

// The single-component collision of lbm-collision.cpp:23-97 for one cell.
// fl/gl hold the (possibly shifted) populations and are overwritten with the
// post-collision ones in the same storage convention.
// kfac scales the conductivity: 1 for gas, 10 for the solid cells of the
// conjugate variant (lbm-collision.cpp:107).
template <typename P>
__device__ inline bool collide_cell(const SingleView<P>& v, const Prim<P>& s, const P dQ[3], P fl[Q], P gl[Q],
                                    P kfac = P(1))
{
    const P R = P(v.R), cp = P(v.cp);
    const P rho = s.rho, p = s.p;
    const P theta = R * s.T;
    const P mu = P(v.nu) * rho;
    const P kappa = kfac == P(1) ? mu * cp / P(v.prtl) : kfac * mu * cp / P(v.prtl);
    const P om1 = P(2) * p * cp / (p * cp + P(2) * kappa);
    const P om = P(2) * p / (p + P(2) * mu);
    const bool warn = (P(1) / om <= P(0.5)) || (P(1) / om1 <= P(0.5));
    const P* u = s.u;
    const P Uint = s.rhoe / rho - P(0.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    // eq. pressure tensor p delta + rho u u, and the shift of g* (d_str_heat_flux)
    P dq[3], corr[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        P acc = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int idx = (a == b) ? a : (3 + a + b - 1);  // xy=3, xz=4, yz=5
            const P eqP = (a == b ? p : P(0)) + rho * u[a] * u[b];
            acc += u[a] * (s.Pt[idx] - eqP);
        }
        dq[b] = acc + P(0.5) * u[b] * dQ[b];
        corr[b] = (P(2) - om) / (P(2) * rho * om) * dQ[b];
    }
    Axis3<P> F[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = psi_axis<P>(u[d], theta + u[d] * u[d] + corr[d]);
    const GeqAxes<P> G = geq_axes<P>(theta, u);
    const P Ushift = Uint - P(1.5) * theta;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const P feq = rho * pick(F[0], cx(l)) * pick(F[1], cy(l)) * pick(F[2], cz(l));
        const P geq = geq_dir(G, l, rho, Ushift);
        const P cdq = P(0.5) * (P(cx(l)) * dq[0] + P(cy(l)) * dq[1] + P(cz(l)) * dq[2]);
        if constexpr (kShift<P>) {
            const P fr = P(v.fref_d(l)), gr = P(v.gref_d(l));
            fl[l] = (P(1) - om) * fl[l] + om * (feq - fr);
            const P dg = gl[l];
            const P geq_rel = geq - gr;
            gl[l] = dg + om1 * (geq_rel - dg) - (c2(l) == 1 ? (om - om1) * cdq : P(0));
        } else {
            fl[l] = (P(1) - om) * fl[l] + om * feq;
            const P gstr = (c2(l) == 1) ? geq + cdq : geq;
            const P g = gl[l];
            gl[l] = g + om1 * (geq - g) + (om - om1) * (geq - gstr);
        }
    }
    return warn;
}

template <typename P>
__device__ inline void count_warn(const SingleView<P>& v, bool warn)
{
    // one atomic per wavefront that saw a warning
    const unsigned long long m = __ballot(warn);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(v.tau_count, (unsigned long long)__popcll(m));
}

// -------------------------------------------------------------------- Init
// Moment-form energy equilibrium of lbm-edf.cpp:100-138 (used by Init only).
__device__ inline double geq_moment(int l, double rhoe, const double q[3], const double Rt[9], double theta)
{
    const double c[3] = {double(cx(l)), double(cy(l)), double(cz(l))};
    double g = rhoe + (q[0] * c[0] + q[1] * c[1] + q[2] * c[2]) / theta;
    double tr = 0;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        double col = 0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const double A = Rt[3 * m + n] - (m == n ? rhoe * theta : 0.0);
            const double B = c[m] * c[n] - (m == n ? theta : 0.0);
            col += A * B;
        }
        tr += col;
    }
    g += tr / (2.0 * theta * theta);
    double w = 1.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) w *= (c[m] == 0.0) ? (1.0 - theta) : 0.5 * theta;
    g = w * g;
    const int cc = c2(l);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double B = (cc == 0) ? 1.0 : (cc == 1 ? -0.5 * fabs(c[m]) : 0.0);
        const double Z = (1.0 - 3.0 * theta) / (2.0 * theta) * (Rt[4 * m] - theta * rhoe);
        g += B * Z;
    }
    return g;
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_init(SingleView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    const uint8_t t = v.type[c];
    if (!(t == TYPE_F || t == TYPE_S || t == TYPE_I || t == TYPE_O || t == TYPE_Q)) return;
    double* const* F = v.fld;
    const double p = F[LBM_FIELD_P][c], T = F[LBM_FIELD_T][c];
    const double rho = p / (v.R * T);
    const double theta = v.R * T;
    F[LBM_FIELD_RHO][c] = rho;
    F[LBM_FIELD_PXX][c] = p;
    F[LBM_FIELD_PYY][c] = p;
    F[LBM_FIELD_PZZ][c] = p;
    F[LBM_FIELD_PXY][c] = 0.0;
    F[LBM_FIELD_PXZ][c] = 0.0;
    F[LBM_FIELD_PYZ][c] = 0.0;
    const double rhoe = rho * (v.cv * T);
    F[LBM_FIELD_RHOE][c] = rhoe;
    const double H = v.cp * T;
    const double q[3] = {H * rho * F[LBM_FIELD_U][c], H * rho * F[LBM_FIELD_V][c], H * rho * F[LBM_FIELD_W][c]};
    const double Rt[9] = {H * p, 0, 0, 0, H * p, 0, 0, 0, H * p};
    const Axis3<double> z = psi_axis<double>(0.0, theta);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const double feq = rho * pick(z, cx(l)) * pick(z, cy(l)) * pick(z, cz(l));
        const double geq = geq_moment(l, rhoe, q, Rt, theta);
        stp(v.fo, l, v.d.n, c, P(feq - v.fref_d(l)));
        stp(v.go, l, v.d.n, c, P(geq - v.gref_d(l)));
    }
}

// ------------------------------------------------ strict, one kernel per call
template <typename P>
__device__ inline P field_q(const SingleView<P>& v, int c, int d)
{
    return qflux<P>(P(v.fld[LBM_FIELD_RHO][c]), P(v.fld[LBM_FIELD_U + d][c]), P(v.fld[LBM_FIELD_T][c]), P(v.R));
}

// A conducting solid cell of the conjugate variant: TYPE_S off the ghost
// layer (ghost-layer TYPE_S cells stay isothermal walls; the reference's solid
// loops would index outside the lattice there).
template <typename P>
__device__ inline bool conj_solid(const SingleView<P>& v, int c)
{
    if (!v.conjugate || v.type[c] != TYPE_S) return false;
    const Ijk x = ijk_of(v.d, c);
    return x.i > 0 && x.j > 0 && x.k > 0 && x.i < v.d.nx - 1 && x.j < v.d.ny - 1 && x.k < v.d.nz - 1;
}

// Collide from explicit f/g and the stored moments (the reference's Collide).
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_collide_explicit(SingleView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    bool warn = false;
    // CONJUGATE: solid cells collide too, from their stored fields, with ten
    // times the conductivity (lbm-collision.cpp:101-180)
    const bool solid = c < v.d.n && conj_solid(v, c);
    if (c < v.d.n && (v.type[c] == TYPE_F || solid)) {
        const Ijk x = ijk_of(v.d, c);
        double* const* F = v.fld;
        Prim<P> s;
        s.rho = P(F[LBM_FIELD_RHO][c]);
        s.u[0] = P(F[LBM_FIELD_U][c]);
        s.u[1] = P(F[LBM_FIELD_V][c]);
        s.u[2] = P(F[LBM_FIELD_W][c]);
        s.p = P(F[LBM_FIELD_P][c]);
        s.T = P(F[LBM_FIELD_T][c]);
        s.rhoe = P(F[LBM_FIELD_RHOE][c]);
#pragma unroll
        for (int a = 0; a < 6; ++a) s.Pt[a] = P(F[LBM_FIELD_PXX + a][c]);
        const int nm[3] = {cell_of(v.d, x.i - 1, x.j, x.k), cell_of(v.d, x.i, x.j - 1, x.k), c - 1};
        const int np[3] = {cell_of(v.d, x.i + 1, x.j, x.k), cell_of(v.d, x.i, x.j + 1, x.k), c + 1};
        P dQ[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            dQ[d] = fd_central<P>(field_q(v, nm[d], d), field_q(v, c, d), field_q(v, np[d], d), v.type[nm[d]],
                                  v.type[np[d]]);
        P fl[Q], gl[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            fl[l] = ldp(v.f, l, v.d.n, c);
            gl[l] = ldp(v.g, l, v.d.n, c);
        }
        warn = collide_cell(v, s, dQ, fl, gl, solid ? P(10) : P(1));
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            stp(v.fo, l, v.d.n, c, fl[l]);
            stp(v.go, l, v.d.n, c, gl[l]);
        }
    }
    count_warn(v, warn);
}

// CONJUGATE streaming of a solid cell (lbm-streaming.cpp:120-165): g only;
// nothing from gas cells (TMS_BC supplies those directions), bounce-back from
// adiabatic walls, a pull from solid and periodic cells through the ghost wrap.
template <typename P>
__device__ inline void stream_solid(const SingleView<P>& v, int c, const Ijk& x)
{
    const int nx = v.d.nx, ny = v.d.ny, nz = v.d.nz;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        int in = x.i - cx(l), jn = x.j - cy(l), kn = x.k - cz(l);
        const uint8_t tn = v.type[cell_of(v.d, in, jn, kn)];
        if (tn == TYPE_A) {
            stp(v.go, l, v.d.n, c, ldp(v.g, opp(l), v.d.n, c));
        } else if (tn == TYPE_S || tn == TYPE_P) {
            if (in == 0) in = nx - 2; else if (in == nx - 1) in = 1;
            if (jn == 0) jn = ny - 2; else if (jn == ny - 1) jn = 1;
            if (kn == 0) kn = nz - 2; else if (kn == nz - 1) kn = 1;
            stp(v.go, l, v.d.n, c, ldp(v.g, l, v.d.n, cell_of(v.d, in, jn, kn)));
        }
    }
}

// Streaming: v.f/v.g = post-collision source, v.fo/v.go = f/g destination.
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_stream(SingleView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    if (conj_solid(v, c)) {
        stream_solid(v, c, ijk_of(v.d, c));
        return;
    }
    if (v.type[c] != TYPE_F) return;
    const Ijk x = ijk_of(v.d, c);
    P fl[Q], gl[Q];
    uint32_t kf = 0, kg = 0;
    if (v.nbflag[c] & NB_SIMPLE)
        pull_wrap(v, x.i, x.j, x.k, fl, gl);
    else
        pull_general(v, x.i, x.j, x.k, c, fl, gl, kf, kg);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if (!(kf >> l & 1u)) stp(v.fo, l, v.d.n, c, fl[l]);
        if (!(kg >> l & 1u)) stp(v.go, l, v.d.n, c, gl[l]);
    }
}

template <typename P>
__device__ inline void store_fields(const SingleView<P>& v, int c, const Mom<P>& M, const Prim<P>& s)
{
    double* const* F = v.fld;
    F[LBM_FIELD_RHO][c] = s.rho;
    F[LBM_FIELD_U][c] = s.u[0];
    F[LBM_FIELD_V][c] = s.u[1];
    F[LBM_FIELD_W][c] = s.u[2];
    F[LBM_FIELD_P][c] = s.p;
    F[LBM_FIELD_T][c] = s.T;
    F[LBM_FIELD_RHOE][c] = s.rhoe;
#pragma unroll
    for (int a = 0; a < 6; ++a) F[LBM_FIELD_PXX + a][c] = M.Pt[a];
    F[LBM_FIELD_QX][c] = M.q[0];
    F[LBM_FIELD_QY][c] = M.q[1];
    F[LBM_FIELD_QZ][c] = M.q[2];
}

// CONJUGATE moments of a solid cell (lbm-moment.cpp:85-120): energy and its
// flux from g; T = (rho e / rho) / c_v and p = rho R T at the stored density.
template <typename P>
__device__ inline void moments_solid(const SingleView<P>& v, int c)
{
    double e = 0, qx = 0, qy = 0, qz = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const double g = double(ldp(v.g, l, v.d.n, c)) + v.gref_d(l);
        e += g;
        if (cx(l)) qx += cx(l) * g;
        if (cy(l)) qy += cy(l) * g;
        if (cz(l)) qz += cz(l) * g;
    }
    double* const* F = v.fld;
    const double rho = F[LBM_FIELD_RHO][c];
    const double T = e / rho / v.cv;
    F[LBM_FIELD_RHOE][c] = e;
    F[LBM_FIELD_QX][c] = qx;
    F[LBM_FIELD_QY][c] = qy;
    F[LBM_FIELD_QZ][c] = qz;
    F[LBM_FIELD_T][c] = T;
    F[LBM_FIELD_P][c] = rho * v.R * T;
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_moments(SingleView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    if (conj_solid(v, c)) {
        moments_solid(v, c);
        return;
    }
    if (v.type[c] != TYPE_F) return;
    P fl[Q], gl[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        fl[l] = ldp(v.f, l, v.d.n, c);
        gl[l] = ldp(v.g, l, v.d.n, c);
    }
    const Mom<P> M = moments_of(v, fl, gl);
    const Prim<P> s = prim_of(v, M);
    store_fields(v, c, M, s);
}

// ------------------------------------------- fused pull-stream + collide
// State between steps: post-collision populations in v.f/v.g and Q_d of
// every fluid cell in v.qd (written by k_fused_moments).  One pass pulls the
// streamed populations, recomputes the cell's moments, reads the neighbours'
// Q_d for the finite-difference correction and writes the next
// post-collision populations to v.fo/v.go.
template <typename P>
__device__ inline void pull_cell(const SingleView<P>& v, const Ijk& x, int c, P fl[Q], P gl[Q])
{
    if (v.nbflag[c] & NB_SIMPLE) {
        pull_wrap(v, x.i, x.j, x.k, fl, gl);
    } else {
        uint32_t kf, kg;  // fused mode rejects keep-type neighbours at set-up
        pull_general(v, x.i, x.j, x.k, c, fl, gl, kf, kg);
    }
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_fused_collide(SingleView<P> v)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    bool warn = false;
    if (c < v.d.n && v.type[c] == TYPE_F) {
        const Ijk x = ijk_of(v.d, c);
        P fl[Q], gl[Q];
        pull_cell(v, x, c, fl, gl);
        const Mom<P> M = moments_of(v, fl, gl);
        const Prim<P> s = prim_of(v, M);
        const int nm[3] = {cell_of(v.d, x.i - 1, x.j, x.k), cell_of(v.d, x.i, x.j - 1, x.k), c - 1};
        const int np[3] = {cell_of(v.d, x.i + 1, x.j, x.k), cell_of(v.d, x.i, x.j + 1, x.k), c + 1};
        P dQ[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const uint8_t tm = v.type[nm[d]], tp = v.type[np[d]];
            const P qm = fd_wallish(tm) ? P(0) : v.qd[d][nm[d]];
            const P qp = fd_wallish(tp) ? P(0) : v.qd[d][np[d]];
            dQ[d] = fd_central<P>(qm, qflux<P>(s.rho, s.u[d], s.T, P(v.R)), qp, tm, tp);
        }
        warn = collide_cell(v, s, dQ, fl, gl);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            stp(v.fo, l, v.d.n, c, fl[l]);
            stp(v.go, l, v.d.n, c, gl[l]);
        }
    }
    count_warn(v, warn);
}

// Moment pass of the fused mode: Q_d for the next collision; with
// `Mat` also every host field and the streamed populations (into v.fo/v.go).
template <typename P, bool Mat>
__global__ __launch_bounds__(BLOCK) void k_fused_moments(SingleView<P> v)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    const Ijk x = ijk_of(v.d, c);
    P fl[Q], gl[Q];
    pull_cell(v, x, c, fl, gl);
    const Mom<P> M = moments_of(v, fl, gl);
    const Prim<P> s = prim_of(v, M);
#pragma unroll
    for (int d = 0; d < 3; ++d) v.qd[d][c] = qflux<P>(s.rho, s.u[d], s.T, P(v.R));
    if constexpr (Mat) {
        store_fields(v, c, M, s);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            stp(v.fo, l, v.d.n, c, fl[l]);
            stp(v.go, l, v.d.n, c, gl[l]);
        }
    }
}

// ------------------------------------------------------------ diagnostics
// Per-block partial sums (mass, energy, KE, enstrophy, nan count) and max
// Mach, reduced across the wavefront with shuffles, then in LDS.
constexpr int DIAG_Q = 6;

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    return x;
}
__device__ inline double wave_max(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
    return x;
}

__global__ __launch_bounds__(BLOCK) void k_diag(Dims d, const uint8_t* type, const double* rho, const double* u,
                                                const double* vv, const double* w, const double* T,
                                                const double* rhoe, const double* p, double gamma, double* part)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    double acc[DIAG_Q] = {0, 0, 0, 0, 0, 0};
    if (c < d.n && type[c] == TYPE_F) {
        const Ijk x = ijk_of(d, c);
        const double r = rho[c];
        const double uu = u[c] * u[c] + vv[c] * vv[c] + w[c] * w[c];
        int ia = x.i + 1, ib = x.i - 1, ja = x.j + 1, jb = x.j - 1, ka = x.k + 1, kb = x.k - 1;
        if (ib < 1) ib = d.nx - 2; else if (ia > d.nx - 2) ia = 1;
        if (jb < 1) jb = d.ny - 2; else if (ja > d.ny - 2) ja = 1;
        if (kb < 1) kb = d.nz - 2; else if (ka > d.nz - 2) ka = 1;
        const double uy = (u[cell_of(d, x.i, ja, x.k)] - u[cell_of(d, x.i, jb, x.k)]) / 2;
        const double uz = (u[cell_of(d, x.i, x.j, ka)] - u[cell_of(d, x.i, x.j, kb)]) / 2;
        const double vx = (vv[cell_of(d, ia, x.j, x.k)] - vv[cell_of(d, ib, x.j, x.k)]) / 2;
        const double vz = (vv[cell_of(d, x.i, x.j, ka)] - vv[cell_of(d, x.i, x.j, kb)]) / 2;
        const double wx = (w[cell_of(d, ia, x.j, x.k)] - w[cell_of(d, ib, x.j, x.k)]) / 2;
        const double wy = (w[cell_of(d, x.i, ja, x.k)] - w[cell_of(d, x.i, jb, x.k)]) / 2;
        const double vort = (wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy);
        acc[0] = r;
        acc[1] = rhoe[c];
        acc[2] = r * uu;
        acc[3] = r * vort;
        const bool bad = !isfinite(r) || !isfinite(T[c]);
        acc[4] = bad ? 1.0 : 0.0;
        const double mach = sqrt(uu) / sqrt(gamma * p[c] / r);  // p / rho = R T of the mixture
        acc[5] = isfinite(mach) ? mach : 0.0;
    }
    __shared__ double sm[BLOCK / 64][DIAG_Q];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < DIAG_Q; ++q) {
        const double s = (q == 5) ? wave_max(acc[q]) : wave_sum(acc[q]);
        if (lane == 0) sm[wid][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < DIAG_Q) {
        const int q = threadIdx.x;
        double s = sm[0][q];
        for (int k = 1; k < BLOCK / 64; ++k) s = (q == 5) ? fmax(s, sm[k][q]) : s + sm[k][q];
        part[(size_t)blockIdx.x * DIAG_Q + q] = s;
    }
}

// Fixed-order final reduction: deterministic run to run.
__global__ __launch_bounds__(BLOCK) void k_diag_final(const double* part, int nblk, double* out)
{
    __shared__ double sm[BLOCK];
    for (int q = 0; q < DIAG_Q; ++q) {
        double s = (q == 5) ? 0.0 : 0.0;
        for (int b = threadIdx.x; b < nblk; b += BLOCK) {
            const double x = part[(size_t)b * DIAG_Q + q];
            s = (q == 5) ? fmax(s, x) : s + x;
        }
        sm[threadIdx.x] = s;
        __syncthreads();
        for (int o = BLOCK / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) sm[threadIdx.x] = (q == 5) ? fmax(sm[threadIdx.x], sm[threadIdx.x + o])
                                                            : sm[threadIdx.x] + sm[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = sm[0];
        __syncthreads();
    }
}

// ----------------------------------------------------------------- TMS_BC
// The three single-component sweeps of TMS_BC (lbm-bc.cpp:341-460 outflow,
// 809-938 inflow, 1111-1230 isothermal wall).  Each sweep rewrites only the
// populations of its own interface cell and reads the macroscopic fields of
// neighbours, which TMS_BC never writes, so applying the three sweeps in
// order to one cell at a time (one thread per interface cell) is the
// reference's sweep-after-sweep result.  Arithmetic is FP64 for both storage
// types.

// calculate_feq_geq (lbm-edf.cpp:153-164) of one state, direction by direction.
struct EqSet {
    GeqAxes<double> G;
    double rho, Ushift;
};
__device__ inline EqSet eq_set(double rho, const double u[3], double T, double R, double cv)
{
    EqSet e;
    e.G = geq_axes<double>(R * T, u);
    e.rho = rho;
    e.Ushift = cv * T - 1.5 * (R * T);
    return e;
}
__device__ inline double eq_f(const EqSet& e, int l)
{
    return e.rho * pick(e.G.P[0], cx(l)) * pick(e.G.P[1], cy(l)) * pick(e.G.P[2], cz(l));
}
__device__ inline double eq_g(const EqSet& e, int l) { return geq_dir(e.G, l, e.rho, e.Ushift); }

// One reconstruction: the unknown directions `unk` take the boundary set `b`
// in the local moments; then every direction becomes
//   tgt + f - loc (known)   or   tgt + b - loc (unknown),
// with tgt = b (outflow, wall) or the equilibrium of the local density at the
// boundary velocity and temperature (inflow, `inflow` set).
__device__ inline void tms_reconstruct(const EqSet& b, uint32_t unk, bool inflow, const double u_b[3], double T_b,
                                       double R, double cv, double f[Q], double g[Q])
{
    double r = 0, mx = 0, my = 0, mz = 0, e = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const bool un = (unk >> l) & 1u;
        const double a = un ? eq_f(b, l) : f[l];
        const double q = un ? eq_g(b, l) : g[l];
        r += a;
        if (cx(l)) mx += a * cx(l);
        if (cy(l)) my += a * cy(l);
        if (cz(l)) mz += a * cz(l);
        e += q;
    }
    const double ul[3] = {mx / r, my / r, mz / r};
    const double T = (e / r - 0.5 * (ul[0] * ul[0] + ul[1] * ul[1] + ul[2] * ul[2])) / cv;
    const EqSet loc = eq_set(r, ul, T, R, cv);
    const EqSet tgt = inflow ? eq_set(r, u_b, T_b, R, cv) : b;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const bool un = (unk >> l) & 1u;
        const double fk = un ? eq_f(b, l) : f[l];
        const double gk = un ? eq_g(b, l) : g[l];
        f[l] = (eq_f(tgt, l) + fk) - eq_f(loc, l);
        g[l] = (eq_g(tgt, l) + gk) - eq_g(loc, l);
    }
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_tms(SingleView<P> v, const int* cells, int count)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t];
    const Ijk x = ijk_of(v.d, c);
    double* const* F = v.fld;
    const double R = v.R, cv = v.cv;
    // unknown directions per boundary kind and the first interface direction
    // (upstream of the kind, downstream fluid) of the outflow / inflow sweeps
    uint32_t unkO = 0, unkI = 0, unkS = 0, wall = 0;
    int lo = -1, li = -1;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const uint8_t tu = v.type[cell_of(v.d, x.i - cx(l), x.j - cy(l), x.k - cz(l))];
        const bool dnF = v.type[cell_of(v.d, x.i + cx(l), x.j + cy(l), x.k + cz(l))] == TYPE_F;
        if (tu == TYPE_O) {
            unkO |= 1u << l;
            if (dnF && lo < 0) lo = l;
        } else if (tu == TYPE_I) {
            unkI |= 1u << l;
            if (dnF && li < 0) li = l;
        } else if (tu == TYPE_S) {
            unkS |= 1u << l;
            if (dnF) wall |= 1u << l;
        }
    }
    double f[Q], g[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        f[l] = double(ldp(v.f, l, v.d.n, c)) + v.fref_d(l);
        g[l] = double(ldp(v.g, l, v.d.n, c)) + v.gref_d(l);
    }
    if (lo >= 0) {
        // velocity and temperature downstream, pressure of the outflow ghost
        const int dn = cell_of(v.d, x.i + cx(lo), x.j + cy(lo), x.k + cz(lo));
        const int up = cell_of(v.d, x.i - cx(lo), x.j - cy(lo), x.k - cz(lo));
        const double u[3] = {F[LBM_FIELD_U][dn], F[LBM_FIELD_V][dn], F[LBM_FIELD_W][dn]};
        const double T = F[LBM_FIELD_T][dn];
        const EqSet b = eq_set(F[LBM_FIELD_P][up] / (R * T), u, T, R, cv);
        tms_reconstruct(b, unkO, false, u, T, R, cv, f, g);
    }
    if (li >= 0) {
        // velocity and temperature of the inflow ghost, pressure downstream
        const int dn = cell_of(v.d, x.i + cx(li), x.j + cy(li), x.k + cz(li));
        const int up = cell_of(v.d, x.i - cx(li), x.j - cy(li), x.k - cz(li));
        const double u[3] = {F[LBM_FIELD_U][up], F[LBM_FIELD_V][up], F[LBM_FIELD_W][up]};
        const double T = F[LBM_FIELD_T][up];
        const EqSet b = eq_set(F[LBM_FIELD_P][dn] / (R * T), u, T, R, cv);
        tms_reconstruct(b, unkI, true, u, T, R, cv, f, g);
    }
    if (wall) {
        // wall state at q = 1/2 between the cell downstream and the wall cell,
        // averaged over the interface links; the cell's own density
        double u[3] = {0, 0, 0}, T = 0, rbb = 0;
        int nd = 0;
        const double q = 0.5;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            rbb += f[l];
            if (!((wall >> l) & 1u)) continue;
            ++nd;
            const int dn = cell_of(v.d, x.i + cx(l), x.j + cy(l), x.k + cz(l));
            const int up = cell_of(v.d, x.i - cx(l), x.j - cy(l), x.k - cz(l));
            u[0] += (q * F[LBM_FIELD_U][dn] + F[LBM_FIELD_U][up]) / (1.0 + q);
            u[1] += (q * F[LBM_FIELD_V][dn] + F[LBM_FIELD_V][up]) / (1.0 + q);
            u[2] += (q * F[LBM_FIELD_W][dn] + F[LBM_FIELD_W][up]) / (1.0 + q);
            T += (q * F[LBM_FIELD_T][dn] + F[LBM_FIELD_T][up]) / (1.0 + q);
        }
        const double inv = 1.0 / nd;
        u[0] = inv * u[0];
        u[1] = inv * u[1];
        u[2] = inv * u[2];
        T = inv * T;
        const EqSet b = eq_set(rbb, u, T, R, cv);
        tms_reconstruct(b, unkS, false, u, T, R, cv, f, g);
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        stp(v.f, l, v.d.n, c, P(f[l] - v.fref_d(l)));
        stp(v.g, l, v.d.n, c, P(g[l] - v.gref_d(l)));
    }
}

// CONJUGATE part of TMS_BC (lbm-bc.cpp:1230-1346) for one solid interface
// cell: the target is the state at rest with the interface temperature
// interpolated at q = 1/2 (averaged over the links with gas upstream and solid
// downstream) and the cell's own density; every direction with gas upstream
// is unknown.  Only g is rewritten (f keeps its value).
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_tms_conj(SingleView<P> v, const int* cells, int count)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t];
    const Ijk x = ijk_of(v.d, c);
    const double* Tf = v.fld[LBM_FIELD_T];
    uint32_t unk = 0;
    double T = 0, rbb = 0;
    int nd = 0;
    const double q = 0.5;
    double f[Q], g[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        f[l] = double(ldp(v.f, l, v.d.n, c)) + v.fref_d(l);
        g[l] = double(ldp(v.g, l, v.d.n, c)) + v.gref_d(l);
        rbb += f[l];
        const int up = cell_of(v.d, x.i - cx(l), x.j - cy(l), x.k - cz(l));
        const uint8_t tu = v.type[up];
        if (tu != TYPE_F && tu != TYPE_Q) continue;
        unk |= 1u << l;
        const int dn = cell_of(v.d, x.i + cx(l), x.j + cy(l), x.k + cz(l));
        if (v.type[dn] != TYPE_S) continue;
        ++nd;
        T += (q * Tf[dn] + Tf[up]) / (1.0 + q);
    }
    if (nd == 0) return;
    T = 1.0 / nd * T;
    const double zero[3] = {0, 0, 0};
    const EqSet b = eq_set(rbb, zero, T, v.R, v.cv);
    double fk[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) fk[l] = f[l];
    tms_reconstruct(b, unk, false, zero, T, v.R, v.cv, fk, g);
#pragma unroll
    for (int l = 0; l < Q; ++l) stp(v.g, l, v.d.n, c, P(g[l] - v.gref_d(l)));
}

// ---------------------------------------------------------------- launchers
static inline dim3 grid_of(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

template <typename P>
void launch_init(const SingleView<P>& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_init<P>, grid_of(v.d.n), dim3(BLOCK), 0, s, v);
}
template <typename P>
void launch_collide_explicit(const SingleView<P>& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_collide_explicit<P>, grid_of(v.d.n), dim3(BLOCK), 0, s, v);
}
template <typename P>
void launch_stream(const SingleView<P>& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_stream<P>, grid_of(v.d.n), dim3(BLOCK), 0, s, v);
}
template <typename P>
void launch_moments(const SingleView<P>& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_moments<P>, grid_of(v.d.n), dim3(BLOCK), 0, s, v);
}
template <typename P>
void launch_fused_collide(const SingleView<P>& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fused_collide<P>, dim3(xcd_grid(v.d.n, BLOCK)), dim3(BLOCK), 0, s, v);
}
template <typename P>
void launch_fused_moments(const SingleView<P>& v, bool materialize, hipStream_t s)
{
    if (materialize)
        hipLaunchKernelGGL((k_fused_moments<P, true>), dim3(xcd_grid(v.d.n, BLOCK)), dim3(BLOCK), 0, s, v);
    else
        hipLaunchKernelGGL((k_fused_moments<P, false>), dim3(xcd_grid(v.d.n, BLOCK)), dim3(BLOCK), 0, s, v);
}

template <typename P>
void launch_tms_conj(const SingleView<P>& v, const int* cells, int count, hipStream_t s)
{
    if (count > 0) hipLaunchKernelGGL(k_tms_conj<P>, grid_of(count), dim3(BLOCK), 0, s, v, cells, count);
}

template <typename P>
void launch_tms(const SingleView<P>& v, const int* cells, int count, hipStream_t s)
{
    if (count > 0) hipLaunchKernelGGL(k_tms<P>, grid_of(count), dim3(BLOCK), 0, s, v, cells, count);
}

int diag_partials(const Dims& d) { return (int)((d.n + BLOCK - 1) / BLOCK); }

void launch_diag_partials(const Dims& d, const uint8_t* type, double* const* fld, double gamma, double* partials,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_diag, grid_of(d.n), dim3(BLOCK), 0, s, d, type, fld[LBM_FIELD_RHO], fld[LBM_FIELD_U],
                       fld[LBM_FIELD_V], fld[LBM_FIELD_W], fld[LBM_FIELD_T], fld[LBM_FIELD_RHOE], fld[LBM_FIELD_P],
                       gamma, partials);
}

void launch_diag_final(const double* partials, int npart, double* out8, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(BLOCK), 0, s, partials, npart, out8);
}

void launch_diag(const Dims& d, const uint8_t* type, double* const* fld, double gamma, double* partials, int npart,
                 double* out8, hipStream_t s)
{
    launch_diag_partials(d, type, fld, gamma, partials, s);
    launch_diag_final(partials, npart, out8, s);
}

#define LBM_INSTANTIATE(P)                                                                 \
    template void launch_init<P>(const SingleView<P>&, hipStream_t);                       \
    template void launch_collide_explicit<P>(const SingleView<P>&, hipStream_t);           \
    template void launch_stream<P>(const SingleView<P>&, hipStream_t);                     \
    template void launch_moments<P>(const SingleView<P>&, hipStream_t);                    \
    template void launch_fused_collide<P>(const SingleView<P>&, hipStream_t);              \
    template void launch_fused_moments<P>(const SingleView<P>&, bool, hipStream_t);        \
    template void launch_tms<P>(const SingleView<P>&, const int*, int, hipStream_t);                 \
    template void launch_tms_conj<P>(const SingleView<P>&, const int*, int, hipStream_t);
LBM_INSTANTIATE(double)
LBM_INSTANTIATE(float)

}  // namespace lbm
