// Multicomponent reacting D3Q27 kernels for gfx950 (the reference's
// MULTICOMP path), one kernel per reference call.
//
// Storage: species populations [NS][27][ncell], mixture energy populations
// [27][ncell], per-species fields [NS][ncell] (X, rho_a) and [3][NS][ncell]
// (species velocity), mixture fields as in the single-component path.  One
// thread per cell; species are processed one after another so that only 27
// populations are live in registers at a time.
//
// Reference routines restated here:
//   Init             lbm-init.cpp:85-146
//   Collide_Species  lbm-collision.cpp:334-526
//   Collide          lbm-collision.cpp:191-332
//   Streaming        lbm-streaming.cpp:3-119
//   calculate_moment lbm-moment.cpp:193-377
#include <type_traits>

#include "chem_device.h"
#include "lbm_multi.h"
#include "surrogate_device.h"

namespace lbm {

using namespace chem;

namespace {

constexpr int BLOCK = 256;

struct Ijk {
    int i, j, k;
};
__device__ inline Ijk ijk(const Dims& d, int c)
{
    const int k = c % d.nz;
    const int t = c / d.nz;
    return {t / d.ny, t % d.ny, k};
}
__device__ inline int cell(const Dims& d, int i, int j, int k) { return (i * d.ny + j) * d.nz + k; }

// Element i of a plane as a wave-uniform base plus a 32-bit byte offset: the
// access compiles to one global load or store with a scalar base and a
// 32-bit VGPR offset, not per-lane 64-bit address arithmetic (grids hold
// fewer than 2^29 cells, so a plane of doubles stays below 2^32 bytes).
template <typename T>
__device__ __forceinline__ T& at32(T* base, int i)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (uint32_t)i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, int i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint32_t)i * (uint32_t)sizeof(T));
}

__device__ inline bool fluidlike(uint8_t t)
{
    return t == TYPE_F || t == TYPE_I || t == TYPE_O || t == TYPE_I_C || t == TYPE_O_C;
}

template <typename P>
__device__ inline P& spop(P* base, int a, int l, int64_t n, int c)
{
    return base[((int64_t)a * Q + l) * n + c];
}

// Population values.  FP32 storage holds deviations from the rest-state
// equilibria fref [NS][27] and gref [27]; FP64 storage holds the values.
template <typename P>
__device__ inline double fshift(const MultiView<P>& v, int a, int l)
{
    if constexpr (sizeof(P) == 4) return v.fref[a * Q + l];
    else return 0.0;
}
template <typename P>
__device__ inline double gshift(const MultiView<P>& v, int l)
{
    if constexpr (sizeof(P) == 4) return v.gref[l];
    else return 0.0;
}
template <typename P>
__device__ inline double ldf(const MultiView<P>& v, const P* base, int a, int l, int c)
{
    const double x = (double)base[((int64_t)a * Q + l) * v.d.n + c];
    if constexpr (sizeof(P) == 4) return x + fshift(v, a, l);
    else return x;
}
template <typename P>
__device__ inline void stf(const MultiView<P>& v, P* base, int a, int l, int c, double x)
{
    if constexpr (sizeof(P) == 4) x -= fshift(v, a, l);
    base[((int64_t)a * Q + l) * v.d.n + c] = (P)x;
}
template <typename P>
__device__ inline double ldg(const MultiView<P>& v, const P* base, int l, int c)
{
    const double x = (double)base[(int64_t)l * v.d.n + c];
    if constexpr (sizeof(P) == 4) return x + gshift(v, l);
    else return x;
}
template <typename P>
__device__ inline void stg(const MultiView<P>& v, P* base, int l, int c, double x)
{
    if constexpr (sizeof(P) == 4) x -= gshift(v, l);
    base[(int64_t)l * v.d.n + c] = (P)x;
}

// product-form equilibrium of one direction from the per-axis factors
template <typename T>
__device__ inline T feq_dir(const Axis3<T> F[3], int l, T rho)
{
    return rho * pick(F[0], cx(l)) * pick(F[1], cy(l)) * pick(F[2], cz(l));
}

template <typename T>
__device__ inline void feq_axes(Axis3<T> F[3], const T u[3], T theta, const T corr[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = psi_axis<T>(u[d], theta + u[d] * u[d] + corr[d]);
}

// ------------------------------------------------- fused mode (push)
// Species streaming (lbm-streaming.cpp:3-119) read backwards.  The pull of
// cell d in direction l takes the post-collision population (wrap(d - e_l),
// l) when the upstream cell d - e_l is not a wall or an in/outflow ghost, the
// bounce-back (d, opp(l)) when it is TYPE_S or TYPE_A, and nothing when it is
// an in/outflow ghost: that slot keeps its value.  Free-slip neighbours are
// not handled (the fused mode refuses them).
__device__ inline bool pullable(uint8_t t)
{
    return !(t == TYPE_S || t == TYPE_A || t == TYPE_FS || t == TYPE_O || t == TYPE_O_C || t == TYPE_I ||
             t == TYPE_I_C);
}
__device__ inline int wrap1(int i, int n) { return i == 0 ? n - 2 : (i == n - 1 ? 1 : i); }

// The wrapped coordinates x - 1, x, x + 1 of a cell along each axis.
struct Wrap3 {
    int i[3], j[3], k[3];
    __device__ Wrap3(const Dims& d, const Ijk& x)
        : i{wrap1(x.i - 1, d.nx), x.i, wrap1(x.i + 1, d.nx)},
          j{wrap1(x.j - 1, d.ny), x.j, wrap1(x.j + 1, d.ny)},
          k{wrap1(x.k - 1, d.nz), x.k, wrap1(x.k + 1, d.nz)}
    {
    }
    // cell x + s e_l (s = +1 or -1), wrapped
    __device__ int at(const Dims& d, int l, int s) const
    {
        return cell(d, i[s * cx(l) + 1], j[s * cy(l) + 1], k[s * cz(l) + 1]);
    }
};

// In-place mode: the cell whose slot opp(l) holds the incoming population l
// of cell c in layout 1, c - e_l, wrapped when that cell is a periodic ghost
// (walls and in/outflow ghosts keep their own slots).
template <typename P>
__device__ inline int aa_source(const MultiView<P>& v, int c, int l)
{
    const Ijk x = ijk(v.d, c);
    const int i = x.i - cx(l), j = x.j - cy(l), k = x.k - cz(l);
    const int s = cell(v.d, i, j, k);
    if (v.type[s] != TYPE_P) return s;
    return cell(v.d, wrap1(i, v.d.nx), wrap1(j, v.d.ny), wrap1(k, v.d.nz));
}

// Slot of the incoming species population (a, l) of cell c in the view's
// layout (always layout 0 outside the in-place mode).
template <typename P>
__device__ inline int64_t in_slot(const MultiView<P>& v, int a, int l, int c)
{
    const int64_t n = v.d.n;
    if (!v.swapped) return ((int64_t)a * Q + l) * n + c;
    return ((int64_t)a * Q + opp(l)) * n + aa_source(v, c, l);
}

// Incoming species population (a, l) of cell c (buffer v.f, view's layout).
// The FP32 rest-state shifts are symmetric in l <-> opp(l), so a stored value
// means the same in either slot.
template <typename P>
__device__ inline double ldf_in(const MultiView<P>& v, int a, int l, int c)
{
    const double x = (double)v.f[in_slot(v, a, l, c)];
    if constexpr (sizeof(P) == 4) return x + fshift(v, a, l);
    else return x;
}
template <typename P>
__device__ inline void stf_in(const MultiView<P>& v, int a, int l, int c, double x)
{
    if constexpr (sizeof(P) == 4) x -= fshift(v, a, l);
    v.f[in_slot(v, a, l, c)] = (P)x;
}

// Writes the post-collision value pv of (c, l) of one species (A, B: the
// species' planes in the source and destination buffers) to every slot the
// pull would fill from it, and keeps the slot (c, l) of a cell whose
// upstream is an in/outflow ghost as the pull does.
template <typename P>
__device__ inline void push_general(const MultiView<P>& v, const P* A, P* B, int c, const Ijk& x, int l, P pv)
{
    const int64_t n = v.d.n;
    const uint8_t tt = v.type[cell(v.d, x.i + cx(l), x.j + cy(l), x.k + cz(l))];
    if (tt == TYPE_A || tt == TYPE_S) B[(int64_t)opp(l) * n + c] = pv;
    const int di = wrap1(x.i + cx(l), v.d.nx), dj = wrap1(x.j + cy(l), v.d.ny), dk = wrap1(x.k + cz(l), v.d.nz);
    const int dc = cell(v.d, di, dj, dk);
    if (v.type[dc] == TYPE_F && pullable(v.type[cell(v.d, di - cx(l), dj - cy(l), dk - cz(l))]))
        B[(int64_t)l * n + dc] = pv;
    const uint8_t tu = v.type[cell(v.d, x.i - cx(l), x.j - cy(l), x.k - cz(l))];
    if (tu == TYPE_O || tu == TYPE_O_C || tu == TYPE_I || tu == TYPE_I_C) B[(int64_t)l * n + c] = A[(int64_t)l * n + c];
}

// Post-collision species population in storage form (shared by the strict
// and the fused collisions, so that both round alike).
template <typename P>
__device__ __forceinline__ P species_post(const MultiView<P>& v, int a, int l, P fl, double om, double fe, double fdu,
                                          double fr)
{
    if constexpr (sizeof(P) == 4) {
        // the same update on the stored deviation f - shift:
        // (1 - om) (fl - fe) + fdu + fr - om shift, five operations for eight
        const double x = (1.0 - om) * ((double)fl - fe) + (fdu + fr - om * fshift(v, a, l));
        return (P)x;
    } else {
        const double f = (double)fl;
        return f + om * (fe - f) + (fdu - fe) + fr;
    }
}

}  // namespace

// ------------------------------------------------------------------ Init
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_init(MultiView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n || !fluidlike(v.type[c])) return;
    const int64_t n = v.d.n;
    double X[NS], Y[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) X[a] = v.X[a * n + c];
    double s = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a) s += (X[a] > 0 ? X[a] : 0.0);
#pragma unroll
    for (int a = 0; a < NS; ++a) X[a] = (X[a] > 0 ? X[a] : 0.0) / s;
    Y_from_X(X, Y);
    const double Tsi = v.u.si_temp(v.fld[LBM_FIELD_T][c]);
    const double psi = v.u.si_p(v.fld[LBM_FIELD_P][c]);
    const double Wm = mean_W(Y);
    const double rho = v.u.rho(psi * Wm / (R_UNIV * Tsi));
    const double vel[3] = {v.fld[LBM_FIELD_U][c], v.fld[LBM_FIELD_V][c], v.fld[LBM_FIELD_W][c]};
    const double U = v.u.energy_mass(u_mass(Tsi, Y));
    v.fld[LBM_FIELD_RHO][c] = rho;
    v.fld[LBM_FIELD_RHOE][c] = rho * (U + 0.5 * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]));
    const double theta = v.u.energy_mass(R_UNIV * Tsi / Wm);
    const GeqAxes<double> G = geq_axes<double>(theta, vel);
#pragma unroll
    for (int l = 0; l < Q; ++l) stg(v, v.go, l, c, geq_dir<double>(G, l, rho, U - 1.5 * theta));
    const double zero[3] = {0, 0, 0};
    for (int a = 0; a < NS; ++a) {
        const double yr = Y[a] * rho;
        const double ra = yr > SPECIES_MIN ? yr : SPECIES_MIN;
        v.rhoa[a * n + c] = ra;
#pragma unroll
        for (int d = 0; d < 3; ++d) v.ua[(d * NS + a) * n + c] = vel[d];
        const double th = v.u.energy_mass(R_UNIV * Tsi / W[a]);
        Axis3<double> F[3];
        feq_axes<double>(F, vel, th, zero);
#pragma unroll
        for (int l = 0; l < Q; ++l) stf(v, v.fo, a, l, c, feq_dir<double>(F, l, ra));
    }
}

// ------------------------------------------------------------ chemistry
__device__ inline unsigned mix_hash(unsigned c, unsigned step)
{
    unsigned h = c * 0x9E3779B1u ^ (step * 0x85EBCA77u + 0x165667B1u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

// Thermochemical state of a cell as Collide_Species sees it
// (lbm-collision.cpp:350-369): species densities, normalised mass fractions,
// SI temperature and density, and the SI internal energy without the
// species diffusion kinetic energy.
struct ChemState {
    double rho, rho_si, Tsi, usi;
    double ra[NS], Y0[NS];
};

template <typename P>
__device__ inline void chem_state(const MultiView<P>& v, int c, bool with_energy, ChemState& s)
{
    const int64_t n = v.d.n;
    s.rho = v.fld[LBM_FIELD_RHO][c];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        s.ra[a] = v.rhoa[a * n + c];
        s.Y0[a] = s.ra[a] / s.rho;
    }
    normalize(s.Y0);
    s.Tsi = v.u.si_temp(v.fld[LBM_FIELD_T][c]);
    s.rho_si = v.u.si_rho(s.rho);
    s.usi = 0.0;
    if (with_energy) {
        double Uint = v.fld[LBM_FIELD_RHOE][c] / s.rho;
        for (int a = 0; a < NS; ++a) {
            const double ux = v.ua[(0 * NS + a) * n + c], uy = v.ua[(1 * NS + a) * n + c],
                         uz = v.ua[(2 * NS + a) * n + c];
            Uint = Uint - 0.5 * s.ra[a] / s.rho * (ux * ux + uy * uy + uz * uz);
        }
        s.usi = v.u.si_energy_mass(Uint);
    }
}

// The reference's chemistry source (lbm-collision.cpp:370-393): two explicit
// Euler sub-steps of dt/2 of the net production rates, each followed by a
// UV temperature re-solve at the new composition.  src[a] is the species
// mass added over the step, *Tout the final temperature [K].
template <typename P>
__device__ inline void plain_chemistry(const MultiView<P>& v, const ChemState& s, double src[NS], double* Tout)
{
    double Tcur = s.Tsi, Ycur[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        Ycur[a] = s.Y0[a];
        src[a] = 0.0;
    }
    for (int sub = 0; sub < 2; ++sub) {
        double wdot[NS];
        net_production(Tcur, s.rho_si, Ycur, wdot);
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            src[a] += 1.0 / 2.0 * (double)v.u.rho_dot(wdot[a] * W[a]);
            Ycur[a] = (s.ra[a] + src[a]) / s.rho;
        }
        normalize(Ycur);
        Tcur = T_from_UV(s.usi, Ycur, Tcur);
    }
    *Tout = Tcur;
}

// HRR = -sum_a h_a(T) rho_dot_a (lbm-collision.cpp:395-399)
template <typename P>
__device__ inline double heat_release(const MultiView<P>& v, const double src[NS], double T)
{
    double hrr = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a) hrr += -1.0 * (double)v.u.energy_mass(h_RT(a, T) * R_UNIV * T / W[a]) * src[a];
    return hrr;
}

template <typename P>
__device__ inline void store_chem(const MultiView<P>& v, int c, const double src[NS], double T)
{
    const int64_t n = v.d.n;
#pragma unroll
    for (int a = 0; a < NS; ++a) v.src[a * n + c] = src[a];
    v.hrr[c] = heat_release(v, src, T);
    if (v.tchem) v.tchem[c] = T;
}

// Appends one (state, result) row of the plain integrator to the surrogate's
// training set (raw features T, p, Y_k, dt; targets dT, dY_k), while it has room.
template <typename P>
__device__ inline void capture_row(const MultiView<P>& v, const ChemState& s, const double src[NS], double Tcur)
{
    const unsigned long long slot = atomicAdd(v.cap_count, 1ull);
    if ((long long)slot >= v.cap_max) return;
    float* row = v.cap_buf + slot * sur::CAP_STRIDE;
    row[0] = (float)s.Tsi;
    row[1] = (float)pressure(s.Tsi, s.rho_si, s.Y0);
#pragma unroll
    for (int a = 0; a < NS; ++a) row[2 + a] = (float)s.Y0[a];
    row[sur::NIN - 1] = v.u.s;
#pragma unroll
    for (int q = sur::NIN; q < 16; ++q) row[q] = 0.f;
    // dT from the exact temperature of the initial state (Tsi carries the
    // binary32 rounding of the lattice-unit crossing)
    row[16] = (float)(Tcur - T_from_UV(s.usi, s.Y0, s.Tsi));
#pragma unroll
    for (int a = 0; a < NS; ++a) row[17 + a] = (float)(src[a] / s.rho);
#pragma unroll
    for (int q = 16 + sur::NOUT; q < 32; ++q) row[q] = 0.f;
}

// Plain chemistry of every fluid cell; optionally samples (state, result)
// rows for the surrogate's training set.
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_chemistry(MultiView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    ChemState s;
    chem_state(v, c, true, s);
    double src[NS], Tcur;
    plain_chemistry(v, s, src, &Tcur);
    store_chem(v, c, src, Tcur);
    if (v.cap_buf && mix_hash(c, v.step) % (unsigned)v.cap_stride == 0) capture_row(v, s, src, Tcur);
}

// Surrogate chemistry in three passes, so that the FP64 feature and decode
// work runs at full occupancy and the network pass is pure matrix-core work:
//   k_mc_sur_features: normalised BF16 features [n][16] and envelope flags;
//   sur::mlp_infer_bf16: the network, outputs [n][16] FP32;
//   k_mc_sur_decode: sources, HRR and final temperature; cells outside the
//   envelope and audited 32-cell chunks go to the plain integrator's list.
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_sur_features(MultiView<P> v, uint16_t* X, uint8_t* flag)
{
    using namespace sur;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    const float* st = v.sur_stats;
    float raw[NIN] = {};
    const bool mine = v.type[c] == TYPE_F;
    if (mine) {
        ChemState s;
        chem_state(v, c, false, s);
        raw[0] = (float)s.Tsi;
        raw[1] = (float)pressure(s.Tsi, s.rho_si, s.Y0);
#pragma unroll
        for (int a = 0; a < NS; ++a) raw[2 + a] = (float)s.Y0[a];
        raw[NIN - 1] = v.u.s;
    }
    bool outside = false;
    uint32_t packed[8];
#pragma unroll
    for (int q = 0; q < 16; q += 2) {
        float x0 = 0.f, x1 = 0.f;
        if (mine && q < NIN) {
            outside |= raw[q] < st[SS_LO + q] || raw[q] > st[SS_HI + q];
            x0 = (raw[q] - st[SS_FMEAN + q]) / st[SS_FSTD + q];
        }
        if (mine && q + 1 < NIN) {
            outside |= raw[q + 1] < st[SS_LO + q + 1] || raw[q + 1] > st[SS_HI + q + 1];
            x1 = (raw[q + 1] - st[SS_FMEAN + q + 1]) / st[SS_FSTD + q + 1];
        }
        packed[q / 2] = (uint32_t)f2bf(x0) | ((uint32_t)f2bf(x1) << 16);
    }
    uint4* dst = reinterpret_cast<uint4*>(X + (int64_t)c * 16);
    dst[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    dst[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
    flag[c] = outside ? 1 : 0;
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_sur_decode(MultiView<P> v, const float* Y, const uint8_t* flag,
                                                         unsigned* list, unsigned* count)
{
    using namespace sur;
    __shared__ unsigned blk_used, blk_fb;
    const int64_t n = v.d.n;
    const float* st = v.sur_stats;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) blk_used = blk_fb = 0u;
    __syncthreads();
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    const bool mine = c < n && v.type[c] == TYPE_F;
    bool audit = false, fallback = false;
    if (mine) {
        ChemState s;
        chem_state(v, c, false, s);
        const float4* pr4 = reinterpret_cast<const float4*>(Y + (int64_t)c * 16);
        const float4 y0 = pr4[0], y1 = pr4[1], y2 = pr4[2];
        static_assert(!sur::ENABLED || NOUT <= 12, "three float4 loads cover the real outputs");
        const float pr[12] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w, y2.x, y2.y, y2.z, y2.w};
        // the imbalance of the predicted sources is removed in proportion to each
        // species' target scale (an inert species, scale 0, keeps its mean)
        double src[NS], sum = 0, wsum = 0;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            src[a] = s.rho * ((double)pr[1 + a] * st[SS_TSTD + 1 + a] + st[SS_TMEAN + 1 + a]);
            sum += src[a];
            wsum += st[SS_TSTD + 1 + a];
        }
#pragma unroll
        for (int a = 0; a < NS; ++a) src[a] -= st[SS_TSTD + 1 + a] / wsum * sum;  // no net mass
        const double Tcur = s.Tsi + ((double)pr[0] * st[SS_TSTD] + st[SS_TMEAN]);
        const bool finite = isfinite(sum) && isfinite(Tcur) && Tcur > 0.0;
        fallback = flag[c] != 0 || !finite;
        if (!finite) {
#pragma unroll
            for (int a = 0; a < NS; ++a) src[a] = 0.0;
        }
        store_chem(v, c, src, finite ? Tcur : s.Tsi);
        audit = !fallback && v.audit_stride > 0 && mix_hash((unsigned)(c >> 5), v.step) % (unsigned)v.audit_stride == 0;
    }
    // wavefront-aggregated append of the cells the plain integrator must see
    const unsigned long long need = __ballot(fallback || audit);
    const unsigned used = (unsigned)__popcll(__ballot(mine && !fallback));
    const unsigned fb = (unsigned)__popcll(__ballot(fallback));
    unsigned slot0 = 0;
    if (lane == 0) {
        if (need) slot0 = atomicAdd(count, (unsigned)__popcll(need));
        if (used) atomicAdd(&blk_used, used);
        if (fb) atomicAdd(&blk_fb, fb);
    }
    slot0 = __shfl(slot0, 0);
    if (fallback || audit) {
        const unsigned below = (unsigned)__popcll(need & ((1ull << lane) - 1ull));
        list[slot0 + below] = (unsigned)c | (audit ? CHEM_AUDIT : 0u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // one update per block
        if (blk_used) atomicAdd(&v.sur_count[SUR_PREDICTED], (unsigned long long)blk_used);
        if (blk_fb) atomicAdd(&v.sur_count[SUR_FALLBACK], (unsigned long long)blk_fb);
    }
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_chem_list(MultiView<P> v, const unsigned* list, const unsigned* count)
{
    using namespace sur;
    const unsigned total = *count;
    const int64_t n = v.d.n;
    const float* st = v.sur_stats;
    unsigned long long n_aud = 0, n_fail = 0;
    float emax = 0.f;
    for (unsigned e = blockIdx.x * BLOCK + threadIdx.x; e < total; e += gridDim.x * BLOCK) {
        const unsigned ent = list[e];
        const int c = (int)(ent & ~CHEM_AUDIT);
        ChemState s;
        chem_state(v, c, true, s);
        double src[NS], Tp;
        plain_chemistry(v, s, src, &Tp);
        // while capturing, the states the network could not take (and the
        // audited ones) join the training set
        if (v.cap_buf) capture_row(v, s, src, Tp);
        if (!(ent & CHEM_AUDIT)) {
            store_chem(v, c, src, Tp);
            continue;
        }
        // errors in target standard deviations; the temperature's scale is at
        // least 1e-6 T (the binary32 rounding of the initial temperature)
        const double tscale = fmax((double)st[SS_TSTD], 1e-6 * s.Tsi);
        float err = (float)(fabs(Tp - v.tchem[c]) / tscale);
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            if (!(st[SS_TSTD + 1 + a] > 0.f)) continue;  // no spread: not predicted
            const float ea = (float)(fabs(src[a] - v.src[a * n + c]) / (s.rho * st[SS_TSTD + 1 + a]));
            err = ea > err ? ea : err;
        }
        if (!(err == err)) err = INFINITY;
        ++n_aud;
        emax = err > emax ? err : emax;
        if (err > v.audit_tol) {  // the audit overrules the network for this cell
            ++n_fail;
            store_chem(v, c, src, Tp);
        }
    }
    // one update per wavefront
    for (int o = 32; o > 0; o >>= 1) {
        n_aud += __shfl_down(n_aud, o, 64);
        n_fail += __shfl_down(n_fail, o, 64);
        emax = fmaxf(emax, __shfl_down(emax, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && n_aud) {
        atomicAdd(&v.sur_count[SUR_AUDITED], n_aud);
        atomicAdd(&v.sur_count[SUR_AUDIT_FAIL], n_fail);
        atomicMax(&v.sur_count[SUR_MAX_ERR], (unsigned long long)__float_as_uint(emax));
    }
}

// ------------------------------------------------------- Collide_Species
// Two kernels.  The transport pass couples the species of a cell: species
// viscosities, Wilke weights and omega_a (lbm-collision.cpp:405-429), binary
// diffusivities and the Maxwell-Stefan friction impulse (:433-470); it stores
// omega_a and the impulse per species in v.tsp.  The collide pass then runs
// one thread per (cell, species) (:452-520) with a small register footprint;
// its per-species inputs are the species' own planes.
// The mixture collide's relaxation rates (lbm-collision.cpp:241-252) from
// the mixture viscosity mu_si and conductivity kap_si [SI] of the cell's
// state, into v.omix: omega = 2 p / (p + 2 mu), omega_1 = 2 p c_p / (p c_p +
// 2 kappa) in lattice units.
template <typename P>
__device__ __forceinline__ void mixture_rates(const MultiView<P>& v, int c, double Tsi, const double Y0[NS],
                                              double mu_si, double kap_si)
{
    const double p = v.fld[LBM_FIELD_P][c];
    const double cp = v.u.cp(cp_mass(Tsi, Y0));
    const double mu = v.u.mu(mu_si);
    const double kap = v.u.thermal_conductivity(kap_si);
    v.omix[c] = 2 * p * 1 / (p * 1 + 2 * mu);
    v.omix[v.d.n + c] = 2 * p * cp * 1 / (p * cp * 1 + 2 * kap);
}

// The transport pass of the FP32 path, all in binary32: species
// viscosities and Wilke weights with sqrt(mu_a / mu_b) = sqrt(mu_a) / sqrt(mu_b)
// (9 square roots and reciprocals in place of 45 divisions and square roots),
// and the Maxwell-Stefan friction with one reciprocal per diffusivity.  The
// same quantities as the plain kernel below, to binary32 rounding.
template <typename P>
__device__ __forceinline__ void species_transport_fp32(const MultiView<P>& v, int c)
{
    const int64_t n = v.d.n;
    double* const* Fd = v.fld;
    const double rho = Fd[LBM_FIELD_RHO][c], T = Fd[LBM_FIELD_T][c];
    const float p = (float)Fd[LBM_FIELD_P][c];
    const double irho = 1.0 / rho;
    float ra[NS], Xa[NS];
    double Y0[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        ra[a] = v.rhoa[a * n + c];
        Xa[a] = v.X[a * n + c];
        Y0[a] = ra[a] * irho;
    }
    normalize(Y0);
    const double Tsi = v.u.si_temp(T);
    const double rho_si = v.u.si_rho(rho);
    const float lnT = (float)log(Tsi);
    // lattice units of viscosity and diffusivity: s m / kg and s / m^2
    const float kmu = v.u.s * v.u.m / v.u.kg;
    const float knu = (float)((double)v.u.s / ((double)v.u.m * (double)v.u.m));
    float mu[NS], sm[NS], rsm[NS], den[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        mu[a] = species_viscosity_r<float>(a, lnT);
        sm[a] = sqrtf(mu[a]);
        rsm[a] = rcp_r(sm[a]);
        den[a] = 0.0f;
    }
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = a; b < NS; ++b) {
            const float pab = wilke_phi(a, b, sm[a] * rsm[b]);
            den[a] += Xa[b] * pab;
            if (b != a) den[b] += Xa[a] * (mu[b] * rcp_r(mu[a]) * (float)(W[a] / W[b]) * pab);
        }
    float mu_mix = 0.0f;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const float visc = Xa[a] * (mu[a] * kmu) * rcp_r(den[a]);
        v.tsp[a * n + c] = P(2.0f * Xa[a] * p * rcp_r(Xa[a] * p + 2.0f * visc));
        mu_mix += Xa[a] * mu[a] * rcp_r(den[a]);
    }
    mixture_rates(v, c, Tsi, Y0, (double)mu_mix, (double)mix_conductivity_r<float>(lnT, Xa));
    const float rp_si = rcp_r((float)pressure<true>(Tsi, rho_si, Y0));
    float ua[3][NS], acc[NS][3];
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ua[d][a] = v.ua[(d * NS + a) * n + c];
            acc[a][d] = 0.0f;
        }
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = a + 1; b < NS; ++b) {
            const float D = exp(poly4r<float>(DIFF_FIT[a][b], lnT)) * rp_si * knu;
            const float k = -p * Xa[a] * Xa[b] * rcp_r(D);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float du = ua[d][a] - ua[d][b];
                acc[a][d] += k * du;
                acc[b][d] -= k * du;
            }
        }
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const float ir = rcp_r(ra[a]);
#pragma unroll
        for (int d = 0; d < 3; ++d) v.tsp[((1 + d) * NS + a) * n + c] = P(acc[a][d] * ir);
    }
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_species_transport(MultiView<P> v)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    if constexpr (sizeof(P) == 4) {
        species_transport_fp32(v, c);
        return;
    }
    const int64_t n = v.d.n;
    double* const* Fd = v.fld;
    const double rho = Fd[LBM_FIELD_RHO][c], T = Fd[LBM_FIELD_T][c], p = Fd[LBM_FIELD_P][c];
    double ra[NS], Xa[NS], Y0[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        ra[a] = v.rhoa[a * n + c];
        Xa[a] = v.X[a * n + c];
        Y0[a] = ra[a] / rho;
    }
    normalize(Y0);
    const double Tsi = v.u.si_temp(T);
    const double rho_si = v.u.si_rho(rho);
    // species viscosities and Wilke weights at (T, rho); transport arithmetic
    // R is binary64 on the plain path and binary32 on the FP32 path
    using R = typename std::conditional<sizeof(P) == 4, float, double>::type;
    const double lnT = log(Tsi);
    const R lnTr = (R)lnT;
    R mu[NS], den[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        mu[a] = species_viscosity_r<R>(a, lnTr);
        den[a] = 0;
    }
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = a; b < NS; ++b) {
            const R f1 = R(1) + sqrt(mu[a] / mu[b]) * (R)sqrt(sqrt(W[b] / W[a]));
            const R pab = f1 * f1 / (R)sqrt(8.0 * (1.0 + W[a] / W[b]));
            den[a] += (R)Xa[b] * pab;
            if (b != a) den[b] += (R)Xa[a] * (mu[b] / mu[a] * (R)W[a] / (R)W[b] * pab);
        }
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const double visc = Xa[a] * (double)v.u.mu(mu[a]) / (double)den[a];
        v.tsp[a * n + c] = P(2 * Xa[a] * p / (Xa[a] * p + 2 * visc));
    }

    const double p_si = pressure(Tsi, rho_si, Y0);
    const R p_sir = (R)p_si;
    // Maxwell-Stefan friction with the binary diffusivities of the state.  The
    // fit table is symmetric, so D_ab = D_ba is evaluated once per pair; the
    // b = a term, k (u_a - u_a), is zero and skipped.  Each acc[a] still sums
    // its terms in the order b = 0, 1, ..., NS-1.
    double ua[3][NS], acc[NS][3];
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ua[d][a] = v.ua[(d * NS + a) * n + c];
            acc[a][d] = 0.0;
        }
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = a + 1; b < NS; ++b) {
            const double D = v.u.nu(binary_diff_r<R>(a, b, lnTr, p_sir));
            const double kab = -1.0 * p * Xa[a] * Xa[b] / D;
            const double kba = -1.0 * p * Xa[b] * Xa[a] / D;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                acc[a][d] += kab * (ua[d][a] - ua[d][b]);
                acc[b][d] += kba * (ua[d][b] - ua[d][a]);
            }
        }
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) v.tsp[((1 + d) * NS + a) * n + c] = P(acc[a][d] * 1.0 / ra[a]);
}

// Collide_Species of species a at cell c (lbm-collision.cpp:452-520) from
// its incoming populations fl (storage form).  emit(l, value) receives each
// post-collision population in storage form; keep(l) is called instead where
// the reference leaves the post-collision value as it was (a species at the
// SPECIES_MIN floor without a positive chemistry source, :512-513).
template <typename P, typename Emit, typename Keep>
__device__ __forceinline__ void species_collision(const MultiView<P>& v, int c, int a, const Ijk& x, const P fl[Q],
                                                  Emit emit, Keep keep)
{
    const int64_t n = v.d.n;
    double* const* Fd = v.fld;
    const auto* rhoa = v.rhoa + a * n;
    const double T = at32(Fd[LBM_FIELD_T], c);
    const double umix[3] = {at32(Fd[LBM_FIELD_U], c), at32(Fd[LBM_FIELD_V], c), at32(Fd[LBM_FIELD_W], c)};
    const double ra = at32(rhoa, c);
    const double src = v.reaction ? (double)at32(v.src + a * n, c) : 0.0;
    const double om = at32(v.tsp + a * n, c);
    const double Fa[3] = {at32(v.tsp + (NS + a) * n, c), at32(v.tsp + (2 * NS + a) * n, c),
                          at32(v.tsp + (3 * NS + a) * n, c)};
    const int nm[3] = {cell(v.d, x.i - 1, x.j, x.k), cell(v.d, x.i, x.j - 1, x.k), c - 1};
    const int np[3] = {cell(v.d, x.i + 1, x.j, x.k), cell(v.d, x.i, x.j + 1, x.k), c + 1};
    const double* Tf = Fd[LBM_FIELD_T];
    double mom[3] = {0, 0, 0};
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const double f = (double)fl[l] + fshift(v, a, l);
        if (cx(l)) mom[0] += cx(l) * f;
        if (cy(l)) mom[1] += cy(l) * f;
        if (cz(l)) mom[2] += cz(l) * f;
    }
    // FP32 path: one reciprocal for the three divisions
    constexpr bool FAST = sizeof(P) == 4;
    const double ira = FAST ? 1.0 / ra : 0.0;
    const double uraw[3] = {FAST ? mom[0] * ira : mom[0] / ra, FAST ? mom[1] * ira : mom[1] / ra,
                            FAST ? mom[2] * ira : mom[2] / ra};
    const double udu[3] = {uraw[0] + Fa[0], uraw[1] + Fa[1], uraw[2] + Fa[2]};
    const double Ra = v.u.cp(R_UNIV / W[a]);
    const double theta = Ra * T;
    double dQ[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const auto* uad = v.ua + (d * NS + a) * n;
        auto Qf = [&](int m) {
            const double r = at32(rhoa, m), ud = at32(uad, m);
            return r * ud * (1.0 - 3.0 * Ra * at32(Tf, m)) - r * (ud * ud * ud);
        };
        dQ[d] = fd_central<double>(Qf(nm[d]), Qf(c), Qf(np[d]), at32(v.type, nm[d]), at32(v.type, np[d]));
    }
    double corr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) corr[d] = 1 * (2.0 - om) / (2.0 * ra * om) * dQ[d];
    Axis3<double> Fe[3], Fu[3];
    feq_axes<double>(Fe, uraw, theta, corr);
    feq_axes<double>(Fu, udu, theta, corr);
    // the chemistry source's equilibrium: binary32 on the FP32 path (a
    // correction far below the stored deviations' resolution), binary64 else
    using RS = typename std::conditional<sizeof(P) == 4, float, double>::type;
    const RS zero[3] = {0, 0, 0}, um[3] = {(RS)umix[0], (RS)umix[1], (RS)umix[2]};
    Axis3<RS> Fr[3];
    feq_axes<RS>(Fr, um, (RS)theta, zero);
    const bool low = ra <= SPECIES_MIN;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const double fe = feq_dir<double>(Fe, l, ra);
        const double fdu = feq_dir<double>(Fu, l, ra);
        const double fr = (double)feq_dir<RS>(Fr, l, (RS)src);
        if (low && fr <= 0.0) keep(l);  // post-collision value left as it was
        else emit(l, species_post(v, a, l, fl[l], om, fe, fdu, fr));
    }
}

// grid: xcd_species_grid(n, BLOCK, NS).  PUSH: the fused mode's form, each
// post-collision population written to its streamed slot.
template <typename P, bool PUSH>
__global__ __launch_bounds__(BLOCK) void k_mc_collide_species(MultiView<P> v)
{
    int cb, a;
    xcd_species_block(NS, cb, a);
    const int64_t c64 = (int64_t)cb * BLOCK + threadIdx.x;
    if (c64 >= v.d.n) return;
    const int c = (int)c64;
    if (v.type[c] != TYPE_F) return;
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    const P* A = v.f + (int64_t)a * Q * n;
    P* B = v.fo + (int64_t)a * Q * n;
    P fl[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) fl[l] = at32(A + l * n, c);
    auto none = [](int) {};
    if constexpr (!PUSH) {
        species_collision(v, c, a, x, fl, [&](int l, P pv) { at32(B + l * n, c) = pv; }, none);
    } else {
        // A cell with NB_PUSH streams every population to a fluid cell that
        // TMS_BC leaves alone: the post-collision value a low species keeps
        // is the one streamed there last step.  The other cells collide into
        // the side store as the strict form does into fo; k_mc_side_push
        // then streams them.
        if (!(at32(v.nbflag, c) & NB_PUSH)) {
            P* side = v.fside + (int64_t)a * Q * v.nside;
            const int slot = at32(v.side_slot, c);
            species_collision(v, c, a, x, fl, [&](int l, P pv) { at32(side + l * v.nside, slot) = pv; }, none);
            return;
        }
        const Wrap3 w(v.d, x);
        species_collision(
            v, c, a, x, fl, [&](int l, P pv) { at32(B + l * n, w.at(v.d, l, 1)) = pv; },
            [&](int l) {
                const int dst = w.at(v.d, l, 1);
                at32(B + l * n, dst) = at32(A + l * n, dst);
            });
    }
}

// ------------------------------------------- in-place mode (AA pattern)
// One species buffer.  A step that reads layout SW writes layout 1 - SW:
//   SW = 0: a plain (NB_PUSH) cell reads its own slots (l, x) and writes its
//           post-collision population l to (opp(l), x), where layout 1 keeps
//           the incoming population l of x + e_l;
//   SW = 1: it reads (opp(l), x - e_l) and writes (l, x + e_l), coordinates
//           wrapped through periodic ghosts.
// Each plain cell writes exactly the slots it reads, so the cells of a step
// touch disjoint sets.  Side cells (near walls, in/outflow faces and TMS_BC
// cells) collide into the side store and k_mc_side_push_aa applies the
// strict stream's rules forwards from there.  A low species of a plain cell
// keeps last step's post-collision value, which sits in a slot another cell
// overwrites this step: k_mc_aa_save copies it out first and the listed pairs
// collide in k_mc_collide_species_low.

// The post-collision value a plain cell's population l kept from last step:
// the incoming population l of x + e_l in layout SW.
template <int SW>
__device__ inline int64_t aa_kept_slot(const Dims& d, const Wrap3& w, int c, int l)
{
    return SW ? (int64_t)opp(l) * d.n + c : (int64_t)l * d.n + w.at(d, l, 1);
}

// Incoming populations of a plain cell in layout SW (F: the species' planes)
// and the slot its post-collision population l goes to.
template <typename P, int SW>
__device__ inline void aa_plain_load(const Dims& d, const P* F, const Wrap3& w, int c, P fl[Q])
{
#pragma unroll
    for (int l = 0; l < Q; ++l) fl[l] = SW ? F[(int64_t)opp(l) * d.n + w.at(d, l, -1)] : F[(int64_t)l * d.n + c];
}
template <int SW>
__device__ inline int64_t aa_plain_dst(const Dims& d, const Wrap3& w, int c, int l)
{
    return SW ? (int64_t)l * d.n + w.at(d, l, 1) : (int64_t)opp(l) * d.n + c;
}

template <typename P, int SW>
__global__ __launch_bounds__(BLOCK) void k_mc_aa_save(MultiView<P> v, int first, int a0)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    const int a = a0 + (int)blockIdx.y;
    const int64_t n = v.d.n;
    const bool listed = c < n && v.type[c] == TYPE_F && (v.nbflag[c] & NB_PUSH) && v.rhoa[a * n + c] <= SPECIES_MIN;
    // wavefront-aggregated append
    const unsigned long long m = __ballot(listed);
    if (!m) return;
    const int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
    unsigned slot0 = 0;
    if (lane == leader) slot0 = atomicAdd(v.low_count, (unsigned)__popcll(m));
    slot0 = __shfl(slot0, leader);
    if (!listed) return;
    const int64_t e = slot0 + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    if (e >= v.low_cap) return;
    v.low_cell[e] = c;
    v.low_sp[e] = (uint8_t)a;
    const P* F = v.f + (int64_t)a * Q * n;
    const Wrap3 w(v.d, ijk(v.d, c));
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        P kept;
        if (first) kept = v.pending ? v.pending[((int64_t)a * Q + l) * n + c] : P(0);
        else kept = F[aa_kept_slot<SW>(v.d, w, c, l)];
        v.low_keep[(int64_t)l * v.low_cap + e] = kept;
    }
}

// grid: xcd_species_grid(n, BLOCK, NS).  Plain cells (not listed) and side cells.
template <typename P, int SW>
__global__ __launch_bounds__(BLOCK) void k_mc_collide_species_aa(MultiView<P> v, int a0, int na)
{
    int cb, a;
    xcd_species_block(na, cb, a);
    a += a0;
    const int64_t c64 = (int64_t)cb * BLOCK + threadIdx.x;
    if (c64 >= v.d.n) return;
    const int c = (int)c64;
    if (v.type[c] != TYPE_F) return;
    const int64_t n = v.d.n;
    const bool plain = v.nbflag[c] & NB_PUSH;
    if (plain && v.rhoa[a * n + c] <= SPECIES_MIN) return;  // listed: k_mc_collide_species_low
    const Ijk x = ijk(v.d, c);
    P* F = v.f + (int64_t)a * Q * n;
    P fl[Q];
    auto none = [](int) {};
    if (plain) {
        const Wrap3 w(v.d, x);
        aa_plain_load<P, SW>(v.d, F, w, c, fl);
        species_collision(v, c, a, x, fl, [&](int l, P pv) { F[aa_plain_dst<SW>(v.d, w, c, l)] = pv; }, none);
        return;
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) fl[l] = v.f[in_slot(v, a, l, c)];
    P* side = v.fside + (int64_t)a * Q * v.nside + v.side_slot[c];
    species_collision(v, c, a, x, fl, [&](int l, P pv) { side[(int64_t)l * v.nside] = pv; }, none);
}

// The listed (cell, species) pairs: grid-stride over the list.
template <typename P, int SW>
__global__ __launch_bounds__(BLOCK) void k_mc_collide_species_low(MultiView<P> v)
{
    const int64_t total = min((int64_t)*v.low_count, v.low_cap);
    const int64_t n = v.d.n;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * BLOCK) {
        const int c = v.low_cell[e];
        const int a = v.low_sp[e];
        const Ijk x = ijk(v.d, c);
        const Wrap3 w(v.d, x);
        P* F = v.f + (int64_t)a * Q * n;
        P fl[Q];
        aa_plain_load<P, SW>(v.d, F, w, c, fl);
        species_collision(
            v, c, a, x, fl, [&](int l, P pv) { F[aa_plain_dst<SW>(v.d, w, c, l)] = pv; },
            [&](int l) { F[aa_plain_dst<SW>(v.d, w, c, l)] = v.low_keep[(int64_t)l * v.low_cap + e]; });
    }
}

// Side cells: the strict stream's rules (lbm-streaming.cpp:3-119) applied
// forwards from the side store into layout 1 - SW.  Population m of cell x
// goes to the incoming slot (m, z) of z = wrap(x + e_m) when z is fluid and
// pulls through a plain move; bounces back into the incoming slot (opp(m), x)
// at a TYPE_S / TYPE_A neighbour; an in/outflow neighbour takes nothing.  A
// population of x whose upstream cell is an in/outflow ghost keeps its value.
// grid: (ceil(nside / BLOCK), NS).
template <typename P, int SW>
__global__ __launch_bounds__(BLOCK) void k_mc_side_push_aa(MultiView<P> v)
{
    const int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (s >= v.nside) return;
    const int c = v.side_cell[s];
    const int a = blockIdx.y;
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    const P* S = v.fside + (int64_t)a * Q * v.nside;
    P* F = v.f + (int64_t)a * Q * n;
    for (int m = 0; m < Q; ++m) {
        const P pv = S[(int64_t)m * v.nside + s];
        const int t = cell(v.d, x.i + cx(m), x.j + cy(m), x.k + cz(m));
        const uint8_t tt = v.type[t];
        if (tt == TYPE_S || tt == TYPE_A) {
            // the incoming population opp(m) of x: slot (m, t) in layout 1, (opp(m), x) in layout 0
            if (SW) F[(int64_t)opp(m) * n + c] = pv;
            else F[(int64_t)m * n + t] = pv;
        } else if (pullable(tt)) {
            const int di = wrap1(x.i + cx(m), v.d.nx), dj = wrap1(x.j + cy(m), v.d.ny),
                      dk = wrap1(x.k + cz(m), v.d.nz);
            const int z = cell(v.d, di, dj, dk);
            if (v.type[z] == TYPE_F && pullable(v.type[cell(v.d, di - cx(m), dj - cy(m), dk - cz(m))])) {
                // the incoming population m of z: slot (opp(m), x) in layout 1, (m, z) in layout 0
                if (SW) F[(int64_t)m * n + z] = pv;
                else F[(int64_t)opp(m) * n + c] = pv;
            }
        }
    }
    for (int l = 0; l < Q; ++l) {
        const int y = cell(v.d, x.i - cx(l), x.j - cy(l), x.k - cz(l));
        const uint8_t ty = v.type[y];
        if (ty == TYPE_O || ty == TYPE_O_C || ty == TYPE_I || ty == TYPE_I_C) {
            // incoming population l of x: slot (l, x) in layout 0, (opp(l), y) in layout 1
            if (SW) F[(int64_t)l * n + c] = F[(int64_t)opp(l) * n + y];
            else F[(int64_t)opp(l) * n + y] = F[(int64_t)l * n + c];
        }
    }
}

// --------------------------------------------------------------- Collide
// PART 0: the whole collide, the relaxation rates from the cell's own
// transport evaluation.  With the FP32 transport pass's rates in v.omix the
// collide runs in two kernels: PART 1 evaluates the cell terms (the rates'
// tau check, theta, U - 3/2 theta and the energy-flux correction dq, which
// read the neighbours and every species) into v.omix; PART 2 relaxes the 27
// energy populations from them, a streaming pass of few registers.  Both
// forms do the same arithmetic in the same order.
enum : int { OMIX_OM = 0, OMIX_OM1 = 1, OMIX_THETA = 2, OMIX_USHIFT = 3, OMIX_DQ = 4, OMIX_PLANES = 7 };

template <typename P, int PART>
__global__ __launch_bounds__(BLOCK) void k_mc_collide(MultiView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    bool warn = false;
    if constexpr (PART == 2) {
        if (c >= v.d.n || v.type[c] != TYPE_F) return;
        const int64_t n = v.d.n;
        double* const* Fd = v.fld;
        const double rho = Fd[LBM_FIELD_RHO][c];
        const double vel[3] = {Fd[LBM_FIELD_U][c], Fd[LBM_FIELD_V][c], Fd[LBM_FIELD_W][c]};
        const double om = v.omix[OMIX_OM * n + c], om1 = v.omix[OMIX_OM1 * n + c];
        const double theta = v.omix[OMIX_THETA * n + c], ushift = v.omix[OMIX_USHIFT * n + c];
        const double dq[3] = {v.omix[OMIX_DQ * n + c], v.omix[(OMIX_DQ + 1) * n + c], v.omix[(OMIX_DQ + 2) * n + c]};
        const GeqAxes<double> G = geq_axes<double>(theta, vel);
        // all loads before the first store (v.go might alias v.g for all the
        // compiler knows, so interleaved each load would wait for the last store)
        double g[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) g[l] = ldg(v, v.g, l, c);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const double ge = geq_dir<double>(G, l, rho, ushift);
            const double cdq = 0.5 * (cx(l) * dq[0] + cy(l) * dq[1] + cz(l) * dq[2]);
            const double gs = (c2(l) == 1) ? ge + cdq : ge;
            stg(v, v.go, l, c, g[l] + om1 * (ge - g[l]) + (om - om1) * (ge - gs));
        }
        return;
    }
    if (c < v.d.n && v.type[c] == TYPE_F) {
        const int64_t n = v.d.n;
        const Ijk x = ijk(v.d, c);
        double* const* Fd = v.fld;
        const double rho = Fd[LBM_FIELD_RHO][c], T = Fd[LBM_FIELD_T][c], p = Fd[LBM_FIELD_P][c];
        const double vel[3] = {Fd[LBM_FIELD_U][c], Fd[LBM_FIELD_V][c], Fd[LBM_FIELD_W][c]};
        // FP32: products with reciprocals in place of divisions
        constexpr bool FAST = sizeof(P) == 4;
        const double irho = 1.0 / rho;
        double ra[NS], Y0[NS], X0[NS];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            ra[a] = v.rhoa[a * n + c];
            Y0[a] = FAST ? ra[a] * irho : ra[a] / rho;
        }
        normalize(Y0);
        const double Tsi = v.u.si_temp(T);
        const double Wm = mean_W(Y0);
        const double Rmix = v.u.cp(R_UNIV / Wm);
        const double theta = v.u.energy_mass(R_UNIV * Tsi / Wm);
        double om, om1;
        if constexpr (PART == 1) {
            om = v.omix[OMIX_OM * n + c];
            om1 = v.omix[OMIX_OM1 * n + c];
        } else {
            X_from_Y(Y0, X0);
            const double lnT = log(Tsi);
            const double cp = v.u.cp(cp_mass(Tsi, Y0));
            // transport arithmetic R: binary64 on the plain path, binary32 on the FP32 path
            using R = typename std::conditional<sizeof(P) == 4, float, double>::type;
            const double mu = v.u.mu(mix_viscosity_r<R>((R)lnT, X0));
            const double kap = v.u.thermal_conductivity(mix_conductivity_r<R>((R)lnT, X0));
            om1 = 2 * p * cp * 1 / (p * cp * 1 + 2 * kap);
            om = 2 * p * 1 / (p * 1 + 2 * mu);
        }
        warn = (1 / om <= 0.5) || (1 / om1 <= 0.5);
        double Uint = Fd[LBM_FIELD_RHOE][c] / rho;
        for (int a = 0; a < NS; ++a) {
            const double ux = v.ua[(0 * NS + a) * n + c], uy = v.ua[(1 * NS + a) * n + c],
                         uz = v.ua[(2 * NS + a) * n + c];
            Uint = Uint - (FAST ? 0.5 * ra[a] * irho : 0.5 * ra[a] / rho) * (ux * ux + uy * uy + uz * uz);
        }
        const int nm[3] = {cell(v.d, x.i - 1, x.j, x.k), cell(v.d, x.i, x.j - 1, x.k), c - 1};
        const int np[3] = {cell(v.d, x.i + 1, x.j, x.k), cell(v.d, x.i, x.j + 1, x.k), c + 1};
        double qdiff[3] = {0, 0, 0}, qcorr[3] = {0, 0, 0};
        // neighbour mass fractions through one reciprocal density per
        // neighbour (a multiply instead of a division per species)
        double irn[3], irp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            irn[d] = 1.0 / Fd[LBM_FIELD_RHO][nm[d]];
            irp[d] = 1.0 / Fd[LBM_FIELD_RHO][np[d]];
        }
        const double invT = 1.0 / Tsi;
        for (int a = 0; a < NS; ++a) {
            const double ha = v.u.energy_mass((FAST ? h_RT_inv(a, Tsi, invT) : h_RT(a, Tsi)) * R_UNIV * Tsi * inv_w(a));
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double uad = v.ua[(d * NS + a) * n + c];
                qdiff[d] += om1 / (-om + om1) * ha * ra[a] * (uad - vel[d]);
                const double Yn = v.rhoa[a * n + nm[d]] * irn[d];
                const double Yp = v.rhoa[a * n + np[d]] * irp[d];
                const double dY = fd_fuw<double>(Yn, ra[a] * irho, Yp, uad - vel[d], v.type[nm[d]], v.type[np[d]]);
                qcorr[d] += 0.5 * (2.0 - om1) / (-om + om1) * 1 * p * ha * dY;
            }
        }
        // mixture gas constant of a neighbour, R_univ sum(Y/W) / sum(Y): the
        // ratio is the same with rho_a in place of Y_a
        auto Rn = [&](int m) -> double {
            if (!fluidlike(v.type[m])) return 0.0;
            double sYW = 0, sY = 0;
#pragma unroll
            for (int a = 0; a < NS; ++a) {
                const double ra_m = v.rhoa[a * n + m];
                sYW += ra_m * inv_w(a);
                sY += ra_m;
            }
            return v.u.cp(R_UNIV * sYW / sY);
        };
        double dQ[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* ud = Fd[LBM_FIELD_U + d];
            auto Qf = [&](int m, double R) {
                const double r = Fd[LBM_FIELD_RHO][m], u = ud[m];
                return r * u * (1.0 - 3.0 * R * Fd[LBM_FIELD_T][m]) - r * (u * u * u);
            };
            dQ[d] = fd_central<double>(Qf(nm[d], Rn(nm[d])), Qf(c, Rmix), Qf(np[d], Rn(np[d])), v.type[nm[d]],
                                       v.type[np[d]]);
        }
        double dq[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double acc = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int idx = (a == b) ? a : (3 + a + b - 1);
                const double eqP = (a == b ? p : 0.0) + rho * vel[a] * vel[b];
                acc += vel[a] * (Fd[LBM_FIELD_PXX + idx][c] - eqP);
            }
            dq[b] = acc + qdiff[b] + qcorr[b] + 0.5 * 1 * vel[b] * dQ[b];
        }
        if constexpr (PART == 1) {
            v.omix[OMIX_THETA * n + c] = theta;
            v.omix[OMIX_USHIFT * n + c] = Uint - 1.5 * theta;
#pragma unroll
            for (int b = 0; b < 3; ++b) v.omix[(OMIX_DQ + b) * n + c] = dq[b];
        } else {
        const GeqAxes<double> G = geq_axes<double>(theta, vel);
        double g[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) g[l] = ldg(v, v.g, l, c);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const double ge = geq_dir<double>(G, l, rho, Uint - 1.5 * theta);
            const double cdq = 0.5 * (cx(l) * dq[0] + cy(l) * dq[1] + cz(l) * dq[2]);
            const double gs = (c2(l) == 1) ? ge + cdq : ge;
            stg(v, v.go, l, c, g[l] + om1 * (ge - g[l]) + (om - om1) * (ge - gs));
        }
        }
    }
    const unsigned long long m = __ballot(warn);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1)
        atomicAdd(v.tau_count, (unsigned long long)__popcll(m));
}

// ------------------------------------------------------------- Streaming
// v.f/v.g: post-collision (source); v.fo/v.go: streamed (destination).  The
// rest-state shifts of FP32 storage are the same for opposite directions, so
// stored values move unchanged.
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_stream(MultiView<P> v, int first)
{
    // one thread per (cell, population set): first + blockIdx.y = species
    // 0..NS-1, NS = g
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    const int set = first + blockIdx.y;
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    const int nx = v.d.nx, ny = v.d.ny, nz = v.d.nz;
    const P* src_base = set < NS ? v.f + (int64_t)set * Q * n : v.g;
    P* dst_base = set < NS ? v.fo + (int64_t)set * Q * n : v.go;
    if (v.nbflag[c] & NB_SIMPLE) {  // fluid and periodic neighbours only: plain pull with the wrap
        const int si[3] = {x.i + 1 == nx - 1 ? 1 : x.i + 1, x.i, x.i - 1 == 0 ? nx - 2 : x.i - 1};
        const int sj[3] = {x.j + 1 == ny - 1 ? 1 : x.j + 1, x.j, x.j - 1 == 0 ? ny - 2 : x.j - 1};
        const int sk[3] = {x.k + 1 == nz - 1 ? 1 : x.k + 1, x.k, x.k - 1 == 0 ? nz - 2 : x.k - 1};
        // all 27 pulls before the first store: the buffers are distinct, but
        // interleaved each load would wait for the previous store
        P val[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) val[l] = at32(src_base + l * n, cell(v.d, si[cx(l) + 1], sj[cy(l) + 1], sk[cz(l) + 1]));
#pragma unroll
        for (int l = 0; l < Q; ++l) at32(dst_base + l * n, c) = val[l];
        return;
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        int in = x.i - cx(l), jn = x.j - cy(l), kn = x.k - cz(l);
        const uint8_t tn = v.type[cell(v.d, in, jn, kn)];
        if (tn == TYPE_S || tn == TYPE_A) {
            // bounce-back; the energy populations bounce only at adiabatic walls
            if (set < NS || tn == TYPE_A) dst_base[(int64_t)l * n + c] = src_base[(int64_t)opp(l) * n + c];
        } else if (tn == TYPE_FS) {
            // free-slip wall (lbm-streaming.cpp:34-49): the mirrored direction
            // from the mirrored source (dirSlip, lbm-bc.cpp:1352-1399)
            int a = cx(l), b = cy(l), e = cz(l);
            int ip = x.i - a, jp = x.j - b, kp = x.k - e;
            if (v.type[cell(v.d, x.i - a, x.j, x.k)] == TYPE_FS) {
                a = -a;
                ip = x.i;
            } else if (v.type[cell(v.d, x.i, x.j - b, x.k)] == TYPE_FS) {
                b = -b;
                jp = x.j;
            } else if (v.type[cell(v.d, x.i, x.j, x.k - e)] == TYPE_FS) {
                e = -e;
                kp = x.k;
            }
            ip = ((ip - 1 + (nx - 2)) % (nx - 2)) + 1;
            jp = ((jp - 1 + (ny - 2)) % (ny - 2)) + 1;
            kp = ((kp - 1 + (nz - 2)) % (nz - 2)) + 1;
            dst_base[(int64_t)l * n + c] = src_base[(int64_t)dir_index(a, b, e) * n + cell(v.d, ip, jp, kp)];
        } else if (tn == TYPE_O || tn == TYPE_O_C || tn == TYPE_I || tn == TYPE_I_C) {
            // inflow/outflow sources are left for TMS_BC
        } else {
            if (in == 0) in = nx - 2; else if (in == nx - 1) in = 1;
            if (jn == 0) jn = ny - 2; else if (jn == ny - 1) jn = 1;
            if (kn == 0) kn = nz - 2; else if (kn == nz - 1) kn = 1;
            dst_base[(int64_t)l * n + c] = src_base[(int64_t)l * n + cell(v.d, in, jn, kn)];
        }
    }
}

// Fused mode: the post-collision species sets of the side cells, copied out
// of the strict step's fo.  grid: (ceil(nside / BLOCK), NS).
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_side_fill(MultiView<P> v)
{
    const int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (s >= v.nside) return;
    const int c = v.side_cell[s];
    const int a = blockIdx.y;
    const int64_t n = v.d.n;
    const P* src = v.fo + (int64_t)a * Q * n;
    P* dst = v.fside + (int64_t)a * Q * v.nside;
#pragma unroll
    for (int l = 0; l < Q; ++l) dst[(int64_t)l * v.nside + s] = src[(int64_t)l * n + c];
}

// Fused mode: the side cells' post-collision species sets streamed from the
// side store into fo (push_general).  grid: (ceil(nside / BLOCK), NS).
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_side_push(MultiView<P> v)
{
    const int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (s >= v.nside) return;
    const int c = v.side_cell[s];
    const int a = blockIdx.y;
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    const P* S = v.fside + (int64_t)a * Q * v.nside;
    const P* A = v.f + (int64_t)a * Q * n;
    P* B = v.fo + (int64_t)a * Q * n;
    for (int l = 0; l < Q; ++l) push_general<P>(v, A, B, c, x, l, S[(int64_t)l * v.nside + s]);
}

// Fused mode: the post-collision species sets rebuilt into fo from the
// streamed f (NB_PUSH cells) and the side store (the others).  grid:
// (xcd_grid(n), NS).
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_unstream(MultiView<P> v)
{
    const int c = xcd_block() * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    const int a = blockIdx.y;
    const int64_t n = v.d.n;
    const P* A = v.f + (int64_t)a * Q * n;
    P* B = v.fo + (int64_t)a * Q * n;
    if (v.nbflag[c] & NB_PUSH) {
        const Ijk x = ijk(v.d, c);
        const int ti[3] = {wrap1(x.i - 1, v.d.nx), x.i, wrap1(x.i + 1, v.d.nx)};
        const int tj[3] = {wrap1(x.j - 1, v.d.ny), x.j, wrap1(x.j + 1, v.d.ny)};
        const int tk[3] = {wrap1(x.k - 1, v.d.nz), x.k, wrap1(x.k + 1, v.d.nz)};
#pragma unroll
        for (int l = 0; l < Q; ++l)
            B[(int64_t)l * n + c] = A[(int64_t)l * n + cell(v.d, ti[cx(l) + 1], tj[cy(l) + 1], tk[cz(l) + 1])];
        return;
    }
    const int s = v.side_slot[c];
    const P* S = v.fside + (int64_t)a * Q * v.nside;
#pragma unroll
    for (int l = 0; l < Q; ++l) B[(int64_t)l * n + c] = S[(int64_t)l * v.nside + s];
}

// ------------------------------------------------------- calculate_moment
// Dense symmetric positive definite solve A x = b (Cholesky, lower factor);
// A is overwritten by the factor, B holds the three right-hand sides.
__device__ inline void cholesky_solve3(double A[NS][NS], double B[NS][3])
{
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double s = A[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= A[j][k] * A[j][k];
        const double d = sqrt(s);
        A[j][j] = d;
#pragma unroll
        for (int i = j + 1; i < NS; ++i) {
            double t = A[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= A[i][k] * A[j][k];
            A[i][j] = t / d;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            double t = B[i][r];
#pragma unroll
            for (int k = 0; k < i; ++k) t -= A[i][k] * B[k][r];
            B[i][r] = t / A[i][i];
        }
#pragma unroll
        for (int i = NS - 1; i >= 0; --i) {
            double t = B[i][r];
#pragma unroll
            for (int k = i + 1; k < NS; ++k) t -= A[k][i] * B[k][r];
            B[i][r] = t / A[i][i];
        }
    }
}

// The same solve with A held as its lower triangle, L[tri(i, j)] = A[i][j]
// for j <= i (36 fewer live doubles); same operations in the same order.
constexpr int NTRI = NS * (NS + 1) / 2;
__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// (T = double: the FP64 path; T = float: the FP32 path's binary32 solve.)
template <typename T>
__device__ inline void cholesky_solve3(T L[NTRI], T B[NS][3])
{
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        T s = L[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[tri(j, k)] * L[tri(j, k)];
        const T d = sqrt(s);
        L[tri(j, j)] = d;
#pragma unroll
        for (int i = j + 1; i < NS; ++i) {
            T t = L[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[tri(i, k)] * L[tri(j, k)];
            L[tri(i, j)] = t / d;
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            T t = B[i][r];
#pragma unroll
            for (int k = 0; k < i; ++k) t -= L[tri(i, k)] * B[k][r];
            B[i][r] = t / L[tri(i, i)];
        }
#pragma unroll
        for (int i = NS - 1; i >= 0; --i) {
            T t = B[i][r];
#pragma unroll
            for (int k = i + 1; k < NS; ++k) t -= L[tri(k, i)] * B[k][r];
            B[i][r] = t / L[tri(i, i)];
        }
    }
}

// The binary32 solve of the FP32 path: the same factorisation with the
// reciprocal of each pivot taken once (hardware reciprocal, 1 ulp) and
// multiplied in place of the divisions.
__device__ inline void cholesky_solve3_rcp(float L[NTRI], float B[NS][3])
{
    float rd[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        float s = L[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[tri(j, k)] * L[tri(j, k)];
        const float d = sqrtf(s);
        L[tri(j, j)] = d;
        rd[j] = rcp_r(d);
#pragma unroll
        for (int i = j + 1; i < NS; ++i) {
            float t = L[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[tri(i, k)] * L[tri(j, k)];
            L[tri(i, j)] = t * rd[j];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            float t = B[i][r];
#pragma unroll
            for (int k = 0; k < i; ++k) t -= L[tri(i, k)] * B[k][r];
            B[i][r] = t * rd[i];
        }
#pragma unroll
        for (int i = NS - 1; i >= 0; --i) {
            float t = B[i][r];
#pragma unroll
            for (int k = i + 1; k < NS; ++k) t -= L[tri(k, i)] * B[k][r];
            B[i][r] = t * rd[i];
        }
    }
}

// The FP64 kernel holds more live doubles than fit two waves per SIMD: it
// runs at one wave per SIMD with the accumulation registers and no scratch
// spills (2.9x faster than the spilling default).  The FP32 one runs at two
// waves per SIMD without scratch: its per-species sums go through LDS, the
// Maxwell-Stefan solve is binary32 and the species planes are written before
// the temperature iteration (2.10 -> 1.49 ms on the 256x128x128 layer).
// SW: the in-place mode's layout 1 (species populations read through their
// sources, with the wrapped neighbour coordinates when every neighbour is
// fluid or periodic).
template <typename P, int SW>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(P) == 8 ? 1 : 2, sizeof(P) == 8 ? 1 : 8))) void k_mc_moments(MultiView<P> v)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n || v.type[c] != TYPE_F) return;
    const int64_t n = v.d.n;
    const bool simple = SW && (v.nbflag[c] & NB_SIMPLE);
    const Wrap3 w(v.d, ijk(v.d, c));
    auto species_f = [&](int a, int l) -> double {
        if constexpr (!SW) return ldf(v, v.f, a, l, c);
        const int src = simple ? w.at(v.d, l, -1) : aa_source(v, c, l);
        const double x = (double)v.f[((int64_t)a * Q + opp(l)) * n + src];
        if constexpr (sizeof(P) == 4) return x + fshift(v, a, l);
        else return x;
    };
    double* const* Fd = v.fld;
    // Cantera state of the previous moments (old T, rho, composition)
    const double rho_old = Fd[LBM_FIELD_RHO][c];
    const double Tsi_old = v.u.si_temp(Fd[LBM_FIELD_T][c]);
    const double rho_old_si = v.u.si_rho(rho_old);
    const double p_old = Fd[LBM_FIELD_P][c];
    constexpr bool FAST = sizeof(P) == 4;  // FP32: reciprocals in place of divisions
    const double irho_old = 1.0 / rho_old;
    double Yold[NS], Xold[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        Yold[a] = FAST ? v.rhoa[a * n + c] * irho_old : v.rhoa[a * n + c] / rho_old;
        Xold[a] = v.X[a * n + c];
    }
    normalize(Yold);
    const double p_old_si = pressure<FAST>(Tsi_old, rho_old_si, Yold);
    const double lnT_old = log(Tsi_old);
    // energy moments
    double re = 0, qx = 0, qy = 0, qz = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const double g = ldg(v, v.g, l, c);
        re += g;
        if (cx(l)) qx += cx(l) * g;
        if (cy(l)) qy += cy(l) * g;
        if (cz(l)) qz += cz(l) * g;
    }
    // species moments.  FP32: the per-species results go through LDS, one
    // column per thread (the rolled loop would put register arrays it indexes
    // by species into scratch memory)
    constexpr bool LDS_STAGE = sizeof(P) == 4 && NS <= 16;
    __shared__ double s_rr[LDS_STAGE ? NS : 1][BLOCK];
    __shared__ float s_B[LDS_STAGE ? 3 * NS : 1][BLOCK];
    double Pt[6] = {0, 0, 0, 0, 0, 0};
    double rr[NS], B[NS][3];
    double rho = 0, m0 = 0, m1 = 0, m2 = 0;
    for (int a = 0; a < NS; ++a) {
        double r = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const double f = species_f(a, l);
            r += f;
            if (cx(l)) mx += cx(l) * f;
            if (cy(l)) my += cy(l) * f;
            if (cz(l)) mz += cz(l) * f;
            if (cx(l)) Pt[0] += f;
            if (cy(l)) Pt[1] += f;
            if (cz(l)) Pt[2] += f;
            if (cx(l) * cy(l)) Pt[3] += (cx(l) * cy(l)) * f;
            if (cx(l) * cz(l)) Pt[4] += (cx(l) * cz(l)) * f;
            if (cy(l) * cz(l)) Pt[5] += (cy(l) * cz(l)) * f;
        }
        const double rclamp = r > SPECIES_MIN ? (r > SPECIES_MIN ? r : SPECIES_MIN) : SPECIES_MIN;
        if constexpr (LDS_STAGE) {
            s_rr[a][threadIdx.x] = rclamp;
            s_B[3 * a + 0][threadIdx.x] = (float)mx;
            s_B[3 * a + 1][threadIdx.x] = (float)my;
            s_B[3 * a + 2][threadIdx.x] = (float)mz;
        } else {
            rr[a] = rclamp;
            B[a][0] = mx;
            B[a][1] = my;
            B[a][2] = mz;
        }
        rho += r;
        m0 += mx;
        m1 += my;
        m2 += mz;
    }
    const double vel[3] = {m0 / rho, m1 / rho, m2 / rho};
    // Maxwell-Stefan velocities: A u = (rho u)_a with the old state's D, p, X;
    // then the internal energy U = rho e / rho - sum 1/2 rho_a |u_a|^2 / rho and
    // the species velocities are written (not live across the T iteration)
    double Uint = re / rho;
    if constexpr (sizeof(P) == 8) {
        double A[NTRI];  // lower triangle; A[a][b] = A[b][a] = A[tri(max, min)]
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int b = 0; b < NS; ++b) {
                if (b < a) continue;
                if (b == a) {
                    A[tri(a, a)] = rr[a] == 0.0 ? 1.0 : rr[a];
                } else {
                    const double D = v.u.nu(binary_diff(a, b, lnT_old, p_old_si));
                    A[tri(b, a)] = -1.0 * 1 / 2.0 * p_old * Xold[a] * Xold[b] / D;
                }
            }
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int b = 0; b < NS; ++b)
                if (b != a) A[tri(a, a)] -= A[b > a ? tri(b, a) : tri(a, b)];
        cholesky_solve3(A, B);
    } else {
        // binary32 diffusivities (as the transport pass of the FP32 path) and a
        // binary32 solve: the velocities are stored in binary32
#pragma unroll
        for (int a = 0; a < NS; ++a) rr[a] = s_rr[a][threadIdx.x];
        const float lnTf = (float)lnT_old, rpf = rcp_r((float)p_old_si);
        const float knu = (float)((double)v.u.s / ((double)v.u.m * (double)v.u.m));
        const float hp = -0.5f * (float)p_old;
        float A[NTRI], Bf[NS][3];
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int b = 0; b < NS; ++b) {
                if (b < a) continue;
                if (b == a) {
                    A[tri(a, a)] = rr[a] == 0.0 ? 1.0f : (float)rr[a];
                } else {
                    const float D = exp(poly4r<float>(DIFF_FIT[a][b], lnTf)) * rpf * knu;
                    A[tri(b, a)] = hp * (float)Xold[a] * (float)Xold[b] * rcp_r(D);
                }
            }
#pragma unroll
        for (int a = 0; a < NS; ++a) {
#pragma unroll
            for (int b = 0; b < NS; ++b)
                if (b != a) A[tri(a, a)] -= A[b > a ? tri(b, a) : tri(a, b)];
#pragma unroll
            for (int d = 0; d < 3; ++d) Bf[a][d] = s_B[3 * a + d][threadIdx.x];
        }
        cholesky_solve3_rcp(A, Bf);
        const double irho = 1.0 / rho;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            const double b0 = Bf[a][0], b1 = Bf[a][1], b2 = Bf[a][2];
            Uint = Uint - 0.5 * rr[a] * irho * (b0 * b0 + b1 * b1 + b2 * b2);
        }
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) v.ua[(d * NS + a) * n + c] = Bf[a][d];
    }
    // temperature from energy at the new composition
    double Y[NS], X[NS];
    const double irho_new = 1.0 / rho;
#pragma unroll
    for (int a = 0; a < NS; ++a) Y[a] = FAST ? rr[a] * irho_new : rr[a] / rho;
    normalize(Y);
    if constexpr (sizeof(P) == 8) {  // (write order measured fastest at one wave per SIMD)
        for (int a = 0; a < NS; ++a)
            Uint = Uint - 0.5 * rr[a] / rho * (B[a][0] * B[a][0] + B[a][1] * B[a][1] + B[a][2] * B[a][2]);
        const double usi = v.u.si_energy_mass(Uint);
        const double rho_si = v.u.si_rho(rho);
        const double Tsi = T_from_UV(usi, Y, Tsi_old);
        X_from_Y(Y, X);
        Fd[LBM_FIELD_RHO][c] = rho;
        Fd[LBM_FIELD_U][c] = vel[0];
        Fd[LBM_FIELD_V][c] = vel[1];
        Fd[LBM_FIELD_W][c] = vel[2];
        Fd[LBM_FIELD_RHOE][c] = re;
        Fd[LBM_FIELD_QX][c] = qx;
        Fd[LBM_FIELD_QY][c] = qy;
        Fd[LBM_FIELD_QZ][c] = qz;
#pragma unroll
        for (int a = 0; a < 6; ++a) Fd[LBM_FIELD_PXX + a][c] = Pt[a];
        Fd[LBM_FIELD_T][c] = v.u.temp(Tsi);
        Fd[LBM_FIELD_P][c] = v.u.p(pressure(Tsi, rho_si, Y));
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            v.rhoa[a * n + c] = rr[a];
            v.X[a * n + c] = X[a];
#pragma unroll
            for (int d = 0; d < 3; ++d) v.ua[(d * NS + a) * n + c] = B[a][d];
        }
        return;
    }
    X_from_Y<true>(Y, X);
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        v.rhoa[a * n + c] = rr[a];
        v.X[a * n + c] = X[a];
    }
    Fd[LBM_FIELD_RHO][c] = rho;
    Fd[LBM_FIELD_U][c] = vel[0];
    Fd[LBM_FIELD_V][c] = vel[1];
    Fd[LBM_FIELD_W][c] = vel[2];
    Fd[LBM_FIELD_RHOE][c] = re;
    Fd[LBM_FIELD_QX][c] = qx;
    Fd[LBM_FIELD_QY][c] = qy;
    Fd[LBM_FIELD_QZ][c] = qz;
#pragma unroll
    for (int a = 0; a < 6; ++a) Fd[LBM_FIELD_PXX + a][c] = Pt[a];
    const double usi = v.u.si_energy_mass(Uint);
    const double rho_si = v.u.si_rho(rho);
    const double Tsi = T_from_UV<true>(usi, Y, Tsi_old);
    Fd[LBM_FIELD_T][c] = v.u.temp(Tsi);
    Fd[LBM_FIELD_P][c] = v.u.p(pressure<true>(Tsi, rho_si, Y));
}

// calculate_moment_point (lbm-moment.cpp:379-537) of cell c, writing its
// fields.  Unlike calculate_moment: no SPECIES_MIN clamp; the binary
// diffusivities at the T of the cell's fields and the new density and
// composition; the Maxwell-Stefan matrix from the fields' p and X, not yet
// updated; internal energy rho e / rho - (sum 1/2 rho_a |u_a|^2) / rho.  The
// reference solves by partial-pivoting LU, this by Cholesky (the matrix is
// symmetric positive definite).
template <typename P>
__device__ __noinline__ void mc_moment_point(const MultiView<P>& v, int c)
{
    const int64_t n = v.d.n;
    double* const* Fd = v.fld;
    double re = 0, qx = 0, qy = 0, qz = 0;
    for (int l = 0; l < Q; ++l) {
        const double g = ldg(v, v.g, l, c);
        re += g;
        if (cx(l)) qx += cx(l) * g;
        if (cy(l)) qy += cy(l) * g;
        if (cz(l)) qz += cz(l) * g;
    }
    Fd[LBM_FIELD_RHOE][c] = re;
    Fd[LBM_FIELD_QX][c] = qx;
    Fd[LBM_FIELD_QY][c] = qy;
    Fd[LBM_FIELD_QZ][c] = qz;
    double Pt[6] = {0, 0, 0, 0, 0, 0};
    double ra[NS], B[NS][3];
    double rho = 0, m0 = 0, m1 = 0, m2 = 0;
    for (int a = 0; a < NS; ++a) {
        double r = 0, mx = 0, my = 0, mz = 0;
        for (int l = 0; l < Q; ++l) {
            const double f = ldf_in(v, a, l, c);
            r += f;
            if (cx(l)) mx += cx(l) * f;
            if (cy(l)) my += cy(l) * f;
            if (cz(l)) mz += cz(l) * f;
            if (cx(l)) Pt[0] += f;
            if (cy(l)) Pt[1] += f;
            if (cz(l)) Pt[2] += f;
            if (cx(l) * cy(l)) Pt[3] += (cx(l) * cy(l)) * f;
            if (cx(l) * cz(l)) Pt[4] += (cx(l) * cz(l)) * f;
            if (cy(l) * cz(l)) Pt[5] += (cy(l) * cz(l)) * f;
        }
        ra[a] = r;
        B[a][0] = mx;
        B[a][1] = my;
        B[a][2] = mz;
        rho += r;
        m0 += mx;
        m1 += my;
        m2 += mz;
    }
    for (int a = 0; a < 6; ++a) Fd[LBM_FIELD_PXX + a][c] = Pt[a];
    Fd[LBM_FIELD_RHO][c] = rho;
    Fd[LBM_FIELD_U][c] = m0 / rho;
    Fd[LBM_FIELD_V][c] = m1 / rho;
    Fd[LBM_FIELD_W][c] = m2 / rho;
    double Y[NS];
    for (int a = 0; a < NS; ++a) Y[a] = ra[a] / rho;
    normalize(Y);
    const double Tsi_old = v.u.si_temp(Fd[LBM_FIELD_T][c]);
    const double rho_si = v.u.si_rho(rho);
    const double p_si = pressure(Tsi_old, rho_si, Y);
    const double lnT = log(Tsi_old);
    const double p_fld = Fd[LBM_FIELD_P][c];
    double A[NS][NS];
    for (int a = 0; a < NS; ++a)
        for (int b = 0; b < NS; ++b) {
            if (b != a) {
                const double D = v.u.nu(binary_diff(a, b, lnT, p_si));
                A[a][b] = -1.0 * 1 / 2.0 * p_fld * v.X[a * n + c] * v.X[b * n + c] / D;
            } else {
                A[a][a] = ra[a] == 0.0 ? 1.0 : ra[a];
            }
        }
    for (int a = 0; a < NS; ++a)
        for (int b = 0; b < NS; ++b)
            if (b != a) A[a][a] -= A[a][b];
    cholesky_solve3(A, B);
    double ke = 0;
    for (int a = 0; a < NS; ++a) ke += 0.5 * ra[a] * (B[a][0] * B[a][0] + B[a][1] * B[a][1] + B[a][2] * B[a][2]);
    const double Uint = re / rho - ke / rho;
    const double Tsi = T_from_UV(v.u.si_energy_mass(Uint), Y, Tsi_old);
    double X[NS];
    X_from_Y(Y, X);
    for (int a = 0; a < NS; ++a) {
        v.rhoa[a * n + c] = ra[a];
        v.X[a * n + c] = X[a];
        for (int d = 0; d < 3; ++d) v.ua[(d * NS + a) * n + c] = B[a][d];
    }
    Fd[LBM_FIELD_T][c] = v.u.temp(Tsi);
    Fd[LBM_FIELD_P][c] = v.u.p(pressure(Tsi, rho_si, Y));
}

// ---------------------------------------------------------------- TMS_BC
// Outflow sweep of TMS_BC (lbm-bc.cpp:126-340) for TYPE_O_C ghosts with the
// NSCBC target state of impose_nscbc.cpp:45-546.  One thread per interface
// cell of the compact list built at lbm_set_types; each thread rewrites only
// its own cell's populations, reading neighbours' moments of the last
// calculate_moment.
namespace {

constexpr int NV = 5 + NS;  // p, u, v, w, rho, Y[NS]

template <typename P>
__device__ inline void cell_vals(const MultiView<P>& v, int c, double out[NV])
{
    const int64_t n = v.d.n;
    out[0] = v.fld[LBM_FIELD_P][c];
    out[1] = v.fld[LBM_FIELD_U][c];
    out[2] = v.fld[LBM_FIELD_V][c];
    out[3] = v.fld[LBM_FIELD_W][c];
    const double r = v.fld[LBM_FIELD_RHO][c];
    out[4] = r;
#pragma unroll
    for (int a = 0; a < NS; ++a) out[5 + a] = v.rhoa[a * n + c] / r;
}

// one-sided second order into the domain (normal_derivative) or the
// tangential rule of tangential_derivative (z repeats its k-1 test)
template <typename P>
__device__ inline void deriv(const MultiView<P>& v, int c, int stride, int mode, double D[NV])
{
    double a[NV], b[NV], e[NV];
    if (mode == 0) {  // central
        cell_vals(v, c + stride, a);
        cell_vals(v, c - stride, b);
#pragma unroll
        for (int q = 0; q < NV; ++q) D[q] = (a[q] - b[q]) / (2.0 * 1);
        return;
    }
    cell_vals(v, c, a);
    cell_vals(v, c + mode * stride, b);
    cell_vals(v, c + 2 * mode * stride, e);
    const double s = (double)mode;  // +1 forward, -1 backward
#pragma unroll
    for (int q = 0; q < NV; ++q) D[q] = s * (-1.5 * a[q] + 2.0 * b[q] - 0.5 * e[q]) / 1;
}

// T from (rho, p, Y) in SI (setState_DP) with the lattice crossings.
__device__ inline double T_from_DP(const Units& u, double rho, double p, double Y[NS], double* Tsi_out)
{
    normalize(Y);
    const double rho_si = u.si_rho(rho), p_si = u.si_p(p);
    const double Tsi = p_si * mean_W(Y) / (R_UNIV * rho_si);
    *Tsi_out = Tsi;
    return u.temp(Tsi);
}

// frozen sound speed (lattice) and gamma at the DP state of a cell.  The
// state takes a normalised copy of Y (setMassFractions clips and scales its
// own copy): the caller's mass fractions stay as they were, since the
// boundary update builds the outflow densities from them
// (impose_nscbc.cpp:479-545)
__device__ inline void sound_DP(const Units& u, double rho, double p, const double Y[NS], double* a_lat, double* gam)
{
    double Yn[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) Yn[k] = Y[k];
    double Tsi;
    T_from_DP(u, rho, p, Yn, &Tsi);
    const double cp = cp_mass(Tsi, Yn);
    const double Wm = mean_W(Yn);
    const double cv = cp - R_UNIV / Wm;
    *gam = cp / cv;
    *a_lat = u.u(sqrt(*gam * R_UNIV * Tsi / Wm));
}

}  // namespace

// The sweep runs in four launches over a record of per-entry values
// (tms_rec, [TR_*][count] doubles), one thread per entry or per (entry,
// population set), with the arithmetic of a single per-cell pass:
//   k_mc_tms_target   the target state (TYPE_O: calculate_moment_point of the
//                     downstream cell; TYPE_O_C: the NSCBC LODI step);
//   k_mc_tms_local    per (entry, species): the local moments with the target
//                     populations in the unknown directions;
//   k_mc_tms_close    per entry: mixture sums, calculate_temp, local state;
//   k_mc_tms_rebuild  per (entry, species or energy): the reconstruction.
// Each entry is a distinct cell, and only the last launch writes populations,
// each thread those of its own cell.
enum {
    TR_RHO_OUT = 0, TR_VOUT = 1, TR_TSO = 4, TR_U_OUT = 5, TR_TH_OUT = 6, TR_MASK = 7, TR_RA_OUT = 8,
    TR_RA_LOC = TR_RA_OUT + NS, TR_MA = TR_RA_LOC + NS, TR_RHO_LOC = TR_MA + 3 * NS, TR_VLOC = TR_RHO_LOC + 1,
    TR_TSL = TR_VLOC + 3, TR_U_LOC = TR_TSL + 1, TR_TH_LOC = TR_U_LOC + 1, TR_COUNT = TR_TH_LOC + 1
};

template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_target(MultiView<P> v, const int* cells, const int* dirs, int count,
                                                      double* rec)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t], l = dirs[t];
    const Ijk x = ijk(v.d, c);
    const int strides[3] = {v.d.ny * v.d.nz, v.d.nz, 1};
    const int Ns[3] = {v.d.nx - 2, v.d.ny - 2, v.d.nz - 2};
    const int dims[3] = {v.d.nx, v.d.ny, v.d.nz};
    const int cl[3] = {cx(l), cy(l), cz(l)};
    const int idir = cl[0] ? 0 : (cl[1] ? 1 : 2);
    const int isign = cl[idir];
    double* const* Fd = v.fld;
    const int64_t n = v.d.n;
    const int cg = c - isign * strides[idir];  // the ghost
    double rho_out, vout[3], ra_out[NS], Yo[NS], T_out;
    if (v.type[cg] == TYPE_O) {
        // TYPE_O target (lbm-bc.cpp:168-198): the moments of the cell
        // downstream, recomputed by calculate_moment_point, the ghost's
        // pressure, T from setState_DP
        const int cd = c + isign * strides[idir];
        mc_moment_point(v, cd);
        rho_out = Fd[LBM_FIELD_RHO][cd];
        vout[0] = Fd[LBM_FIELD_U][cd];
        vout[1] = Fd[LBM_FIELD_V][cd];
        vout[2] = Fd[LBM_FIELD_W][cd];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            ra_out[a] = v.rhoa[a * n + cd];
            Yo[a] = ra_out[a] / rho_out;
        }
        double Tsi_o;
        T_out = T_from_DP(v.u, rho_out, Fd[LBM_FIELD_P][cg], Yo, &Tsi_o);
    } else {
    // ---- impose_NSCBC: LODI waves at the boundary cell
    double V0[NV];
    cell_vals(v, c, V0);
    const double p = V0[0], rho = V0[4];
    const double vel[3] = {V0[1], V0[2], V0[3]};
    double Yc[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yc[a] = V0[5 + a];
    double a_s, gam;
    sound_DP(v.u, rho, p, Yc, &a_s, &gam);
    double Dn[NV];
    deriv(v, c, strides[idir], isign, Dn);
    double T1 = 0, T5 = 0;
    {
        int t1 = (idir + 1) % 3, t2 = (idir + 2) % 3;
        if (t1 > t2) { const int tmp = t1; t1 = t2; t2 = tmp; }
        double Dt[2][NV];
        const int tt[2] = {t1, t2};
        bool any = false;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int d = tt[q];
#pragma unroll
            for (int r = 0; r < NV; ++r) Dt[q][r] = 0.0;
            if (Ns[d] > 3) {
                any = true;
                const int s = strides[d];
                int mode;
                if (v.type[c - s] != TYPE_F) mode = 1;
                else if (d != 2 && v.type[c + s] != TYPE_F) mode = -1;
                else mode = 0;
                deriv(v, c, s, mode, Dt[q]);
            }
        }
        if (any) {
            // Yoo & Im transverse terms for the normal direction idir
            T1 = vel[t1] * (Dt[0][0] - rho * a_s * Dt[0][1 + idir]) + vel[t2] * (Dt[1][0] - rho * a_s * Dt[1][1 + idir]) +
                 gam * p * (Dt[0][1 + t1] + Dt[1][1 + t2]);
            T5 = vel[t1] * (Dt[0][0] + rho * a_s * Dt[0][1 + idir]) + vel[t2] * (Dt[1][0] + rho * a_s * Dt[1][1 + idir]) +
                 gam * p * (Dt[0][1 + t1] + Dt[1][1 + t2]);
        }
    }
    const double un = vel[idir], dun = Dn[1 + idir], dpn = Dn[0], drhon = Dn[4];
    const double mach = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) / a_s;
    double L1 = (un - a_s) * (dpn - rho * a_s * dun);
    double L5 = (un + a_s) * (dpn + rho * a_s * dun);
    const double Ls = un * (a_s * a_s * drhon - dpn);
    const int cb = c - isign * strides[idir];
    if (mach < 1.0) {
        const double Kout = NSCBC_SIGMA_OUT * (1.0 - mach * mach) * (a_s / dims[idir]);
        if (isign == 1) L5 = Kout * (p - Fd[LBM_FIELD_P][cb]) - (1.0 - mach) * T5;
        else L1 = Kout * (p - Fd[LBM_FIELD_P][cb]) - (1.0 - mach) * T1;
    }
    // update_bc_cells, dt = 1
    const double drho = (Ls + 0.5 * (L1 + L5)) / (a_s * a_s);
#pragma unroll
    for (int d = 0; d < 3; ++d) vout[d] = vel[d] - 1 * (d == idir ? (L5 - L1) / (2.0 * a_s * rho) : un * Dn[1 + d]);
    rho_out = rho - 1 * drho;
    const double p_out = p - 1 * (0.5 * (L1 + L5));
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        ra_out[a] = (Yc[a] - 1 * (un * Dn[5 + a])) * rho_out;
        Yo[a] = ra_out[a] / rho_out;
    }
    double Tsi_out;
    T_out = T_from_DP(v.u, rho_out, p_out, Yo, &Tsi_out);
    }
    // ---- equilibria of the target state (calculate_feq_geq, lbm-edf.cpp:198-209)
    const double Tso = v.u.si_temp(T_out);
    uint32_t unknown = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint8_t tq = v.type[cell(v.d, x.i - cx(q), x.j - cy(q), x.k - cz(q))];
        if (tq == TYPE_O || tq == TYPE_O_C) unknown |= 1u << q;
    }
    auto R = [&](int f) -> double& { return rec[(int64_t)f * count + t]; };
    R(TR_RHO_OUT) = rho_out;
#pragma unroll
    for (int d = 0; d < 3; ++d) R(TR_VOUT + d) = vout[d];
    R(TR_TSO) = Tso;
    R(TR_U_OUT) = v.u.energy_mass(u_mass(Tso, Yo));
    R(TR_TH_OUT) = v.u.energy_mass(R_UNIV * Tso / mean_W(Yo));
    R(TR_MASK) = (double)unknown;
#pragma unroll
    for (int a = 0; a < NS; ++a) R(TR_RA_OUT + a) = ra_out[a];
}

// grid (count / 64, NS): local moments of species blockIdx.y
template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_local(MultiView<P> v, const int* cells, int count, double* rec)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int a = blockIdx.y;
    const int c = cells[t];
    auto R = [&](int f) -> double& { return rec[(int64_t)f * count + t]; };
    const uint32_t unknown = (uint32_t)R(TR_MASK);
    const double vout[3] = {R(TR_VOUT), R(TR_VOUT + 1), R(TR_VOUT + 2)};
    const double ra_out = R(TR_RA_OUT + a);
    const double zero[3] = {0, 0, 0};
    const double th_a = v.u.energy_mass(R_UNIV * R(TR_TSO) / W[a]);
    Axis3<double> Fo[3];
    feq_axes<double>(Fo, vout, th_a, zero);
    double r = 0, m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double f = (unknown >> q & 1u) ? feq_dir<double>(Fo, q, ra_out) : ldf_in(v, a, q, c);
        r += f;
        if (cx(q)) m0 += cx(q) * f;
        if (cy(q)) m1 += cy(q) * f;
        if (cz(q)) m2 += cz(q) * f;
    }
    R(TR_RA_LOC + a) = r;
    R(TR_MA + 3 * a) = m0;
    R(TR_MA + 3 * a + 1) = m1;
    R(TR_MA + 3 * a + 2) = m2;
}

template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_close(MultiView<P> v, const int* cells, int count, double* rec)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t];
    double* const* Fd = v.fld;
    auto R = [&](int f) -> double& { return rec[(int64_t)f * count + t]; };
    const uint32_t unknown = (uint32_t)R(TR_MASK);
    const double rho_out = R(TR_RHO_OUT), th_out = R(TR_TH_OUT), U_out = R(TR_U_OUT);
    const double vout[3] = {R(TR_VOUT), R(TR_VOUT + 1), R(TR_VOUT + 2)};
    double re_loc = 0, rho_loc = 0, mloc[3] = {0, 0, 0};
    const GeqAxes<double> Go = geq_axes<double>(th_out, vout);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        re_loc += (unknown >> q & 1u) ? geq_dir<double>(Go, q, rho_out, U_out - 1.5 * th_out)
                                      : ldg(v, v.g, q, c);
    double ra_loc[NS], ma[NS][3];
    for (int a = 0; a < NS; ++a) {
        ra_loc[a] = R(TR_RA_LOC + a);
        rho_loc += ra_loc[a];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ma[a][d] = R(TR_MA + 3 * a + d);
            mloc[d] += ma[a][d];
            ma[a][d] = ra_loc[a] > 0.0 ? ma[a][d] / ra_loc[a] : 0.0;
            R(TR_MA + 3 * a + d) = ma[a][d];
        }
    }
    double vloc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vloc[d] = rho_loc > 0 ? mloc[d] / rho_loc : 0.0;
    double Uloc = re_loc / rho_loc;
    for (int a = 0; a < NS; ++a)
        Uloc = Uloc - 0.5 * ra_loc[a] / rho_loc * (ma[a][0] * ma[a][0] + ma[a][1] * ma[a][1] + ma[a][2] * ma[a][2]);
    // calculate_temp (lbm-moment.cpp:540-553)
    double Yl[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yl[a] = ra_loc[a] / rho_loc;
    normalize(Yl);
    const double T_loc = v.u.temp(T_from_UV(v.u.si_energy_mass(Uloc), Yl, v.u.si_temp(Fd[LBM_FIELD_T][c])));
    const double Tsl = v.u.si_temp(T_loc);
    R(TR_RHO_LOC) = rho_loc;
#pragma unroll
    for (int d = 0; d < 3; ++d) R(TR_VLOC + d) = vloc[d];
    R(TR_TSL) = Tsl;
    R(TR_U_LOC) = v.u.energy_mass(u_mass(Tsl, Yl));
    R(TR_TH_LOC) = v.u.energy_mass(R_UNIV * Tsl / mean_W(Yl));
}

// grid (count / 64, NS + 1): reconstruction of species blockIdx.y (NS: the
// energy populations); unknown 2 f_out - f_loc, known f_out + f - f_loc
template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_rebuild(MultiView<P> v, const int* cells, int count, const double* rec)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int a = blockIdx.y;
    const int c = cells[t];
    auto R = [&](int f) { return rec[(int64_t)f * count + t]; };
    const uint32_t unknown = (uint32_t)R(TR_MASK);
    const double vout[3] = {R(TR_VOUT), R(TR_VOUT + 1), R(TR_VOUT + 2)};
    if (a < NS) {
        const double zero[3] = {0, 0, 0};
        const double ma[3] = {R(TR_MA + 3 * a), R(TR_MA + 3 * a + 1), R(TR_MA + 3 * a + 2)};
        const double ra_out = R(TR_RA_OUT + a), ra_loc = R(TR_RA_LOC + a);
        const double tho = v.u.energy_mass(R_UNIV * R(TR_TSO) / W[a]);
        const double thl = v.u.energy_mass(R_UNIV * R(TR_TSL) / W[a]);
        Axis3<double> Fo[3], Fl[3];
        feq_axes<double>(Fo, vout, tho, zero);
        feq_axes<double>(Fl, ma, thl, zero);
        // every load before the first store: the stores go to the same
        // population set, so interleaved they would serialise the loads
        double fin[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) fin[q] = ldf_in(v, a, q, c);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double fo = feq_dir<double>(Fo, q, ra_out);
            const double flc = feq_dir<double>(Fl, q, ra_loc);
            stf_in(v, a, q, c, (unknown >> q & 1u) ? 2 * fo - flc : fo + fin[q] - flc);
        }
        return;
    }
    const double rho_out = R(TR_RHO_OUT), th_out = R(TR_TH_OUT), U_out = R(TR_U_OUT);
    const double rho_loc = R(TR_RHO_LOC), th_loc = R(TR_TH_LOC), U_loc = R(TR_U_LOC);
    const double vloc[3] = {R(TR_VLOC), R(TR_VLOC + 1), R(TR_VLOC + 2)};
    const GeqAxes<double> Go = geq_axes<double>(th_out, vout);
    const GeqAxes<double> Gl = geq_axes<double>(th_loc, vloc);
    double gin[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) gin[q] = ldg(v, v.g, q, c);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double go = geq_dir<double>(Go, q, rho_out, U_out - 1.5 * th_out);
        const double gl = geq_dir<double>(Gl, q, rho_loc, U_loc - 1.5 * th_loc);
        stg(v, v.g, q, c, (unknown >> q & 1u) ? 2 * go - gl : go + gin[q] - gl);
    }
}

// ------------------------------------------------- TMS_BC inflow sweep
// Inflow sweep of TMS_BC (lbm-bc.cpp:463-938) for TYPE_I_C ghosts: the
// inflow target from the inline LODI relaxation with sigma = 5 towards the
// ghost's state (lbm-bc.cpp:519-677), the density from setState_TP at the
// target temperature and pressure (lbm-bc.cpp:672-680); then unknown
// directions f_tgt + f_in - f_loc and known ones f_tgt + f - f_loc, with
// f_in the equilibrium of the inflow state, f_loc that of the local moments
// (inflow populations substituted) and f_tgt that of the local density and
// composition at the inflow velocity and temperature.
template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_in(MultiView<P> v, const int* cells, const int* dirs, int count)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t], l = dirs[t];
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    const int strides[3] = {v.d.ny * v.d.nz, v.d.nz, 1};
    const int dims[3] = {v.d.nx, v.d.ny, v.d.nz};
    const int cl[3] = {cx(l), cy(l), cz(l)};
    const int idir = cl[0] ? 0 : (cl[1] ? 1 : 2);
    const int isign = cl[idir];
    const int st = strides[idir];
    const int c1 = c + isign * st, cb = c - isign * st;
    double* const* Fd = v.fld;
    double rho_in, vin[3], p_in, ra_in[NS], Yi[NS], T_in;
    if (v.type[cb] == TYPE_I) {
        // TYPE_I target (lbm-bc.cpp:506-518): velocity, temperature and
        // composition of the ghost; the pressure of the cell downstream after
        // calculate_moment_point
        mc_moment_point(v, c1);
        vin[0] = Fd[LBM_FIELD_U][cb];
        vin[1] = Fd[LBM_FIELD_V][cb];
        vin[2] = Fd[LBM_FIELD_W][cb];
        T_in = Fd[LBM_FIELD_T][cb];
        rho_in = Fd[LBM_FIELD_RHO][cb];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            ra_in[a] = v.rhoa[a * n + cb];
            Yi[a] = ra_in[a] / rho_in;
        }
        normalize(Yi);
        p_in = Fd[LBM_FIELD_P][c1];
    } else {
    const double sigma = 5.0;
    // ---- inline LODI target
    const double rho = Fd[LBM_FIELD_RHO][c], p = Fd[LBM_FIELD_P][c];
    const double vel[3] = {Fd[LBM_FIELD_U][c], Fd[LBM_FIELD_V][c], Fd[LBM_FIELD_W][c]};
    const double velb[3] = {Fd[LBM_FIELD_U][cb], Fd[LBM_FIELD_V][cb], Fd[LBM_FIELD_W][cb]};
    double Yc[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yc[a] = v.rhoa[a * n + c] / rho;
    double a_s, gam;
    double Yn[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yn[a] = Yc[a];
    sound_DP(v.u, rho, p, Yn, &a_s, &gam);
    normalize(Yn);  // the DP state's mass fractions (setMassFractions)
    // fd_uw (FD.cpp:64-74) with the interface direction as the velocity
    auto fd_uw = [&](double s0, double s1) { return isign < 0 ? (s0 - s1) / 1 : (-s0 + s1) / 1; };
    const double du = fd_uw(Fd[LBM_FIELD_U + idir][c], Fd[LBM_FIELD_U + idir][c1]);
    const double dp = fd_uw(p, Fd[LBM_FIELD_P][c1]);
    const double rho1 = Fd[LBM_FIELD_RHO][c1];
    const double un = vel[idir];
    // tangential components as paired per normal: x -> (v, w), y -> (w, u), z -> (u, v)
    const int t3 = idir == 0 ? 1 : (idir == 1 ? 2 : 0);
    const int t4 = idir == 0 ? 2 : (idir == 1 ? 0 : 1);
    const double Nd = dims[idir];  // ghost-inclusive length
    const double Rmix = v.u.cp(R_UNIV / mean_W(Yn));
    double L1 = (un - a_s) * (dp - rho * a_s * du);
    const double L2 = sigma * a_s / Nd * rho * Rmix * (Fd[LBM_FIELD_T][c] - Fd[LBM_FIELD_T][cb]);
    const double L3 = sigma * a_s / Nd * (vel[t3] - velb[t3]);
    const double L4 = sigma * a_s / Nd * (vel[t4] - velb[t4]);
    double L5 = (un + a_s) * (dp + rho * a_s * du);
    const double K = sigma * a_s * a_s / Nd * rho * (1.0 - (un / a_s) * (un / a_s));
    // the z-normal branch tests cy (lbm-bc.cpp:652), so it always takes the second form
    const bool left = idir == 2 ? (cy(l) > 0) : (isign > 0);
    if (left) L5 = K * (un - velb[idir]);
    else L1 = K * (un - velb[idir]);
    const double d1 = 1.0 / (a_s * a_s) * (L2 + 0.5 * (L5 + L1));
    const double d2 = 1.0 / (2.0 * rho * a_s) * (L5 - L1);
    const double d5 = 0.5 * (L5 + L1);
    rho_in = rho - 1 * d1;
#pragma unroll
    for (int d = 0; d < 3; ++d) vin[d] = vel[d];
    vin[idir] = vel[idir] - 1 * d2;
    vin[t3] = vel[t3] - 1 * L3;
    vin[t4] = vel[t4] - 1 * L4;
    p_in = p - 1 * d5;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
        const double dY = fd_uw(Yc[a], v.rhoa[a * n + c1] / rho1);
        ra_in[a] = (Yc[a] - 1 * un * dY) * rho_in;
        Yi[a] = ra_in[a] / rho_in;
    }
    double Tsi_in;
    T_in = T_from_DP(v.u, rho_in, p_in, Yi, &Tsi_in);  // normalises Yi
    }
    // setState_TP(T_in, p_in): the inflow density follows
    const double Tsi = v.u.si_temp(T_in);
    rho_in = v.u.rho(v.u.si_p(p_in) * mean_W(Yi) / (R_UNIV * Tsi));
#pragma unroll
    for (int a = 0; a < NS; ++a) ra_in[a] = Yi[a] * rho_in;
    // ---- equilibria of the inflow state (calculate_feq_geq, lbm-edf.cpp:198-209)
    const double U_in = v.u.energy_mass(u_mass(Tsi, Yi));
    const double th_in = v.u.energy_mass(R_UNIV * Tsi / mean_W(Yi));
    uint32_t unknown = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint8_t tq = v.type[cell(v.d, x.i - cx(q), x.j - cy(q), x.k - cz(q))];
        if (tq == TYPE_I || tq == TYPE_I_C) unknown |= 1u << q;
    }
    const double zero[3] = {0, 0, 0};
    // ---- local moments with the inflow populations in the unknown directions
    double ra_loc[NS], ma[NS][3], re_loc = 0;
    double rho_loc = 0, mloc[3] = {0, 0, 0};
    for (int a = 0; a < NS; ++a) {
        const double th_a = v.u.energy_mass(R_UNIV * Tsi / W[a]);
        Axis3<double> Fi[3];
        feq_axes<double>(Fi, vin, th_a, zero);
        double r = 0, m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double f = (unknown >> q & 1u) ? feq_dir<double>(Fi, q, ra_in[a]) : ldf_in(v, a, q, c);
            r += f;
            if (cx(q)) m0 += cx(q) * f;
            if (cy(q)) m1 += cy(q) * f;
            if (cz(q)) m2 += cz(q) * f;
        }
        ra_loc[a] = r;
        ma[a][0] = m0;
        ma[a][1] = m1;
        ma[a][2] = m2;
    }
    const GeqAxes<double> Gi = geq_axes<double>(th_in, vin);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        re_loc += (unknown >> q & 1u) ? geq_dir<double>(Gi, q, rho_in, U_in - 1.5 * th_in) : ldg(v, v.g, q, c);
    for (int a = 0; a < NS; ++a) {
        rho_loc += ra_loc[a];
        mloc[0] += ma[a][0];
        mloc[1] += ma[a][1];
        mloc[2] += ma[a][2];
#pragma unroll
        for (int d = 0; d < 3; ++d) ma[a][d] = ra_loc[a] > 0.0 ? ma[a][d] / ra_loc[a] : 0.0;
    }
    double vloc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vloc[d] = rho_loc > 0 ? mloc[d] / rho_loc : 0.0;
    double Uloc = re_loc / rho_loc;
    for (int a = 0; a < NS; ++a)
        Uloc = Uloc - 0.5 * ra_loc[a] / rho_loc * (ma[a][0] * ma[a][0] + ma[a][1] * ma[a][1] + ma[a][2] * ma[a][2]);
    // calculate_temp (lbm-moment.cpp:540-553)
    double Yl[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yl[a] = ra_loc[a] / rho_loc;
    normalize(Yl);
    const double T_loc = v.u.temp(T_from_UV(v.u.si_energy_mass(Uloc), Yl, v.u.si_temp(Fd[LBM_FIELD_T][c])));
    const double Tsl = v.u.si_temp(T_loc);
    const double U_loc = v.u.energy_mass(u_mass(Tsl, Yl));
    const double th_loc = v.u.energy_mass(R_UNIV * Tsl / mean_W(Yl));
    // target: local density and composition at the inflow velocity and temperature
    const double U_tg = v.u.energy_mass(u_mass(Tsi, Yl));
    const double th_tg = v.u.energy_mass(R_UNIV * Tsi / mean_W(Yl));
    // ---- reconstruction
    for (int a = 0; a < NS; ++a) {
        const double thi = v.u.energy_mass(R_UNIV * Tsi / W[a]);
        const double thl = v.u.energy_mass(R_UNIV * Tsl / W[a]);
        Axis3<double> Fi[3], Fl[3];
        feq_axes<double>(Fi, vin, thi, zero);
        feq_axes<double>(Fl, ma[a], thl, zero);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double fi = feq_dir<double>(Fi, q, ra_in[a]);
            const double ft = feq_dir<double>(Fi, q, ra_loc[a]);
            const double flc = feq_dir<double>(Fl, q, ra_loc[a]);
            const double fin = ldf_in(v, a, q, c);
            stf_in(v, a, q, c, (unknown >> q & 1u) ? ft + fi - flc : ft + fin - flc);
        }
    }
    const GeqAxes<double> Gl = geq_axes<double>(th_loc, vloc);
    const GeqAxes<double> Gt = geq_axes<double>(th_tg, vin);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double gi = geq_dir<double>(Gi, q, rho_in, U_in - 1.5 * th_in);
        const double gt = geq_dir<double>(Gt, q, rho_loc, U_tg - 1.5 * th_tg);
        const double gl = geq_dir<double>(Gl, q, rho_loc, U_loc - 1.5 * th_loc);
        const double gin = ldg(v, v.g, q, c);
        stg(v, v.g, q, c, (unknown >> q & 1u) ? gt + gi - gl : gt + gin - gl);
    }
}

// ------------------------------------------------- TMS_BC wall sweep
// Isothermal-wall sweep of TMS_BC (lbm-bc.cpp:941-1109): every link with a
// TYPE_S cell upstream and fluid downstream is an interface link.  Target:
// the cell's own species densities (sum of its populations), the mixture and
// species velocities and the temperature interpolated at q = 1/2 between the
// cell downstream and the wall cell, averaged over the links (species: the
// downstream species velocity with the wall's mixture velocity); unknown
// directions 2 f_tgt - f_loc, known f_tgt + f - f_loc.
template <typename P>
__global__ __launch_bounds__(64) void k_mc_tms_wall(MultiView<P> v, const int* cells, int count)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    const int c = cells[t];
    const int64_t n = v.d.n;
    const Ijk x = ijk(v.d, c);
    double* const* Fd = v.fld;
    uint32_t unknown = 0, links = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint8_t tu = v.type[cell(v.d, x.i - cx(q), x.j - cy(q), x.k - cz(q))];
        const uint8_t td = v.type[cell(v.d, x.i + cx(q), x.j + cy(q), x.k + cz(q))];
        if (tu == TYPE_S) {
            unknown |= 1u << q;
            if (td == TYPE_F) links |= 1u << q;
        }
    }
    // ---- target state
    double ra_bb[NS], rho_bb = 0;
    for (int a = 0; a < NS; ++a) {
        double r = 0;
        for (int q = 0; q < Q; ++q) r += ldf_in(v, a, q, c);
        ra_bb[a] = r;
        rho_bb += r;
    }
    double vt[3] = {0, 0, 0}, vat[NS][3], Tt = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a) vat[a][0] = vat[a][1] = vat[a][2] = 0.0;
    int nd = 0;
    const double qw = 0.5;
    for (int q = 0; q < Q; ++q) {
        if (!((links >> q) & 1u)) continue;
        ++nd;
        const int dn = cell(v.d, x.i + cx(q), x.j + cy(q), x.k + cz(q));
        const int up = cell(v.d, x.i - cx(q), x.j - cy(q), x.k - cz(q));
        double uw[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            uw[d] = Fd[LBM_FIELD_U + d][up];
            vt[d] += (qw * Fd[LBM_FIELD_U + d][dn] + uw[d]) / (1.0 + qw);
        }
        Tt += (qw * Fd[LBM_FIELD_T][dn] + Fd[LBM_FIELD_T][up]) / (1.0 + qw);
#pragma unroll
        for (int a = 0; a < NS; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) vat[a][d] += (qw * v.ua[(d * NS + a) * n + dn] + uw[d]) / (1.0 + qw);
    }
    const double inv = 1.0 / nd;
#pragma unroll
    for (int d = 0; d < 3; ++d) vt[d] = inv * vt[d];
    Tt = inv * Tt;
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) vat[a][d] = inv * vat[a][d];
    // equilibria of the target (calculate_feq_geq, lbm-edf.cpp:168-196)
    double Yt[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yt[a] = ra_bb[a] / rho_bb;
    normalize(Yt);
    const double Tst = v.u.si_temp(Tt);
    const double U_t = v.u.energy_mass(u_mass(Tst, Yt));
    const double th_t = v.u.energy_mass(R_UNIV * Tst / mean_W(Yt));
    const double zero[3] = {0, 0, 0};
    // ---- local moments with the target populations in the unknown directions
    double ra_loc[NS], ma[NS][3], re_loc = 0;
    double rho_loc = 0, mloc[3] = {0, 0, 0};
    for (int a = 0; a < NS; ++a) {
        const double th_a = v.u.energy_mass(R_UNIV * Tst / W[a]);
        Axis3<double> Ft[3];
        feq_axes<double>(Ft, vat[a], th_a, zero);
        double r = 0, m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double f = (unknown >> q & 1u) ? feq_dir<double>(Ft, q, ra_bb[a]) : ldf_in(v, a, q, c);
            r += f;
            if (cx(q)) m0 += cx(q) * f;
            if (cy(q)) m1 += cy(q) * f;
            if (cz(q)) m2 += cz(q) * f;
        }
        ra_loc[a] = r;
        ma[a][0] = m0;
        ma[a][1] = m1;
        ma[a][2] = m2;
    }
    const GeqAxes<double> Gt = geq_axes<double>(th_t, vt);
#pragma unroll
    for (int q = 0; q < Q; ++q)
        re_loc += (unknown >> q & 1u) ? geq_dir<double>(Gt, q, rho_bb, U_t - 1.5 * th_t) : ldg(v, v.g, q, c);
    for (int a = 0; a < NS; ++a) {
        rho_loc += ra_loc[a];
        mloc[0] += ma[a][0];
        mloc[1] += ma[a][1];
        mloc[2] += ma[a][2];
#pragma unroll
        for (int d = 0; d < 3; ++d) ma[a][d] = ra_loc[a] > 0 ? ma[a][d] / ra_loc[a] : 0.0;
    }
    double vloc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vloc[d] = rho_loc > 0 ? mloc[d] / rho_loc : 0.0;
    double Uloc = re_loc / rho_loc;
    for (int a = 0; a < NS; ++a)
        Uloc = Uloc - 0.5 * ra_loc[a] / rho_loc * (ma[a][0] * ma[a][0] + ma[a][1] * ma[a][1] + ma[a][2] * ma[a][2]);
    // calculate_temp (lbm-moment.cpp:540-553)
    double Yl[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) Yl[a] = ra_loc[a] / rho_loc;
    normalize(Yl);
    const double T_loc = v.u.temp(T_from_UV(v.u.si_energy_mass(Uloc), Yl, v.u.si_temp(Fd[LBM_FIELD_T][c])));
    const double Tsl = v.u.si_temp(T_loc);
    const double U_loc = v.u.energy_mass(u_mass(Tsl, Yl));
    const double th_loc = v.u.energy_mass(R_UNIV * Tsl / mean_W(Yl));
    // ---- reconstruction
    for (int a = 0; a < NS; ++a) {
        const double tht = v.u.energy_mass(R_UNIV * Tst / W[a]);
        const double thl = v.u.energy_mass(R_UNIV * Tsl / W[a]);
        Axis3<double> Ft[3], Fl[3];
        feq_axes<double>(Ft, vat[a], tht, zero);
        feq_axes<double>(Fl, ma[a], thl, zero);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double ft = feq_dir<double>(Ft, q, ra_bb[a]);
            const double flc = feq_dir<double>(Fl, q, ra_loc[a]);
            const double fin = ldf_in(v, a, q, c);
            stf_in(v, a, q, c, (unknown >> q & 1u) ? 2 * ft - flc : ft + fin - flc);
        }
    }
    const GeqAxes<double> Gl = geq_axes<double>(th_loc, vloc);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double gt = geq_dir<double>(Gt, q, rho_bb, U_t - 1.5 * th_t);
        const double gl = geq_dir<double>(Gl, q, rho_loc, U_loc - 1.5 * th_loc);
        const double gin = ldg(v, v.g, q, c);
        stg(v, v.g, q, c, (unknown >> q & 1u) ? 2 * gt - gl : gt + gin - gl);
    }
}

// ------------------------------------------- rest state of FP32 storage
// Per-block sums over TYPE_F cells of 1, T, p and X_a (lattice units), for
// the mean state the FP32 deviations are taken from, and the minima of X_a
// over the TYPE_F cells and the inflow ghosts (TYPE_I, TYPE_I_C).
// ------------------------------------------------------------ diagnostics
// Max Mach number of the multicomponent state with each cell's frozen
// gamma = c_p / (c_p - R_univ / W_mix) at its T and composition (the sound
// speed of impose_nscbc.cpp:358-369): |u| / sqrt(gamma p / rho).  Overwrites
// slot 5 of the diagnostics partials, one block per 256 cells as k_diag.
template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_diag_mach(MultiView<P> v, double* part)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    double mach = 0.0;
    if (c < v.d.n && v.type[c] == TYPE_F) {
        const int64_t n = v.d.n;
        double* const* Fd = v.fld;
        const double rho = Fd[LBM_FIELD_RHO][c];
        double Y[NS];
#pragma unroll
        for (int a = 0; a < NS; ++a) Y[a] = (double)v.rhoa[a * n + c] / rho;
        normalize(Y);
        const double Tsi = v.u.si_temp(Fd[LBM_FIELD_T][c]);
        const double cp = cp_mass(Tsi, Y);
        const double gam = cp / (cp - R_UNIV / mean_W(Y));
        const double u = Fd[LBM_FIELD_U][c], w = Fd[LBM_FIELD_W][c], vv = Fd[LBM_FIELD_V][c];
        const double m = sqrt(u * u + vv * vv + w * w) / sqrt(gam * Fd[LBM_FIELD_P][c] / rho);
        mach = isfinite(m) ? m : 0.0;
    }
    for (int o = 32; o > 0; o >>= 1) mach = fmax(mach, __shfl_down(mach, o, 64));
    __shared__ double sm[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mach;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int k = 1; k < BLOCK / 64; ++k) m = fmax(m, sm[k]);
        part[(size_t)blockIdx.x * 6 + 5] = m;
    }
}

// ------------------------------------------------------------- SMOOTHING
// LBM::Smoothing (lbm-smoothing.cpp:95-173) of species a, race-free: every
// population l of a TYPE_F or TYPE_I cell becomes f + alpha * lap(f) with
// fd_laplace (FD.cpp:40-50): per axis (r - 2c) + l, 0 when either neighbour
// is TYPE_A / TYPE_FS / TYPE_S / TYPE_P; an axis whose neighbours fall
// outside the array takes them as 0 of TYPE_F.  The results go to tmp
// [27][n] (Jacobi: the Laplacians read the populations before the call) and
// k_mc_smooth_store copies them back.  Layout 0 only (strict mode).
__device__ inline bool smooth_dead(uint8_t t) { return t == TYPE_A || t == TYPE_FS || t == TYPE_S || t == TYPE_P; }

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_smooth(MultiView<P> v, int a, P* tmp, double alpha)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    const uint8_t t = v.type[c];
    if (t != TYPE_F && t != TYPE_I) return;
    const Ijk x = ijk(v.d, c);
    const int pos[3] = {x.i, x.j, x.k}, ext[3] = {v.d.nx, v.d.ny, v.d.nz};
    const int stride[3] = {v.d.ny * v.d.nz, v.d.nz, 1};
    for (int l = 0; l < Q; ++l) {
        const double fc = ldf(v, v.f, a, l, c);
        double lap = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double fl = 0.0, fr = 0.0;
            uint8_t tl = TYPE_F, tr = TYPE_F;
            if (pos[d] - 1 >= 0 && pos[d] + 1 < ext[d]) {
                fl = ldf(v, v.f, a, l, c - stride[d]);
                fr = ldf(v, v.f, a, l, c + stride[d]);
                tl = v.type[c - stride[d]];
                tr = v.type[c + stride[d]];
            }
            const double d2 = (smooth_dead(tl) || smooth_dead(tr)) ? 0.0 : (fr - 2.0 * fc + fl) / (1 * 1);
            lap = d == 0 ? d2 : lap + d2;
        }
        double r = fc + alpha * lap;
        if constexpr (sizeof(P) == 4) r -= fshift(v, a, l);
        tmp[(int64_t)l * v.d.n + c] = (P)r;
    }
}

template <typename P>
__global__ __launch_bounds__(BLOCK) void k_mc_smooth_store(MultiView<P> v, int a, const P* tmp)
{
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= v.d.n) return;
    const uint8_t t = v.type[c];
    if (t != TYPE_F && t != TYPE_I) return;
#pragma unroll
    for (int l = 0; l < Q; ++l) spop(v.f, a, l, v.d.n, c) = tmp[(int64_t)l * v.d.n + c];
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_mc_state_sums(const uint8_t* type, const double* T, const double* p,
                                                         const S* X, int64_t n, double* part)
{
    __shared__ double red[BLOCK];
    double acc[MC_STATS] = {};
#pragma unroll
    for (int a = 0; a < NS; ++a) acc[MC_SUMS + a] = HUGE_VAL;
    for (int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < n; c += (int64_t)gridDim.x * BLOCK) {
        if (type[c] == TYPE_I || type[c] == TYPE_I_C) {  // inflow ghosts: TMS_BC imposes their composition
#pragma unroll
            for (int a = 0; a < NS; ++a) acc[MC_SUMS + a] = fmin(acc[MC_SUMS + a], (double)X[a * n + c]);
            continue;
        }
        if (type[c] != TYPE_F) continue;
        acc[0] += 1.0;
        acc[1] += T[c];
        acc[2] += p[c];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            const double x = X[a * n + c];
            acc[3 + a] += x;
            acc[MC_SUMS + a] = fmin(acc[MC_SUMS + a], x);
        }
    }
    for (int q = 0; q < MC_STATS; ++q) {
        red[threadIdx.x] = acc[q];
        __syncthreads();
        for (int w = BLOCK / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                red[threadIdx.x] = q < MC_SUMS ? red[threadIdx.x] + red[threadIdx.x + w]
                                               : fmin(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * MC_STATS + q] = red[0];
        __syncthreads();
    }
}

template <typename S>
void launch_mc_state_sums(const uint8_t* type, const double* T, const double* p, const S* X, int64_t n,
                          double* part, hipStream_t s)
{
    hipLaunchKernelGGL(k_mc_state_sums<S>, dim3(MC_SUM_BLOCKS), dim3(BLOCK), 0, s, type, T, p, X, n, part);
}
template void launch_mc_state_sums<float>(const uint8_t*, const double*, const double*, const float*, int64_t,
                                          double*, hipStream_t);
template void launch_mc_state_sums<double>(const uint8_t*, const double*, const double*, const double*, int64_t,
                                           double*, hipStream_t);

// ------------------------------------------------------------- launchers
static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// FP32 contexts hold at most 16 species (the moment pass stages per-species
// sums in LDS) and lbm_create refuses more: a build for a larger mechanism
// compiles the FP64 kernels only, and the surrogate's kernels only for
// mechanisms its 16-wide layers hold.
template <typename P>
constexpr bool kernels_built()
{
    return sizeof(P) == 8 || NS <= 16;
}

template <typename P>
void launch_mc(McKernel k, const MultiView<P>& v, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        switch (k) {
        case McKernel::Init: hipLaunchKernelGGL(k_mc_init<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v); break;
        case McKernel::Chemistry: hipLaunchKernelGGL(k_mc_chemistry<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v); break;
        case McKernel::SpeciesTransport:
            hipLaunchKernelGGL(k_mc_species_transport<P>, dim3(xcd_grid(v.d.n, BLOCK)), dim3(BLOCK), 0, s, v);
            break;
        case McKernel::CollideSpecies:
            hipLaunchKernelGGL((k_mc_collide_species<P, false>), dim3(xcd_species_grid(v.d.n, BLOCK, NS)), dim3(BLOCK), 0,
                               s, v);
            break;
        case McKernel::CollideSpeciesPush:
            hipLaunchKernelGGL((k_mc_collide_species<P, true>), dim3(xcd_species_grid(v.d.n, BLOCK, NS)), dim3(BLOCK), 0,
                               s, v);
            if (v.nside > 0)
                hipLaunchKernelGGL(k_mc_side_push<P>, dim3((unsigned)((v.nside + BLOCK - 1) / BLOCK), NS), dim3(BLOCK), 0,
                                   s, v);
            break;
        case McKernel::Collide:
            if (v.use_omix) {
                hipLaunchKernelGGL((k_mc_collide<P, 1>), grid1(v.d.n), dim3(BLOCK), 0, s, v);
                hipLaunchKernelGGL((k_mc_collide<P, 2>), grid1(v.d.n), dim3(BLOCK), 0, s, v);
            } else {
                hipLaunchKernelGGL((k_mc_collide<P, 0>), grid1(v.d.n), dim3(BLOCK), 0, s, v);
            }
            break;
        case McKernel::StreamEnergy:
            hipLaunchKernelGGL(k_mc_stream<P>, dim3(xcd_grid(v.d.n, BLOCK), 1), dim3(BLOCK), 0, s, v, NS);
            break;
        case McKernel::SideFill:
            if (v.nside > 0)
                hipLaunchKernelGGL(k_mc_side_fill<P>, dim3((unsigned)((v.nside + BLOCK - 1) / BLOCK), NS), dim3(BLOCK), 0,
                                   s, v);
            break;
        case McKernel::Unstream:
            hipLaunchKernelGGL(k_mc_unstream<P>, dim3(xcd_grid(v.d.n, BLOCK), NS), dim3(BLOCK), 0, s, v);
            break;
        case McKernel::Stream:
            hipLaunchKernelGGL(k_mc_stream<P>, dim3(xcd_grid(v.d.n, BLOCK), NS + 1), dim3(BLOCK), 0, s, v, 0);
            break;
        case McKernel::Moments:
            if (v.swapped) hipLaunchKernelGGL((k_mc_moments<P, 1>), grid1(v.d.n), dim3(BLOCK), 0, s, v);
            else hipLaunchKernelGGL((k_mc_moments<P, 0>), grid1(v.d.n), dim3(BLOCK), 0, s, v);
            break;
        }
    }
}

template <typename P>
void launch_mc_tms(const MultiView<P>& v, const int* cells, const int* dirs, int count, double* rec, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        if (count <= 0) return;
        const unsigned nb = (unsigned)((count + 63) / 64);
        hipLaunchKernelGGL(k_mc_tms_target<P>, dim3(nb), dim3(64), 0, s, v, cells, dirs, count, rec);
        hipLaunchKernelGGL(k_mc_tms_local<P>, dim3(nb, NS), dim3(64), 0, s, v, cells, count, rec);
        hipLaunchKernelGGL(k_mc_tms_close<P>, dim3(nb), dim3(64), 0, s, v, cells, count, rec);
        hipLaunchKernelGGL(k_mc_tms_rebuild<P>, dim3(nb, NS + 1), dim3(64), 0, s, v, cells, count,
                           static_cast<const double*>(rec));
    }
}

int mc_tms_record_fields() { return TR_COUNT; }

template <typename P>
void launch_mc_tms_in(const MultiView<P>& v, const int* cells, const int* dirs, int count, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        if (count > 0) hipLaunchKernelGGL(k_mc_tms_in<P>, dim3((count + 63) / 64), dim3(64), 0, s, v, cells, dirs, count);
    }
}

template <typename P>
void launch_mc_tms_wall(const MultiView<P>& v, const int* cells, int count, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        if (count > 0) hipLaunchKernelGGL(k_mc_tms_wall<P>, dim3((count + 63) / 64), dim3(64), 0, s, v, cells, count);
    }
}

static int cu_count()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus;
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

template <typename P>
void launch_mc_aa_save(const MultiView<P>& v, bool first, int a0, int na, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, v.low_count);
        const dim3 g(xcd_grid(v.d.n, BLOCK), na);
        if (v.swapped) hipLaunchKernelGGL((k_mc_aa_save<P, 1>), g, dim3(BLOCK), 0, s, v, first ? 1 : 0, a0);
        else hipLaunchKernelGGL((k_mc_aa_save<P, 0>), g, dim3(BLOCK), 0, s, v, first ? 1 : 0, a0);
    }
}

template <typename P>
void launch_mc_diag_mach(const MultiView<P>& v, double* partials, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        hipLaunchKernelGGL(k_mc_diag_mach<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v, partials);
    }
}

template <typename P>
void launch_mc_smoothing(const MultiView<P>& v, void* tmp, double alpha, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        for (int a = 0; a < NS; ++a) {
            hipLaunchKernelGGL(k_mc_smooth<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v, a, static_cast<P*>(tmp), alpha);
            hipLaunchKernelGGL(k_mc_smooth_store<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v, a, static_cast<const P*>(tmp));
        }
    }
}

template <typename P>
void launch_mc_aa_collide(const MultiView<P>& v, int a0, int na, bool side, hipStream_t s)
{
    if constexpr (!kernels_built<P>()) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        const dim3 g(xcd_species_grid(v.d.n, BLOCK, na)), gl(4 * cu_count()),
            gs((unsigned)((v.nside + BLOCK - 1) / BLOCK), NS);
        if (v.swapped) {
            if (na > 0) {
                hipLaunchKernelGGL((k_mc_collide_species_aa<P, 1>), g, dim3(BLOCK), 0, s, v, a0, na);
                hipLaunchKernelGGL((k_mc_collide_species_low<P, 1>), gl, dim3(BLOCK), 0, s, v);
            }
            if (side && v.nside > 0) hipLaunchKernelGGL((k_mc_side_push_aa<P, 1>), gs, dim3(BLOCK), 0, s, v);
        } else {
            if (na > 0) {
                hipLaunchKernelGGL((k_mc_collide_species_aa<P, 0>), g, dim3(BLOCK), 0, s, v, a0, na);
                hipLaunchKernelGGL((k_mc_collide_species_low<P, 0>), gl, dim3(BLOCK), 0, s, v);
            }
            if (side && v.nside > 0) hipLaunchKernelGGL((k_mc_side_push_aa<P, 0>), gs, dim3(BLOCK), 0, s, v);
        }
    }
}


template <typename P>
void launch_mc_sur_chemistry(const MultiView<P>& v, const sur::Mlp& net, uint16_t* X, float* Y, uint8_t* flag,
                             int64_t rows, unsigned* list, unsigned* count, hipStream_t s)
{
    if constexpr (!kernels_built<P>() || !sur::ENABLED) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, count);
        hipLaunchKernelGGL(k_mc_sur_features<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v, X, flag);
        sur::mlp_infer_bf16(net, X, rows, Y, s);
        hipLaunchKernelGGL(k_mc_sur_decode<P>, grid1(v.d.n), dim3(BLOCK), 0, s, v, Y, flag, list, count);
    }
}

template <typename P>
void launch_mc_chem_list(const MultiView<P>& v, const unsigned* list, const unsigned* count, hipStream_t s)
{
    if constexpr (!kernels_built<P>() || !sur::ENABLED) {
        return;  // refused before launch (lbm_create, surrogate entry points)
    } else {
        hipLaunchKernelGGL(k_mc_chem_list<P>, dim3(4 * cu_count()), dim3(BLOCK), 0, s, v, list, count);
    }
}

template void launch_mc<double>(McKernel, const MultiView<double>&, hipStream_t);
template void launch_mc<float>(McKernel, const MultiView<float>&, hipStream_t);
template void launch_mc_diag_mach<double>(const MultiView<double>&, double*, hipStream_t);
template void launch_mc_diag_mach<float>(const MultiView<float>&, double*, hipStream_t);
template void launch_mc_smoothing<double>(const MultiView<double>&, void*, double, hipStream_t);
template void launch_mc_smoothing<float>(const MultiView<float>&, void*, double, hipStream_t);
template void launch_mc_aa_save<double>(const MultiView<double>&, bool, int, int, hipStream_t);
template void launch_mc_aa_save<float>(const MultiView<float>&, bool, int, int, hipStream_t);
template void launch_mc_aa_collide<double>(const MultiView<double>&, int, int, bool, hipStream_t);
template void launch_mc_aa_collide<float>(const MultiView<float>&, int, int, bool, hipStream_t);
template void launch_mc_tms<float>(const MultiView<float>&, const int*, const int*, int, double*, hipStream_t);
template void launch_mc_sur_chemistry<float>(const MultiView<float>&, const sur::Mlp&, uint16_t*, float*,
                                             uint8_t*, int64_t, unsigned*, unsigned*, hipStream_t);
template void launch_mc_chem_list<float>(const MultiView<float>&, const unsigned*, const unsigned*, hipStream_t);
template void launch_mc_sur_chemistry<double>(const MultiView<double>&, const sur::Mlp&, uint16_t*, float*,
                                              uint8_t*, int64_t, unsigned*, unsigned*, hipStream_t);
template void launch_mc_chem_list<double>(const MultiView<double>&, const unsigned*, const unsigned*, hipStream_t);
template void launch_mc_tms<double>(const MultiView<double>&, const int*, const int*, int, double*, hipStream_t);
template void launch_mc_tms_in<double>(const MultiView<double>&, const int*, const int*, int, hipStream_t);
template void launch_mc_tms_wall<double>(const MultiView<double>&, const int*, int, hipStream_t);
template void launch_mc_tms_wall<float>(const MultiView<float>&, const int*, int, hipStream_t);
template void launch_mc_tms_in<float>(const MultiView<float>&, const int*, const int*, int, hipStream_t);

}  // namespace lbm
