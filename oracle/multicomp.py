"""CPU oracle of the multicomponent (MULTICOMP) reacting D3Q27 path, numpy.

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels; the product never
imports it.

Restates, cell by cell (vectorised over the fluid cells):
  LBM::Init             lbm-init.cpp:85-146
  LBM::Collide_Species  lbm-collision.cpp:334-526 (two explicit-Euler chemistry
                        sub-steps with a UV temperature re-solve, Wilke species
                        viscosities, Maxwell-Stefan friction, 3 equilibria)
  LBM::Collide          lbm-collision.cpp:191-332 (mixture energy population)
  LBM::Streaming        lbm-streaming.cpp:3-119 (species f and mixture g)
  LBM::calculate_moment lbm-moment.cpp:193-377 (SPECIES_MIN clamp,
                        Maxwell-Stefan velocities by Cholesky, T from energy)
with Units_Conv's binary32 conversions (units.hpp:75-131) at every crossing
between lattice and SI units, as the reference rounds them.  The thermo,
transport and kinetics come from oracle/chem.py.
Parity of this module against the reference binary is unpinned (the
reference's MULTICOMP build needs Cantera and Eigen, both absent); it is
pinned by conservation identities and closed forms in
tests/test_oracle_multicomp.py.
"""
import numpy as np

from .chem import Mechanism

F32 = np.float32
SPECIES_MIN = 1e-8
TYPE_F, TYPE_S, TYPE_P, TYPE_A, TYPE_FS, TYPE_I, TYPE_O, TYPE_I_C, TYPE_O_C, TYPE_Q = range(10)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])
C3 = np.stack([CX, CY, CZ])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25])


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


class Units:
    """Units_Conv with set_m_kg_s(dx, dt) (units.hpp:32-44): m = dx, kg = m^3,
    s = dt, K = 12000, all stored as float; conversions follow the C++
    promotions of units.hpp:75-131 (float arithmetic, or double where a
    double cb()/sq() enters, then rounded to float on return)."""

    def __init__(self, dx, dt, K=12000.0):
        self.m = F32(np.float64(F32(dx)) / 1.0)
        m64 = np.float64(self.m)
        self.kg = F32(1.0 / 1.0 * (m64 * m64 * m64))
        self.s = F32(np.float64(F32(dt)) / 1.0)
        self.K = F32(K)

    # SI -> lattice
    def temp(self, x):
        return f32(x) / self.K

    def energy_mass(self, x):
        return f32(x) * (self.s * self.s) / (self.m * self.m)

    def u(self, x):
        return f32(x) * self.s / self.m

    def rho(self, x):
        m = np.float64(self.m)
        return f32(f32(x).astype(np.float64) * (m * m * m) / np.float64(self.kg))

    def rho_dot(self, x):
        m = np.float64(self.m)
        return f32(f32(x).astype(np.float64) * (m * m * m) * np.float64(self.s) / np.float64(self.kg))

    def nu(self, x):
        m = np.float64(self.m)
        return f32((f32(x) * self.s).astype(np.float64) / (m * m))

    def mu(self, x):
        return f32(x) * self.s * self.m / self.kg

    def thermal_conductivity(self, x):
        return f32(x) * self.s * self.s * self.s * self.K / (self.kg * self.m)

    def cp(self, x):
        return f32(x) * self.s * self.s * self.K / (self.m * self.m)

    def p(self, x):
        s = np.float64(self.s)
        return f32(f32(x).astype(np.float64) * (np.float64(self.m) * (s * s)) / np.float64(self.kg))

    # lattice -> SI
    def si_temp(self, x):
        return f32(x) * self.K

    def si_energy_mass(self, x):
        return f32(x) * (self.m * self.m) / (self.s * self.s)

    def si_rho(self, x):
        m = np.float64(self.m)
        return f32((f32(x) * self.kg).astype(np.float64) / (m * m * m))

    def si_p(self, x):
        s = np.float64(self.s)
        return f32((f32(x) * self.kg).astype(np.float64) / (np.float64(self.m) * (s * s)))


def feq(rho, vel, theta, corr):
    """calculate_feq (lbm.hpp:276-305) for all 27 directions; rho/theta (...),
    vel/corr (3, ...) -> (27, ...)."""
    out = []
    for l in range(27):
        r = np.array(rho, dtype=np.float64, copy=True)
        for d, c in enumerate((CX[l], CY[l], CZ[l])):
            P = theta + vel[d] * vel[d] + corr[d]
            r = r * ((1.0 - P) if c == 0 else ((vel[d] + P) / 2 if c == 1 else (-vel[d] + P) / 2))
        out.append(r)
    return np.stack(out)


def geq(rho, U, theta, vel):
    """calculate_geq(l, rho, U, theta, v) (lbm-edf.cpp:38-98) as the operator
    form (U - 3 theta/2) prod Psi + 1/2 sum_d (O_d^2 Psi_d) prod_{e!=d} Psi_e
    (pinned against the reference's polynomials in test_oracle_single.py)."""
    P, K = [], []
    for d in range(3):
        v = vel[d]
        v2 = v * v
        P.append({0: 1.0 - theta - v2, 1: 0.5 * (v + theta + v2), -1: 0.5 * (theta + v2 - v)})
        even = 3 * theta * theta + 6 * theta * v2 + v2 * v2
        odd = v2 * v + 3 * theta * v
        K.append({0: theta - 3 * theta * theta + v2 - 6 * theta * v2 - v2 * v2, 1: 0.5 * (even + odd),
                  -1: 0.5 * (even - odd)})
    out = []
    for l in range(27):
        cs = (CX[l], CY[l], CZ[l])
        px, py, pz = (P[d][cs[d]] for d in range(3))
        kx, ky, kz = (K[d][cs[d]] for d in range(3))
        out.append(rho * ((U - 1.5 * theta) * px * py * pz + 0.5 * (kx * py * pz + px * ky * pz + px * py * kz)))
    return np.stack(out)


def gstr_shift(dq):
    """g* - g_eq of calculate_gstr (lbm-edf.cpp:140-150): 1/2 c.dq on the six
    axis directions, 0 elsewhere; dq (3, ...) -> (27, ...)."""
    out = []
    for l in range(27):
        if abs(CX[l]) + abs(CY[l]) + abs(CZ[l]) == 1:
            out.append(0.5 * (CX[l] * dq[0] + CY[l] * dq[1] + CZ[l] * dq[2]))
        else:
            out.append(np.zeros_like(dq[0]))
    return np.stack(out)


WALLISH = (TYPE_S, TYPE_A, TYPE_FS, TYPE_P, TYPE_O, TYPE_O_C)
BLOCKED = WALLISH + (TYPE_I, TYPE_I_C)
FLUIDLIKE = (TYPE_F, TYPE_I, TYPE_O, TYPE_I_C, TYPE_O_C)


def fd_central(l, c, r, tl, tr):
    """FD.hpp:14-33 with dx = 1, elementwise."""
    wl = np.isin(tl, WALLISH)
    wr = np.isin(tr, WALLISH)
    out = np.where(wl, np.where(tr == TYPE_P, 0.0, r - c), np.where(wr, np.where(tl == TYPE_P, 0.0, c - l),
                                                                     (r - l) / 2.0))
    return out


def fd_fuw(l, c, r, vel, tl, tr):
    """FD.cpp:5-23 with dx = 1, elementwise."""
    return np.where(vel > 0, np.where(np.isin(tl, BLOCKED), 0.0, c - l), np.where(np.isin(tr, BLOCKED), 0.0, r - c))


class MultiState:
    """MULTICOMP state on a ghost-inclusive grid; arrays are flat over cells
    (index c = (i*ny + j)*nz + k)."""

    def __init__(self, nx, ny, nz, units, mech=None):
        self.mech = mech or Mechanism()
        self.units = units
        self.dims = (nx, ny, nz)
        n = nx * ny * nz
        ns = self.mech.ns
        self.n, self.ns = n, ns
        self.type = np.zeros(n, np.uint8)
        self.rho = np.ones(n)
        self.u = np.zeros(n)
        self.v = np.zeros(n)
        self.w = np.zeros(n)
        self.p = np.full(n, 1.0 / 3.0)
        self.T = np.full(n, 1.0 / 3.0)
        self.rhoe = np.zeros(n)
        self.ptens = np.zeros((3, 3, n))
        self.eflux = np.zeros((3, n))
        self.HRR = np.zeros(n)
        self.g = np.zeros((27, n))
        self.gpc = np.zeros((27, n))
        self.X = np.zeros((ns, n))
        self.rho_a = np.zeros((ns, n))
        self.u_a = np.zeros((3, ns, n))
        self.f = np.zeros((ns, 27, n))
        self.fpc = np.zeros((ns, 27, n))
        self.tau_warnings = 0

    # ----------------------------------------------------------- helpers
    def _idx(self, i, j, k):
        nx, ny, nz = self.dims
        return (i * ny + j) * nz + k

    def _ijk(self, c):
        nx, ny, nz = self.dims
        return c // (ny * nz), (c // nz) % ny, c % nz

    def _neighbours(self, c):
        """flat indices of the -/+ neighbour along x, y, z: arrays (3, 2, m)."""
        nx, ny, nz = self.dims
        s = np.array([ny * nz, nz, 1])
        return np.stack([np.stack([c - s[d], c + s[d]]) for d in range(3)])

    def _vel(self):
        return np.stack([self.u, self.v, self.w])

    def _Y_state(self, c):
        """Mass fractions handed to setMassFractions: rho_a / rho, normalised."""
        return self.mech.normalize_Y((self.rho_a[:, c] / self.rho[c]).T)

    # -------------------------------------------------------------- Init
    def init(self):
        U_ = self.units
        m = self.mech
        c = np.nonzero(np.isin(self.type, FLUIDLIKE))[0]
        X = self.X[:, c].T
        Y = m.Y_from_X(X / X.sum(axis=1, keepdims=True))
        Tsi = U_.si_temp(self.T[c]).astype(np.float64)
        psi = U_.si_p(self.p[c]).astype(np.float64)
        rho_si = m.density(Tsi, psi, Y)
        rho = U_.rho(rho_si).astype(np.float64)
        self.rho[c] = rho
        vel = self._vel()[:, c]
        Uint = U_.energy_mass(m.u_mass(Tsi, Y)).astype(np.float64)
        self.rhoe[c] = rho * (Uint + 0.5 * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
        theta = U_.energy_mass(m.R * Tsi / m.W_mean(Y)).astype(np.float64)
        self.g[:, c] = geq(rho, Uint, theta, vel)
        zero = np.zeros((3, len(c)))
        for a in range(self.ns):
            ra = np.maximum(Y[:, a] * rho, SPECIES_MIN)
            self.rho_a[a, c] = ra
            self.u_a[:, a, c] = vel
            th_a = U_.energy_mass(m.R * Tsi / m.W[a]).astype(np.float64)
            self.f[a, :, c] = feq(ra, vel, th_a, zero).T

    # --------------------------------------------------- Collide_Species
    def collide_species(self, reaction=True, chemistry=None):
        """Collide_Species; `chemistry(c, Tsi, rho_si, Y0, rho, u_si)` may
        return (src [NS][nc], T_final) to replace the plain integration (the
        surrogate path); the sources used are kept in self.src_chem and the
        final temperature in self.T_chem."""
        U_, m = self.units, self.mech
        c = np.nonzero(self.type == TYPE_F)[0]
        nc = len(c)
        rho = self.rho[c]
        Y0 = self._Y_state(c)
        Tsi = U_.si_temp(self.T[c]).astype(np.float64)
        rho_si = U_.si_rho(rho).astype(np.float64)
        rho_a = self.rho_a[:, c]
        ua = self.u_a[:, :, c]
        vmix = self._vel()[:, c]
        src = np.zeros((self.ns, nc))
        if reaction:
            Uint = self.rhoe[c] / rho
            for a in range(self.ns):
                Uint = Uint - 0.5 * rho_a[a] / rho * (ua[0, a] ** 2 + ua[1, a] ** 2 + ua[2, a] ** 2)
            u_si = U_.si_energy_mass(Uint).astype(np.float64)
            if chemistry is None:
                Tcur, Ycur = Tsi.copy(), Y0.copy()
                for _ in range(2):
                    wdot = m.net_production(Tcur, rho_si, Ycur)
                    for a in range(self.ns):
                        src[a] += 1.0 / 2.0 * U_.rho_dot(wdot[:, a] * m.W[a]).astype(np.float64)
                    Ycur = m.normalize_Y(((rho_a + src) / rho).T)
                    Tcur = m.T_from_UV(u_si, Ycur, Tcur)
            else:
                src, Tcur = chemistry(c, Tsi, rho_si, Y0, rho, u_si)
                src = np.array(src, np.float64)
            self.src_chem = np.zeros((self.ns, self.n))
            self.src_chem[:, c] = src
            self.T_chem = np.zeros(self.n)
            self.T_chem[c] = Tcur
            # exact temperature of the initial state (the surrogate's dT origin)
            self.T_exact0 = np.zeros(self.n)
            self.T_exact0[c] = m.T_from_UV(u_si, Y0, Tsi)
            hm = m.h_molar(Tcur)
            hrr = np.zeros(nc)
            for a in range(self.ns):
                hrr += -1.0 * U_.energy_mass(hm[:, a] / m.W[a]).astype(np.float64) * src[a]
            self.HRR[c] = hrr
        # species viscosities and Wilke weights at the (T, rho) state
        mu = m.species_viscosities(Tsi)  # (nc, ns)
        W = m.W
        phi = np.zeros((nc, self.ns, self.ns))
        for a in range(self.ns):
            for b in range(a, self.ns):
                f1 = 1.0 + np.sqrt(mu[:, a] / mu[:, b]) * np.sqrt(np.sqrt(W[b] / W[a]))
                phi[:, a, b] = f1 * f1 / np.sqrt(8.0 * (1.0 + W[a] / W[b]))
                phi[:, b, a] = mu[:, b] / mu[:, a] * W[a] / W[b] * phi[:, a, b]
        X = self.X[:, c]
        p = self.p[c]
        omega_a = np.zeros((self.ns, nc))
        for a in range(self.ns):
            den = np.zeros(nc)
            for b in range(self.ns):
                den = den + X[b] * phi[:, a, b]
            visc = X[a] * U_.mu(mu[:, a]).astype(np.float64) / den
            omega_a[a] = 2 * X[a] * p / (X[a] * p + 2 * visc)
        # binary diffusion at the state's pressure
        p_si = m.pressure(Tsi, rho_si, Y0)
        Dab = U_.nu(m.binary_diff(Tsi, p_si)).astype(np.float64)  # (nc, ns, ns)
        fa = self.f[:, :, c]
        ufa = np.einsum("dl,alc->dac", C3.astype(np.float64), fa) / rho_a[None]
        Fa = np.zeros((3, self.ns, nc))
        for a in range(self.ns):
            acc = np.zeros((3, nc))
            for b in range(self.ns):
                acc = acc + -1.0 * p * X[a] * X[b] / Dab[:, a, b] * (ua[:, a] - ua[:, b])
            Fa[:, a] = acc * 1.0 / rho_a[a]
        nb = self._neighbours(c)
        T = self.T
        for a in range(self.ns):
            Ra = np.float64(U_.cp(m.R / W[a]))
            theta = Ra * T[c]
            dQ = np.zeros((3, nc))
            for d in range(3):
                def Q(idx, d=d):
                    ud = self.u_a[d, a, idx]
                    r = self.rho_a[a, idx]
                    return r * ud * (1.0 - 3.0 * Ra * T[idx]) - r * (ud * ud * ud)
                dQ[d] = fd_central(Q(nb[d, 0]), Q(c), Q(nb[d, 1]), self.type[nb[d, 0]], self.type[nb[d, 1]])
            corr = 1 * (2.0 - omega_a[a]) / (2.0 * rho_a[a] * omega_a[a]) * dQ
            fe = feq(rho_a[a], ufa[:, a], theta, corr)
            fdu = feq(rho_a[a], ufa[:, a] + Fa[:, a], theta, corr)
            fr = feq(src[a], vmix, theta, np.zeros((3, nc)))
            skip = (rho_a[a] <= SPECIES_MIN)[None, :] & (fr <= 0.0)
            new = fa[a] + omega_a[a] * (fe - fa[a]) + (fdu - fe) + fr
            old = self.fpc[a][:, c]
            self.fpc[a][:, c] = np.where(skip, old, new)

    # ------------------------------------------------------------ Collide
    def collide(self):
        U_, m = self.units, self.mech
        c = np.nonzero(self.type == TYPE_F)[0]
        nc = len(c)
        rho = self.rho[c]
        Y0 = self._Y_state(c)
        Tsi = U_.si_temp(self.T[c]).astype(np.float64)
        Xs = m.X_from_Y(Y0)
        cp = U_.cp(m.cp_mass(Tsi, Y0)).astype(np.float64)
        Rmix = U_.cp(m.R / m.W_mean(Y0)).astype(np.float64)
        hm = m.h_molar(Tsi)
        mu = U_.mu(m.mix_viscosity(Tsi, Xs)).astype(np.float64)
        kap = U_.thermal_conductivity(m.mix_conductivity(Tsi, Xs)).astype(np.float64)
        theta = U_.energy_mass(m.R * Tsi / m.W_mean(Y0)).astype(np.float64)
        vel = self._vel()[:, c]
        p = self.p[c]
        om1 = 2 * p * cp / (p * cp + 2 * kap)
        om = 2 * p / (p + 2 * mu)
        self.tau_warnings += int(np.sum((1 / om <= 0.5) | (1 / om1 <= 0.5)))
        rho_a = self.rho_a[:, c]
        ua = self.u_a[:, :, c]
        Uint = self.rhoe[c] / rho
        for a in range(self.ns):
            Uint = Uint - 0.5 * rho_a[a] / rho * (ua[0, a] ** 2 + ua[1, a] ** 2 + ua[2, a] ** 2)
        nb = self._neighbours(c)
        qdiff = np.zeros((3, nc))
        qcorr = np.zeros((3, nc))
        for a in range(self.ns):
            h_a = U_.energy_mass(hm[:, a] / m.W[a]).astype(np.float64)
            for d in range(3):
                qdiff[d] += om1 / (-om + om1) * h_a * rho_a[a] * (ua[d, a] - vel[d])
                Yn = self.rho_a[a, nb[d, 0]] / self.rho[nb[d, 0]]
                Yp = self.rho_a[a, nb[d, 1]] / self.rho[nb[d, 1]]
                dY = fd_fuw(Yn, rho_a[a] / rho, Yp, ua[d, a] - vel[d], self.type[nb[d, 0]], self.type[nb[d, 1]])
                qcorr[d] += 0.5 * (2.0 - om1) / (-om + om1) * 1 * p * h_a * dY
        dQ = np.zeros((3, nc))
        for d in range(3):
            def Rn(idx):
                with np.errstate(invalid="ignore", divide="ignore"):
                    Yn = self.rho_a[:, idx] / self.rho[idx]
                    r = U_.cp(m.R * np.sum(Yn / m.W[:, None], axis=0) / np.sum(Yn, axis=0)).astype(np.float64)
                return np.where(np.isin(self.type[idx], FLUIDLIKE), r, 0.0)
            vv = self._vel()[d]

            def Q(idx, R):
                return self.rho[idx] * vv[idx] * (1.0 - 3.0 * R * self.T[idx]) - self.rho[idx] * (vv[idx] ** 3)
            dQ[d] = fd_central(Q(nb[d, 0], Rn(nb[d, 0])), Q(c, Rmix), Q(nb[d, 1], Rn(nb[d, 1])),
                               self.type[nb[d, 0]], self.type[nb[d, 1]])
        P = self.ptens[:, :, c]
        dq = np.zeros((3, nc))
        for b in range(3):
            acc = np.zeros(nc)
            for a in range(3):
                eqP = (p if a == b else 0.0) + rho * vel[a] * vel[b]
                acc = acc + vel[a] * (P[a, b] - eqP)
            dq[b] = acc + qdiff[b] + qcorr[b] + 0.5 * 1 * vel[b] * dQ[b]
        ge = geq(rho, Uint, theta, vel)
        gs = ge + gstr_shift(dq)
        g = self.g[:, c]
        self.gpc[:, c] = g + om1 * (ge - g) + (om - om1) * (ge - gs)

    # ---------------------------------------------------------- Streaming
    def _dir_slip(self, l, i, j, k):
        """LBM::dirSlip (lbm-bc.cpp:1352-1399): the direction mirrored at the
        first free-slip neighbour along x, y, z, and its source cell wrapped
        into the interior."""
        nx, ny, nz = self.dims
        a, b, e = int(CX[l]), int(CY[l]), int(CZ[l])
        ip, jp, kp = i - a, j - b, k - e
        if self.type[self._idx(i - a, j, k)] == TYPE_FS:
            a, ip = -a, i
        elif self.type[self._idx(i, j - b, k)] == TYPE_FS:
            b, jp = -b, j
        elif self.type[self._idx(i, j, k - e)] == TYPE_FS:
            e, kp = -e, k
        ip = (ip - 1 + (nx - 2)) % (nx - 2) + 1
        jp = (jp - 1 + (ny - 2)) % (ny - 2) + 1
        kp = (kp - 1 + (nz - 2)) % (nz - 2) + 1
        lp = int(np.nonzero((CX == a) & (CY == b) & (CZ == e))[0][0])
        return lp, int(self._idx(ip, jp, kp))

    def stream(self):
        nx, ny, nz = self.dims
        c = np.nonzero(self.type == TYPE_F)[0]
        i, j, k = self._ijk(c)
        for l in range(27):
            inb, jnb, knb = i - CX[l], j - CY[l], k - CZ[l]
            tn = self.type[self._idx(inb, jnb, knb)]
            # bounce-back (S: f only, A: f and g)
            sel = (tn == TYPE_S) | (tn == TYPE_A)
            self.f[:, l, c[sel]] = self.fpc[:, OPP[l], c[sel]]
            selA = tn == TYPE_A
            self.g[l, c[selA]] = self.gpc[OPP[l], c[selA]]
            # free-slip walls (lbm-streaming.cpp:34-49): the population of the
            # mirrored direction from the mirrored source (dirSlip,
            # lbm-bc.cpp:1352-1399), species and energy alike
            for cc, ci, cj, ck in zip(c[tn == TYPE_FS], i[tn == TYPE_FS], j[tn == TYPE_FS], k[tn == TYPE_FS]):
                lp, src = self._dir_slip(l, int(ci), int(cj), int(ck))
                self.f[:, l, cc] = self.fpc[:, lp, src]
                self.g[l, cc] = self.gpc[lp, src]
            per = ~np.isin(tn, (TYPE_S, TYPE_A, TYPE_FS, TYPE_O, TYPE_O_C, TYPE_I, TYPE_I_C))
            ii = np.where(inb == 0, nx - 2, np.where(inb == nx - 1, 1, inb))
            jj = np.where(jnb == 0, ny - 2, np.where(jnb == ny - 1, 1, jnb))
            kk = np.where(knb == 0, nz - 2, np.where(knb == nz - 1, 1, knb))
            src = self._idx(ii, jj, kk)
            self.f[:, l, c[per]] = self.fpc[:, l, src[per]]
            self.g[l, c[per]] = self.gpc[l, src[per]]

    # ------------------------------------------------------------ moments
    def moments(self):
        U_, m = self.units, self.mech
        c = np.nonzero(self.type == TYPE_F)[0]
        nc = len(c)
        # Cantera state of the previous moments: old T, old rho, old Y
        Y_old = self._Y_state(c)
        T_old_si = U_.si_temp(self.T[c]).astype(np.float64)
        rho_old_si = U_.si_rho(self.rho[c]).astype(np.float64)
        g = self.g[:, c]
        self.rhoe[c] = g.sum(axis=0)
        self.eflux[:, c] = C3 @ g
        fa = self.f[:, :, c]
        ra = fa.sum(axis=1)  # (ns, nc)
        mom = np.einsum("dl,alc->dac", C3.astype(np.float64), fa)  # (3, ns, nc)
        P = np.zeros((3, 3, nc))
        for p_ in range(3):
            for q in range(p_, 3):
                P[p_, q] = np.einsum("l,alc->c", (C3[p_] * C3[q]).astype(np.float64), fa)
        P[1, 0], P[2, 0], P[2, 1] = P[0, 1], P[0, 2], P[1, 2]
        self.ptens[:, :, c] = P
        keep = ra > SPECIES_MIN
        self.rho_a[:, c] = np.where(keep, np.maximum(ra, SPECIES_MIN), SPECIES_MIN)
        self.u_a[:, :, c] = np.where(keep[None], mom / np.where(keep, ra, 1.0)[None], 0.0)
        rho = ra.sum(axis=0)
        self.rho[c] = rho
        vel = mom.sum(axis=1) / rho
        self.u[c], self.v[c], self.w[c] = vel
        # Maxwell-Stefan: A u = (rho u)_a with the old state's D_ab, p and X
        p_old_si = m.pressure(T_old_si, rho_old_si, Y_old)
        Dab = U_.nu(m.binary_diff(T_old_si, p_old_si)).astype(np.float64)
        X = self.X[:, c]
        p = self.p[c]
        A = np.zeros((nc, self.ns, self.ns))
        for a in range(self.ns):
            for b in range(self.ns):
                if a != b:
                    A[:, a, b] = -1.0 * 1 / 2.0 * p * X[a] * X[b] / Dab[:, a, b]
                else:
                    r = self.rho_a[a, c]
                    A[:, a, a] = np.where(r == 0.0, 1.0, r)
        for a in range(self.ns):
            for b in range(self.ns):
                if b != a:
                    A[:, a, a] -= A[:, a, b]
        L = np.linalg.cholesky(A)
        for d in range(3):
            y = np.linalg.solve(L, mom[d].T[..., None])
            sol = np.linalg.solve(np.swapaxes(L, 1, 2), y)[..., 0]
            self.u_a[d, :, c] = sol  # (nc, ns): advanced indices d and c are split by a slice
        # temperature from the internal energy at the new composition
        Y = self._Y_state(c)
        Uint = self.rhoe[c] / rho
        ua = self.u_a[:, :, c]
        for a in range(self.ns):
            Uint = Uint - 0.5 * self.rho_a[a, c] / rho * (ua[0, a] ** 2 + ua[1, a] ** 2 + ua[2, a] ** 2)
        u_si = U_.si_energy_mass(Uint).astype(np.float64)
        rho_si = U_.si_rho(rho).astype(np.float64)
        T_si = m.T_from_UV(u_si, Y, T_old_si)
        self.X[:, c] = m.X_from_Y(Y).T
        self.T[c] = U_.temp(T_si).astype(np.float64)
        self.p[c] = U_.p(m.pressure(T_si, rho_si, Y)).astype(np.float64)

    # --------------------------------------------------------- SMOOTHING
    def smoothing(self, alpha=0.1):
        """LBM::Smoothing (lbm-smoothing.cpp:95-173): every species population
        of each TYPE_F and TYPE_I cell moves by alpha times its discrete
        Laplacian, fd_laplace (FD.cpp:40-50) with fd_central_2der (FD.cpp:
        25-37): per axis (r - 2c + l) / dx^2 (dx = 1), 0 when either neighbour
        is TYPE_A, TYPE_FS, TYPE_S or TYPE_P.  An axis without both
        neighbours in the array (a cell on the array's edge) takes 0 for
        their values and TYPE_F (0) for their types, so its term is -2c.
        The reference updates f in place while its loop reads neighbours, an
        order-dependent (and, under OpenMP, racy) result; this restatement is
        the race-free (Jacobi) form: every Laplacian reads the populations as
        they were before the call."""
        nx, ny, nz = self.dims
        typ = self.type.reshape(self.dims)
        f = self.f.reshape((self.ns, 27) + self.dims)
        out = f.copy()
        zero_types = np.isin(typ, (TYPE_A, TYPE_FS, TYPE_S, TYPE_P))
        lap = np.zeros_like(f)
        for ax, m in enumerate((nx, ny, nz)):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            mid = [slice(None)] * 3
            lo[ax], mid[ax], hi[ax] = slice(0, m - 2), slice(1, m - 1), slice(2, m)
            d2 = np.zeros_like(f)
            sl = (Ellipsis,) + tuple(mid)
            d2[sl] = (f[(Ellipsis,) + tuple(hi)] - 2.0 * f[sl]) + f[(Ellipsis,) + tuple(lo)]
            dead = np.zeros(self.dims, bool)
            dead[tuple(mid)] = zero_types[tuple(lo)] | zero_types[tuple(hi)]
            d2[..., dead] = 0.0
            # array edges: both neighbour values 0, types TYPE_F
            for e in (0, m - 1):
                edge = [slice(None)] * 3
                edge[ax] = e
                d2[(Ellipsis,) + tuple(edge)] = 0.0 - 2.0 * f[(Ellipsis,) + tuple(edge)] + 0.0
            lap = lap + d2 if ax else d2
        # fd_laplace sums the three axes in the order x, y, z
        cells = (typ == TYPE_F) | (typ == TYPE_I)
        out[..., cells] = f[..., cells] + alpha * lap[..., cells]
        self.f = out.reshape(self.ns, 27, -1)

    def calculate_moment_smoothing(self, alpha=0.1):
        """lbm.hpp:154-157: moments, Smoothing, moments."""
        self.moments()
        self.smoothing(alpha)
        self.moments()

    def step(self, reaction=True):
        self.collide_species(reaction)
        self.collide()
        self.stream()
        self.moments()

    def diag(self):
        c = np.nonzero(self.type == TYPE_F)[0]
        return {"mass": float(self.rho[c].sum()), "energy": float(self.rhoe[c].sum()),
                "species_mass": self.rho_a[:, c].sum(axis=1), "T_max": float(self.T[c].max()),
                "HRR": float(self.HRR[c].sum())}


# --------------------------------------------------------------------- TMS
def _onlyone(a, b, c):
    """onlyOne_1 (math-util.cpp:89-100): exactly one component is +-1."""
    return (abs(a) + abs(b) + abs(c)) == 1


class MultiStateBC(MultiState):
    """MultiState plus the multicomponent TMS_BC: the outflow sweep
    (lbm-bc.cpp:126-340) with the NSCBC target (impose_nscbc.cpp:45-546) for
    TYPE_O_C ghosts and the downstream-state target for TYPE_O ghosts; the
    inflow sweep (lbm-bc.cpp:463-938) with the inline LODI target for TYPE_I_C
    and the ghost-state target for TYPE_I; the isothermal-wall sweep
    (lbm-bc.cpp:941-1109) for TYPE_S."""

    SIGMA_OUT = 0.25  # NSCBC_SIGMA_OUT, defines.hpp:100-102

    def _Y_at(self, c):
        return self.rho_a[:, c] / self.rho[c]

    def _state_DP(self, rho, p, Y):
        """setState_DP at lattice (rho, p) and Y: returns (T_si, Yn)."""
        U_, m = self.units, self.mech
        Yn = m.normalize_Y(Y)
        rho_si = np.float64(U_.si_rho(rho))
        p_si = np.float64(U_.si_p(p))
        T_si = p_si * m.W_mean(Yn) / (m.R * rho_si)
        return T_si, Yn

    def _sound(self, c):
        """Frozen sound speed (lattice) and gamma at the DP state of cell c."""
        U_, m = self.units, self.mech
        T_si, Yn = self._state_DP(self.rho[c], self.p[c], self._Y_at(c))
        cp = m.cp_mass(np.array(T_si), Yn)
        cv = cp - m.R / m.W_mean(Yn)
        gam = cp / cv
        a_si = np.sqrt(gam * m.R * T_si / m.W_mean(Yn))
        return np.float64(U_.u(a_si)), gam, Yn, T_si

    def _vals(self, c):
        """(p, u, v, w, rho, Y[ns]) of cell c."""
        return np.concatenate([[self.p[c], self.u[c], self.v[c], self.w[c], self.rho[c]], self._Y_at(c)])

    def _normal_derivative(self, c, idir, isign):
        s = self._strides()[idir]
        if isign == 1:
            return (-1.5 * self._vals(c) + 2.0 * self._vals(c + s) - 0.5 * self._vals(c + 2 * s)) / 1
        return (1.5 * self._vals(c) - 2.0 * self._vals(c - s) + 0.5 * self._vals(c - 2 * s)) / 1

    def _tangential_derivative(self, c, idir):
        s = self._strides()[idir]
        tm, tp = self.type[c - s], self.type[c + s]
        if tm != TYPE_F:
            return (-1.5 * self._vals(c) + 2.0 * self._vals(c + s) - 0.5 * self._vals(c + 2 * s)) / 1
        # the z branch tests k-1 twice (impose_nscbc.cpp:233,242): its
        # backward stencil is unreachable and the central one is used instead
        if idir != 2 and tp != TYPE_F:
            return (1.5 * self._vals(c) - 2.0 * self._vals(c - s) + 0.5 * self._vals(c - 2 * s)) / 1
        return (self._vals(c + s) - self._vals(c - s)) / (2.0 * 1)

    def _strides(self):
        nx, ny, nz = self.dims
        return (ny * nz, nz, 1)

    def _nscbc(self, c, l):
        """impose_NSCBC for the interface cell c whose TYPE_O_C neighbour is
        c - e_l: returns rho_out, rhoa_out, vel_out, T_out (lattice)."""
        U_, m = self.units, self.mech
        cvec = (CX[l], CY[l], CZ[l])
        idir = int(np.nonzero(cvec)[0][0])
        isign = int(cvec[idir])
        Ns = [d - 2 for d in self.dims]
        D = [np.zeros(5 + self.ns) for _ in range(3)]
        D[idir] = self._normal_derivative(c, idir, isign)
        others = [d for d in range(3) if d != idir]
        for d in others:
            if Ns[d] > 3:
                D[d] = self._tangential_derivative(c, d)
        Tt = np.zeros(5)
        a, gam, Yn, T_si = self._sound(c)
        rho, p = self.rho[c], self.p[c]
        vel = np.array([self.u[c], self.v[c], self.w[c]])
        if any(Ns[d] > 3 for d in others):
            inv_rho = 1.0 / rho
            dp = [D[d][0] for d in range(3)]
            du = [D[d][1:4] for d in range(3)]  # du[d][comp]
            drho = [D[d][4] for d in range(3)]
            n, t1, t2 = idir, others[0], others[1]
            # Yoo & Im transverse terms (impose_nscbc.cpp:264-335), normal n
            T1 = vel[t1] * (dp[t1] - rho * a * du[t1][n]) + vel[t2] * (dp[t2] - rho * a * du[t2][n]) + \
                gam * p * (du[t1][t1] + du[t2][t2])
            T5 = vel[t1] * (dp[t1] + rho * a * du[t1][n]) + vel[t2] * (dp[t2] + rho * a * du[t2][n]) + \
                gam * p * (du[t1][t1] + du[t2][t2])
            Tt = np.array([T1, 0.0, 0.0, 0.0, T5])
        Dn = D[idir]
        dpn, dvel, drhon, dY = Dn[0], Dn[1:4], Dn[4], Dn[5:]
        un = vel[idir]
        dun = dvel[idir]
        mach = np.sqrt(vel @ vel) / a
        strides = self._strides()
        cb = c - isign * strides[idir]
        L1 = (un - a) * (dpn - rho * a * dun)
        L5 = (un + a) * (dpn + rho * a * dun)
        Ls = un * (a * a * drhon - dpn)
        Lt = un * dvel
        L6 = un * dY
        if self.type[cb] != TYPE_O_C:
            raise ValueError("only TYPE_O_C boundaries are handled")
        if mach < 1.0:
            Kout = self.SIGMA_OUT * (1.0 - mach * mach) * (a / self.dims[idir])
            if isign == 1:
                L5 = Kout * (p - self.p[cb]) - (1.0 - mach) * Tt[4]
            else:
                L1 = Kout * (p - self.p[cb]) - (1.0 - mach) * Tt[0]
        # update_bc_cells (impose_nscbc.cpp:479-545), dt = 1
        drho = (Ls + 0.5 * (L1 + L5)) / (a * a)
        dvel_t = Lt.copy()
        dvel_t[idir] = (L5 - L1) / (2.0 * a * rho)
        dp_t = 0.5 * (L1 + L5)
        rho_out = rho - 1 * drho
        vel_out = vel - 1 * dvel_t
        p_out = p - 1 * dp_t
        rhoa_out = (self.rho_a[:, c] / rho - 1 * L6) * rho_out
        T_si_out, _ = self._state_DP(rho_out, p_out, rhoa_out / rho_out)
        return rho_out, rhoa_out, vel_out, np.float64(U_.temp(T_si_out))

    def _feq_geq(self, rho, rhoa, vel, vela, T):
        """calculate_feq_geq (lbm-edf.cpp:168-196): equilibria at (T, rho, Y)
        with species velocities vela."""
        U_, m = self.units, self.mech
        Y = m.normalize_Y(rhoa / rho)
        T_si = np.float64(U_.si_temp(T))
        Uint = np.float64(U_.energy_mass(m.u_mass(np.array(T_si), Y)))
        theta = np.float64(U_.energy_mass(m.R * T_si / m.W_mean(Y)))
        fa = np.zeros((self.ns, 27))
        for a in range(self.ns):
            th_a = np.float64(U_.energy_mass(m.R * T_si / m.W[a]))
            fa[a] = feq(np.array([rhoa[a]]), vela[a][:, None], np.array([th_a]), np.zeros((3, 1)))[:, 0]
        g = geq(np.array([rho]), np.array([Uint]), np.array([theta]), vel[:, None])[:, 0]
        return fa, g

    def _calculate_temp(self, U, rho, rhoa, T0_si):
        """calculate_temp (lbm-moment.cpp:540-553): T from (U, rho, Y) by UV."""
        U_, m = self.units, self.mech
        Y = m.normalize_Y(rhoa / rho)
        T_si = m.T_from_UV(np.float64(U_.si_energy_mass(U)), Y[None], np.array([T0_si]))[0]
        return np.float64(U_.temp(T_si))

    def interface_cells(self, kinds=(TYPE_O, TYPE_O_C)):
        """(cell, l_interface) of every fluid cell whose first ghost direction
        of the given kinds (outflow by default, lbm-bc.cpp:141-155; inflow
        with (TYPE_I, TYPE_I_C), lbm-bc.cpp:479-494) is axis-aligned."""
        out = []
        for c in np.nonzero(self.type == TYPE_F)[0]:
            i, j, k = self._ijk(c)
            for l in range(27):
                tn = self.type[self._idx(i - CX[l], j - CY[l], k - CZ[l])]
                tp = self.type[self._idx(i + CX[l], j + CY[l], k + CZ[l])]
                if tn in kinds and tp == TYPE_F:
                    if _onlyone(CX[l], CY[l], CZ[l]):
                        out.append((int(c), l))
                    break
        return out

    def _inflow_lodi(self, c, l):
        """TYPE_I_C inflow target (lbm-bc.cpp:519-677): characteristic
        relaxation with sigma = 5 towards the ghost's state, first-order
        one-sided normal derivatives; returns rho_in, rhoa_in, vel_in, T_in
        after the setState_TP density update (lbm-bc.cpp:672-680)."""
        U_, m = self.units, self.mech
        sigma = 5.0
        cvec = (CX[l], CY[l], CZ[l])
        idir = int(np.nonzero(cvec)[0][0])
        isign = int(cvec[idir])
        st = self._strides()[idir]
        c1, c2 = c + isign * st, c + 2 * isign * st
        cb = c - isign * st
        Nd = self.dims[idir]  # ghost-inclusive length (Q7)
        a, gam, Yn, T_si = self._sound(c)
        rho, p = self.rho[c], self.p[c]
        vel = np.array([self.u[c], self.v[c], self.w[c]])
        velb = np.array([self.u[cb], self.v[cb], self.w[cb]])

        def fd_uw(s0, s1):  # FD.cpp:64-74 with vel = the interface direction
            return (s0 - s1) / 1 if isign < 0 else (-s0 + s1) / 1

        comp = [self.u, self.v, self.w][idir]
        du = fd_uw(comp[c], comp[c1])
        dp = fd_uw(self.p[c], self.p[c1])
        dY = np.array([fd_uw(self.rho_a[b, c] / self.rho[c], self.rho_a[b, c1] / self.rho[c1]) for b in range(self.ns)])
        un = vel[idir]
        # tangential components as the reference pairs them: x -> (v, w), y -> (w, u), z -> (u, v)
        t3, t4 = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[idir]
        Rmix = np.float64(U_.cp(m.R / m.W_mean(Yn)))
        L1 = (un - a) * (dp - rho * a * du)
        L2 = sigma * a / Nd * rho * Rmix * (self.T[c] - self.T[cb])
        L3 = sigma * a / Nd * (vel[t3] - velb[t3])
        L4 = sigma * a / Nd * (vel[t4] - velb[t4])
        L5 = (un + a) * (dp + rho * a * du)
        K = sigma * a * a / Nd * rho * (1.0 - (un / a) * (un / a))
        # the z branch tests cy (lbm-bc.cpp:652, Q4): always the "right" form there
        left = (CY[l] > 0) if idir == 2 else (isign > 0)
        if left:
            L5 = K * (un - velb[idir])
        else:
            L1 = K * (un - velb[idir])
        d1 = 1.0 / (a * a) * (L2 + 0.5 * (L5 + L1))
        d2 = 1.0 / (2.0 * rho * a) * (L5 - L1)
        d5 = 0.5 * (L5 + L1)
        rho_in = rho - 1 * d1
        vel_in = vel.copy()
        vel_in[idir] = vel[idir] - 1 * d2
        vel_in[t3] = vel[t3] - 1 * L3
        vel_in[t4] = vel[t4] - 1 * L4
        p_in = p - 1 * d5
        rhoa_in = (self.rho_a[:, c] / rho - 1 * un * dY) * rho_in
        T_si_in, _ = self._state_DP(rho_in, p_in, rhoa_in / rho_in)
        T_in = np.float64(U_.temp(T_si_in))
        # setState_TP(T_in, p_in) at Y = rhoa_in / rho_in: the density follows
        Y = m.normalize_Y(rhoa_in / rho_in)
        T_si = np.float64(U_.si_temp(T_in))
        p_si = np.float64(U_.si_p(p_in))
        rho_in = np.float64(U_.rho(p_si * m.W_mean(Y) / (m.R * T_si)))
        return rho_in, Y * rho_in, vel_in, T_in

    def _moment_point(self, c):
        """calculate_moment_point (lbm-moment.cpp:379-537) of cell c, writing
        its fields.  Unlike calculate_moment: no SPECIES_MIN clamp; the binary
        diffusivities at (T of the cell's fields, the new density and
        composition); the Maxwell-Stefan matrix from the fields' p and X (not
        yet updated); the solve by partial-pivoting LU."""
        U_, m = self.units, self.mech
        g = self.g[:, c]
        self.rhoe[c] = g.sum()
        self.eflux[:, c] = C3 @ g
        fa = self.f[:, :, c]  # (ns, 27)
        ra = fa.sum(axis=1)
        mom = fa @ C3.T.astype(np.float64)  # (ns, 3)
        P = np.zeros((3, 3))
        for p_ in range(3):
            for q in range(p_, 3):
                P[p_, q] = np.sum(fa * (C3[p_] * C3[q]).astype(np.float64)[None, :])
        P[1, 0], P[2, 0], P[2, 1] = P[0, 1], P[0, 2], P[1, 2]
        self.ptens[:, :, c] = P
        rho = ra.sum()
        self.rho_a[:, c] = ra
        self.rho[c] = rho
        self.u[c], self.v[c], self.w[c] = mom.sum(axis=0) / rho
        Y = m.normalize_Y(ra / rho)
        T_old_si = np.float64(U_.si_temp(self.T[c]))
        rho_si = np.float64(U_.si_rho(rho))
        p_si = m.pressure(T_old_si, rho_si, Y)
        Dab = U_.nu(m.binary_diff(T_old_si, p_si)).astype(np.float64)
        A = np.zeros((self.ns, self.ns))
        for a in range(self.ns):
            for b in range(self.ns):
                if b != a:
                    A[a, b] = -1.0 * 1 / 2.0 * self.p[c] * self.X[a, c] * self.X[b, c] / Dab[a, b]
                else:
                    A[a, a] = 1.0 if ra[a] == 0.0 else ra[a]
        for a in range(self.ns):
            for b in range(self.ns):
                if b != a:
                    A[a, a] -= A[a, b]
        ua = np.linalg.solve(A, mom)  # (ns, 3)
        self.u_a[:, :, c] = ua.T
        ke = 0.0
        for a in range(self.ns):
            ke += 0.5 * ra[a] * (ua[a] @ ua[a])
        Uint = self.rhoe[c] / rho - ke / rho
        T_si = m.T_from_UV(np.float64(U_.si_energy_mass(Uint)), Y[None], np.array([T_old_si]))[0]
        self.X[:, c] = m.X_from_Y(Y)
        self.T[c] = np.float64(U_.temp(T_si))
        self.p[c] = np.float64(U_.p(m.pressure(T_si, rho_si, Y)))

    def _reconstruct(self, c, kinds, fa_b, g_b, vel_tgt=None, T_tgt=None, inflow=False):
        """The local part shared by the sweeps: unknown directions (upstream
        ghost of one of `kinds`) take the boundary set (fa_b, g_b) in the local
        moments; T_loc by calculate_temp; then unknown 2 f_b - f_loc and known
        f_b + f - f_loc, or, for the inflow sweep, f_tgt + f_b - f_loc and
        f_tgt + f - f_loc with f_tgt the equilibrium of the local density and
        composition at (vel_tgt, T_tgt)."""
        U_ = self.units
        i, j, k = self._ijk(c)
        unknown = np.array([self.type[self._idx(i - CX[q], j - CY[q], k - CZ[q])] in kinds for q in range(27)])
        fsrc = np.where(unknown[None, :], fa_b, self.f[:, :, c])
        gsrc = np.where(unknown, g_b, self.g[:, c])
        rhoa_loc = fsrc.sum(axis=1)
        mom = fsrc @ C3.T.astype(np.float64)
        rhoe_loc = gsrc.sum()
        rho_loc = rhoa_loc.sum()
        vel_loc = mom.sum(axis=0) / rho_loc if rho_loc > 0 else np.zeros(3)
        vela_loc = np.where(rhoa_loc[:, None] > 0, mom / np.where(rhoa_loc > 0, rhoa_loc, 1.0)[:, None], 0.0)
        Uint = rhoe_loc / rho_loc
        for a in range(self.ns):
            Uint = Uint - 0.5 * rhoa_loc[a] / rho_loc * (vela_loc[a] @ vela_loc[a])
        T_loc = self._calculate_temp(Uint, rho_loc, rhoa_loc, np.float64(U_.si_temp(self.T[c])))
        fa_loc, g_loc = self._feq_geq(rho_loc, rhoa_loc, vel_loc, vela_loc, T_loc)
        if inflow:
            fa_t, g_t = self._feq_geq(rho_loc, rhoa_loc, vel_tgt, np.repeat(vel_tgt[None], self.ns, 0), T_tgt)
            fnew = np.where(unknown[None, :], fa_t + fa_b - fa_loc, fa_t + self.f[:, :, c] - fa_loc)
            gnew = np.where(unknown, g_t + g_b - g_loc, g_t + self.g[:, c] - g_loc)
        else:
            fnew = np.where(unknown[None, :], 2 * fa_b - fa_loc, fa_b + self.f[:, :, c] - fa_loc)
            gnew = np.where(unknown, 2 * g_b - g_loc, g_b + self.g[:, c] - g_loc)
        self.f[:, :, c] = fnew
        self.g[:, c] = gnew

    def _outflow_plain(self, c, l):
        """TYPE_O outflow target (lbm-bc.cpp:168-198): moments of the cell
        downstream (recomputed by calculate_moment_point), the pressure of
        the ghost, T from (rho, p, Y) by setState_DP."""
        U_ = self.units
        d = self._idx(*(np.array(self._ijk(c)) + np.array([CX[l], CY[l], CZ[l]])))
        g_ = self._idx(*(np.array(self._ijk(c)) - np.array([CX[l], CY[l], CZ[l]])))
        self._moment_point(d)
        vel_out = np.array([self.u[d], self.v[d], self.w[d]])
        rho_out = self.rho[d]
        rhoa_out = self.rho_a[:, d].copy()
        T_si, _ = self._state_DP(rho_out, self.p[g_], rhoa_out / rho_out)
        return rho_out, rhoa_out, vel_out, np.float64(U_.temp(T_si))

    def _inflow_plain(self, c, l):
        """TYPE_I inflow target (lbm-bc.cpp:506-518, 672-680): velocity,
        temperature and composition of the ghost, the pressure of the cell
        downstream after calculate_moment_point, the density from
        setState_TP."""
        U_, m = self.units, self.mech
        d = self._idx(*(np.array(self._ijk(c)) + np.array([CX[l], CY[l], CZ[l]])))
        g_ = self._idx(*(np.array(self._ijk(c)) - np.array([CX[l], CY[l], CZ[l]])))
        self._moment_point(d)
        vel_in = np.array([self.u[g_], self.v[g_], self.w[g_]])
        T_in = self.T[g_]
        Y = m.normalize_Y(self.rho_a[:, g_] / self.rho[g_])
        T_si = np.float64(U_.si_temp(T_in))
        p_si = np.float64(U_.si_p(self.p[d]))
        rho_in = np.float64(U_.rho(p_si * m.W_mean(Y) / (m.R * T_si)))
        return rho_in, Y * rho_in, vel_in, T_in

    def _tms_wall(self):
        """TMS_BC isothermal-wall sweep, MULTICOMP (lbm-bc.cpp:941-1109):
        every link with a TYPE_S cell upstream and fluid downstream is an
        interface link; the target takes the cell's own species densities,
        and velocity (mixture and per species) and temperature interpolated at
        q = 1/2 between the cell downstream and the wall, averaged over the
        links; unknown directions 2 f_tgt - f_loc, known f_tgt + f - f_loc."""
        q = 0.5
        for c in np.nonzero(self.type == TYPE_F)[0]:
            i, j, k = self._ijk(c)
            links = []
            for l in range(27):
                up = self._idx(i - CX[l], j - CY[l], k - CZ[l])
                dn = self._idx(i + CX[l], j + CY[l], k + CZ[l])
                if self.type[up] == TYPE_S and self.type[dn] == TYPE_F:
                    links.append((up, dn))
            if not links:
                continue
            rhoa_bb = self.f[:, :, c].sum(axis=1)
            rho_bb = 0.0
            for a in range(self.ns):
                rho_bb += rhoa_bb[a]
            vel = np.zeros(3)
            vela = np.zeros((self.ns, 3))
            T = 0.0
            for up, dn in links:
                mu_up = np.array([self.u[up], self.v[up], self.w[up]])
                mu_dn = np.array([self.u[dn], self.v[dn], self.w[dn]])
                vel += (q * mu_dn + mu_up) / (1.0 + q)
                T += (q * self.T[dn] + self.T[up]) / (1.0 + q)
                vela += (q * self.u_a[:, :, dn].T + mu_up[None, :]) / (1.0 + q)
            nd = len(links)
            vel = 1.0 / nd * vel
            T = 1.0 / nd * T
            vela = 1.0 / nd * vela
            fa_t, g_t = self._feq_geq(rho_bb, rhoa_bb, vel, vela, T)
            self._reconstruct(c, (TYPE_S,), fa_t, g_t)

    def tms_bc(self):
        """TMS_BC: the outflow sweep (lbm-bc.cpp:126-340; TYPE_O_C ghosts with
        the NSCBC target, TYPE_O ghosts with the downstream state), then the
        inflow sweep and the isothermal-wall sweep."""
        m, U_ = self.mech, self.units
        updates = []
        for c, l in self.interface_cells():
            i, j, k = self._ijk(c)
            if self.type[self._idx(i - CX[l], j - CY[l], k - CZ[l])] == TYPE_O:
                # the downstream cell's moments are rewritten in place as the
                # sweep reaches this cell, as calculate_moment_point does
                rho_out, rhoa_out, vel_out, T_out = self._outflow_plain(c, l)
            else:
                rho_out, rhoa_out, vel_out, T_out = self._nscbc(c, l)
            fa_out, g_out = self._feq_geq(rho_out, rhoa_out, vel_out, np.repeat(vel_out[None], self.ns, 0), T_out)
            unknown = np.array([self.type[self._idx(i - CX[q], j - CY[q], k - CZ[q])] in (TYPE_O, TYPE_O_C)
                                for q in range(27)])
            fsrc = np.where(unknown[None, :], fa_out, self.f[:, :, c])
            gsrc = np.where(unknown, g_out, self.g[:, c])
            rhoa_loc = fsrc.sum(axis=1)
            mom = fsrc @ C3.T.astype(np.float64)  # (ns, 3)
            rhoe_loc = gsrc.sum()
            rho_loc = rhoa_loc.sum()
            vel_loc = mom.sum(axis=0) / rho_loc if rho_loc > 0 else np.zeros(3)
            vela_loc = np.where(rhoa_loc[:, None] > 0, mom / np.where(rhoa_loc > 0, rhoa_loc, 1.0)[:, None], 0.0)
            Uint = rhoe_loc / rho_loc
            for a in range(self.ns):
                Uint = Uint - 0.5 * rhoa_loc[a] / rho_loc * (vela_loc[a] @ vela_loc[a])
            T_loc = self._calculate_temp(Uint, rho_loc, rhoa_loc, np.float64(U_.si_temp(self.T[c])))
            fa_loc, g_loc = self._feq_geq(rho_loc, rhoa_loc, vel_loc, vela_loc, T_loc)
            fnew = np.where(unknown[None, :], 2 * fa_out - fa_loc, fa_out + self.f[:, :, c] - fa_loc)
            gnew = np.where(unknown, 2 * g_out - g_loc, g_out + self.g[:, c] - g_loc)
            updates.append((c, fnew, gnew))
        for c, fnew, gnew in updates:
            self.f[:, :, c] = fnew
            self.g[:, c] = gnew
        self._tms_inflow()
        self._tms_wall()

    def _tms_inflow(self):
        """TMS_BC inflow sweep (lbm-bc.cpp:463-938) for TYPE_I_C ghosts: the
        unknown directions get f_tgt + f_in - f_loc, the known ones
        f_tgt + f - f_loc, with f_in the equilibrium of the LODI inflow state,
        f_loc that of the local moments (inflow populations substituted) and
        f_tgt that of the local density and composition at the inflow
        velocity and temperature."""
        U_ = self.units
        updates = []
        for c, l in self.interface_cells((TYPE_I, TYPE_I_C)):
            i, j, k = self._ijk(c)
            if self.type[self._idx(i - CX[l], j - CY[l], k - CZ[l])] == TYPE_I:
                rho_in, rhoa_in, vel_in, T_in = self._inflow_plain(c, l)
            else:
                rho_in, rhoa_in, vel_in, T_in = self._inflow_lodi(c, l)
            fa_in, g_in = self._feq_geq(rho_in, rhoa_in, vel_in, np.repeat(vel_in[None], self.ns, 0), T_in)
            unknown = np.array([self.type[self._idx(i - CX[q], j - CY[q], k - CZ[q])] in (TYPE_I, TYPE_I_C)
                                for q in range(27)])
            fsrc = np.where(unknown[None, :], fa_in, self.f[:, :, c])
            gsrc = np.where(unknown, g_in, self.g[:, c])
            rhoa_loc = fsrc.sum(axis=1)
            mom = fsrc @ C3.T.astype(np.float64)
            rhoe_loc = gsrc.sum()
            rho_loc = rhoa_loc.sum()
            vel_loc = mom.sum(axis=0) / rho_loc if rho_loc > 0 else np.zeros(3)
            vela_loc = np.where(rhoa_loc[:, None] > 0, mom / np.where(rhoa_loc > 0, rhoa_loc, 1.0)[:, None], 0.0)
            Uint = rhoe_loc / rho_loc
            for a in range(self.ns):
                Uint = Uint - 0.5 * rhoa_loc[a] / rho_loc * (vela_loc[a] @ vela_loc[a])
            T_loc = self._calculate_temp(Uint, rho_loc, rhoa_loc, np.float64(U_.si_temp(self.T[c])))
            fa_loc, g_loc = self._feq_geq(rho_loc, rhoa_loc, vel_loc, vela_loc, T_loc)
            fa_tgt, g_tgt = self._feq_geq(rho_loc, rhoa_loc, vel_in, np.repeat(vel_in[None], self.ns, 0), T_in)
            fnew = np.where(unknown[None, :], fa_tgt + fa_in - fa_loc, fa_tgt + self.f[:, :, c] - fa_loc)
            gnew = np.where(unknown, g_tgt + g_in - g_loc, g_tgt + self.g[:, c] - g_loc)
            updates.append((c, fnew, gnew))
        for c, fnew, gnew in updates:
            self.f[:, :, c] = fnew
            self.g[:, c] = gnew

    def step(self, reaction=True):
        self.collide_species(reaction)
        self.collide()
        self.stream()
        self.tms_bc()
        self.moments()
