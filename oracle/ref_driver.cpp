// Reference driver: test infrastructure only (oracle/_ref).
//
// Links the reference's own single-component translation units, compiled
// unmodified from /root/reference/src by oracle/build_ref.sh, and drives them
// through the public LBM class (lbm.hpp:104-274) exactly the way LBM::loop
// does (lbm-main.cpp:153-193): Collide(); Streaming(); TMS_BC();
// calculate_moment().  Its outputs are golden vectors for the oracle and the
// GPU path; it never ships in the product and never runs the product.
//
// Modes
//   fn  <seed> <count> <out.bin>      per-call vectors of calculate_feq,
//                                     both calculate_geq overloads and
//                                     calculate_gstr (lbm.hpp:276, lbm-edf.cpp:38,100,140)
//   tgv <N> <nsteps> <snap csv> <outdir> <U0> <nu> <R> <gamma> <Pr> <T0> <p0>
//                                     periodic 3-D Taylor-Green box set up the
//                                     way setup.cpp:94-148 does (ghost-inclusive
//                                     index, TYPE_P ghosts), Init(), then steps
//   time <N> <nsteps> ...same params  time nsteps of the step loop (CPU baseline)
//   channel <nx> <ny> <nz> <nsteps> <snaps> <outdir> <U> <nu> <R> <gamma> <Pr> <T0> <p0> <Tin> <Tw> <pout>
//                                     inflow/outflow/wall channel (every single-
//                                     component TMS_BC branch); a negative snap
//                                     -s dumps step s before TMS_BC
//   io <nx> <ny> <nz> <nsteps> <outdir> ...channel params
//                                     the channel's VTK and restart files written
//                                     by the reference's OutputVTK/write_restart
#include "lbm.hpp"
#include "output.hpp"
#include "restart_file.hpp"
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <algorithm>
#include <string>
#include <vector>
#include <chrono>
#include <omp.h>

static uint64_t g_state;
static double rnd()  // SplitMix64 -> [0,1)
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
static double uni(double a, double b) { return a + (b - a) * rnd(); }

static int mode_fn(int argc, char** argv)
{
    if (argc < 5) return 2;
    g_state = strtoull(argv[2], nullptr, 10);
    int count = atoi(argv[3]);
    FILE* fo = fopen(argv[4], "wb");
    if (!fo) return 3;
    LBM lb(1, 1, 1, 0.01);
    for (int s = 0; s < count; ++s) {
        double rec[30];
        int l = s % 27;
        double rho = uni(0.5, 2.0);
        double v[3] = {uni(-0.3, 0.3), uni(-0.3, 0.3), uni(-0.3, 0.3)};
        double theta = uni(0.05, 0.5);
        double corr[3] = {uni(-0.01, 0.01), uni(-0.01, 0.01), uni(-0.01, 0.01)};
        double U = uni(0.2, 3.0);
        double dq[3] = {uni(-0.05, 0.05), uni(-0.05, 0.05), uni(-0.05, 0.05)};
        double rhoe = uni(0.5, 3.0);
        double q[3] = {uni(-0.3, 0.3), uni(-0.3, 0.3), uni(-0.3, 0.3)};
        double R[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) {
                R[a][b] = (a == b) ? uni(0.3, 2.0) : uni(-0.2, 0.2);
                R[b][a] = R[a][b];
            }
        double feq = lb.calculate_feq(l, rho, v, theta, corr);
        double geq = lb.calculate_geq(l, rho, U, theta, v);
        double gstr = lb.calculate_gstr(l, geq, dq);
        double geqm = lb.calculate_geq(l, rhoe, q, R, theta);
        int n = 0;
        rec[n++] = l; rec[n++] = rho; rec[n++] = v[0]; rec[n++] = v[1]; rec[n++] = v[2];
        rec[n++] = theta; rec[n++] = corr[0]; rec[n++] = corr[1]; rec[n++] = corr[2];
        rec[n++] = U; rec[n++] = dq[0]; rec[n++] = dq[1]; rec[n++] = dq[2];
        rec[n++] = rhoe; rec[n++] = q[0]; rec[n++] = q[1]; rec[n++] = q[2];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) rec[n++] = R[a][b];
        rec[n++] = feq; rec[n++] = geq; rec[n++] = gstr; rec[n++] = geqm;
        fwrite(rec, sizeof(double), 30, fo);
    }
    fclose(fo);
    return 0;
}

struct TgvParams { int N; double U0, nu, R, gamma, Pr, T0, p0; };

// Taylor-Green set-up of setup.cpp:94-148 with the box parameters passed in:
// ghosts become TYPE_P, interior cells get the analytic fields on the
// ghost-inclusive index.
static void setup_tgv(LBM& lb, const TgvParams& P)
{
    lb.set_gasconst(P.R);
    lb.set_gamma(P.gamma);
    lb.set_prtl(P.Pr);
    int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
    double a = lb.get_NX(), b = lb.get_NY(), c = lb.get_NZ();
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                MIXTURE& m = lb.mixture[i][j][k];
                if (i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1)
                    m.type = TYPE_P;
                if (m.type == TYPE_F) {
                    m.u = P.U0 * sin(2.0 * M_PI / a * i) * cos(2.0 * M_PI / b * j) * cos(2.0 * M_PI / c * k);
                    m.v = -P.U0 * cos(2.0 * M_PI / a * i) * sin(2.0 * M_PI / b * j) * cos(2.0 * M_PI / c * k);
                    m.w = 0.0;
                    m.p = P.p0 + 1.0 * P.U0 * P.U0 / 16.0 * (cos(4.0 * M_PI / a * i) + cos(4.0 * M_PI / b * j)) *
                                     (cos(4.0 * M_PI / c * k) + 2.0);
                    m.temp = P.T0;
                }
            }
}

// KE = sum rho|u|^2 and enstrophy with periodic central differences, as
// OutputKeEns (output.cpp:212-276) defines them; plus sum rho and sum rho*e.
static void diagnostics(LBM& lb, double out[4])
{
    int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
    double ke = 0, ens = 0, mass = 0, en = 0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                MIXTURE& m = lb.mixture[i][j][k];
                if (m.type != TYPE_F) continue;
                ke += m.rho * (m.u * m.u + m.v * m.v + m.w * m.w);
                int ia = i + 1, ib = i - 1, ja = j + 1, jb = j - 1, ka = k + 1, kb = k - 1;
                if (ib < 1) ib = Nx - 2; else if (ia > Nx - 2) ia = 1;
                if (jb < 1) jb = Ny - 2; else if (ja > Ny - 2) ja = 1;
                if (kb < 1) kb = Nz - 2; else if (ka > Nz - 2) ka = 1;
                double uy = (lb.mixture[i][ja][k].u - lb.mixture[i][jb][k].u) / 2;
                double uz = (lb.mixture[i][j][ka].u - lb.mixture[i][j][kb].u) / 2;
                double vx = (lb.mixture[ia][j][k].v - lb.mixture[ib][j][k].v) / 2;
                double vz = (lb.mixture[i][j][ka].v - lb.mixture[i][j][kb].v) / 2;
                double wx = (lb.mixture[ia][j][k].w - lb.mixture[ib][j][k].w) / 2;
                double wy = (lb.mixture[i][ja][k].w - lb.mixture[i][jb][k].w) / 2;
                double vort = (wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy);
                ens += m.rho * vort;
                mass += m.rho;
                en += m.rhoe;
            }
    out[0] = ke; out[1] = ens; out[2] = mass; out[3] = en;
}

// Interior-cell snapshot: 7 moments, p_tensor[9], energy_flux[3], f, g, fpc, gpc.
static void dump_snapshot(LBM& lb, const std::string& path)
{
    FILE* fo = fopen(path.c_str(), "wb");
    if (!fo) { perror(path.c_str()); exit(3); }
    int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
    for (int i = 1; i < Nx - 1; ++i)
        for (int j = 1; j < Ny - 1; ++j)
            for (int k = 1; k < Nz - 1; ++k) {
                MIXTURE& m = lb.mixture[i][j][k];
                double rec[7 + 9 + 3 + 4 * 27];
                int n = 0;
                rec[n++] = m.rho; rec[n++] = m.u; rec[n++] = m.v; rec[n++] = m.w;
                rec[n++] = m.p; rec[n++] = m.temp; rec[n++] = m.rhoe;
                for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) rec[n++] = m.p_tensor[a][b];
                for (int a = 0; a < 3; ++a) rec[n++] = m.energy_flux[a];
                for (int l = 0; l < 27; ++l) rec[n++] = m.f[l];
                for (int l = 0; l < 27; ++l) rec[n++] = m.g[l];
                for (int l = 0; l < 27; ++l) rec[n++] = m.fpc[l];
                for (int l = 0; l < 27; ++l) rec[n++] = m.gpc[l];
                fwrite(rec, sizeof(double), n, fo);
            }
    fclose(fo);
}

static int parse_tgv(int argc, char** argv, int base, TgvParams& P)
{
    if (argc < base + 7) return 2;
    P.U0 = atof(argv[base + 0]); P.nu = atof(argv[base + 1]); P.R = atof(argv[base + 2]);
    P.gamma = atof(argv[base + 3]); P.Pr = atof(argv[base + 4]); P.T0 = atof(argv[base + 5]);
    P.p0 = atof(argv[base + 6]);
    return 0;
}

static int mode_tgv(int argc, char** argv)
{
    TgvParams P;
    if (argc < 6 || parse_tgv(argc, argv, 6, P)) return 2;
    P.N = atoi(argv[2]);
    int nsteps = atoi(argv[3]);
    std::vector<int> snaps;
    {
        std::string s(argv[4]);
        size_t pos = 0;
        while (pos < s.size()) {
            size_t e = s.find(',', pos);
            if (e == std::string::npos) e = s.size();
            snaps.push_back(atoi(s.substr(pos, e - pos).c_str()));
            pos = e + 1;
        }
    }
    std::string outdir(argv[5]);
    LBM lb(P.N, P.N, P.N, P.nu);
    setup_tgv(lb, P);
    lb.Init();
    FILE* fs = fopen((outdir + "/series.bin").c_str(), "wb");
    double d[4];
    diagnostics(lb, d);
    fwrite(d, sizeof(double), 4, fs);
    for (int s : snaps) if (s == 0) dump_snapshot(lb, outdir + "/snap_0.bin");
    for (int step = 1; step <= nsteps; ++step) {
        lb.Collide();
        lb.Streaming();
        lb.TMS_BC();
        lb.calculate_moment();
        diagnostics(lb, d);
        fwrite(d, sizeof(double), 4, fs);
        for (int s : snaps)
            if (s == step) dump_snapshot(lb, outdir + "/snap_" + std::to_string(step) + ".bin");
    }
    fclose(fs);
    return 0;
}

// Channel with every single-component TMS_BC branch: TYPE_I inflow ghost at
// i = 0 (u_in, T_in), TYPE_O outflow ghost at i = Nx-1 (p_out), TYPE_S
// isothermal walls at j = 0 (at rest) and j = Ny-1 (moving with -0.3 u_in),
// both at T_wall, periodic in z; the x faces are typed last.  The inflow
// ghost carries a small transverse velocity 0.05 u_in.  Interior: u_in with
// small transverse waves in v and w, T0, p0.
struct ChannelParams { int nx, ny, nz; double U, nu, R, gamma, Pr, T0, p0, Tin, Tw, pout; };

static void setup_channel(LBM& lb, const ChannelParams& P)
{
    lb.set_gasconst(P.R);
    lb.set_gamma(P.gamma);
    lb.set_prtl(P.Pr);
    int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                MIXTURE& m = lb.mixture[i][j][k];
                if (k == 0 || k == Nz - 1) m.type = TYPE_P;
                if (j == 0 || j == Ny - 1) m.type = TYPE_S;
                if (i == 0) m.type = TYPE_I;
                if (i == Nx - 1) m.type = TYPE_O;
                m.u = 0.0; m.v = 0.0; m.w = 0.0;
                m.temp = P.T0;
                m.p = P.p0;
                if (m.type == TYPE_F) {
                    m.u = P.U;
                    m.v = 0.1 * P.U * sin(2.0 * M_PI * i / (Nx - 2)) * sin(M_PI * j / (Ny - 1));
                    m.w = 0.05 * P.U * cos(2.0 * M_PI * (k - 1) / (Nz - 2)) * sin(M_PI * j / (Ny - 1));
                } else if (m.type == TYPE_I) {
                    m.u = P.U;
                    m.v = 0.05 * P.U;
                    m.temp = P.Tin;
                } else if (m.type == TYPE_O) {
                    m.p = P.pout;
                } else if (m.type == TYPE_S) {
                    m.temp = P.Tw;
                    if (j == Ny - 1) m.u = -0.3 * P.U;  // moving upper wall
                }
                m.rho = m.p / (P.R * m.temp);
            }
}

// channel <nx> <ny> <nz> <nsteps> <snap csv> <outdir> <U> <nu> <R> <gamma> <Pr> <T0> <p0> <Tin> <Tw> <pout>
static int mode_channel(int argc, char** argv)
{
    if (argc < 17) return 2;
    ChannelParams P;
    P.nx = atoi(argv[2]); P.ny = atoi(argv[3]); P.nz = atoi(argv[4]);
    int nsteps = atoi(argv[5]);
    std::vector<int> snaps;
    {
        std::string s(argv[6]);
        size_t pos = 0;
        while (pos < s.size()) {
            size_t e = s.find(',', pos);
            if (e == std::string::npos) e = s.size();
            snaps.push_back(atoi(s.substr(pos, e - pos).c_str()));
            pos = e + 1;
        }
    }
    std::string outdir(argv[7]);
    P.U = atof(argv[8]); P.nu = atof(argv[9]); P.R = atof(argv[10]); P.gamma = atof(argv[11]);
    P.Pr = atof(argv[12]); P.T0 = atof(argv[13]); P.p0 = atof(argv[14]); P.Tin = atof(argv[15]);
    P.Tw = atof(argv[16]); P.pout = argc > 17 ? atof(argv[17]) : P.p0;
    LBM lb(P.nx, P.ny, P.nz, P.nu);
    setup_channel(lb, P);
    lb.Init();
    for (int s : snaps) if (s == 0) dump_snapshot(lb, outdir + "/snap_0.bin");
    for (int step = 1; step <= nsteps; ++step) {
        lb.Collide();
        lb.Streaming();
        if (std::find(snaps.begin(), snaps.end(), -step) != snaps.end())  // state before TMS_BC
            dump_snapshot(lb, outdir + "/pre_" + std::to_string(step) + ".bin");
        lb.TMS_BC();
        lb.calculate_moment();
        for (int s : snaps)
            if (s == step) dump_snapshot(lb, outdir + "/snap_" + std::to_string(step) + ".bin");
    }
    return 0;
}

// io <nx> <ny> <nz> <nsteps> <outdir> <U> <nu> <R> <gamma> <Pr> <T0> <p0> <Tin> <Tw> <pout>
// The channel after nsteps steps, written by the reference's own writers
// (OutputVTK: field%06d.vtr; write_restart: restart%06d.dat) into outdir.
static int mode_io(int argc, char** argv)
{
    if (argc < 16) return 2;
    ChannelParams P;
    P.nx = atoi(argv[2]); P.ny = atoi(argv[3]); P.nz = atoi(argv[4]);
    int nsteps = atoi(argv[5]);
    std::string outdir(argv[6]);
    P.U = atof(argv[7]); P.nu = atof(argv[8]); P.R = atof(argv[9]); P.gamma = atof(argv[10]);
    P.Pr = atof(argv[11]); P.T0 = atof(argv[12]); P.p0 = atof(argv[13]); P.Tin = atof(argv[14]);
    P.Tw = atof(argv[15]); P.pout = argc > 16 ? atof(argv[16]) : P.p0;
    LBM lb(P.nx, P.ny, P.nz, P.nu);
    setup_channel(lb, P);
    lb.Init();
    for (int step = 1; step <= nsteps; ++step) {
        lb.Collide();
        lb.Streaming();
        lb.TMS_BC();
        lb.calculate_moment();
    }
    if (chdir(outdir.c_str()) != 0) { perror(outdir.c_str()); return 3; }
    int nout = nsteps, step = nsteps;
    OutputVTK(nout, &lb);
    write_restart(step, &lb);
    dump_snapshot(lb, "snap.bin");
    return 0;
}

static int mode_time(int argc, char** argv)
{
    TgvParams P;
    if (argc < 4 || parse_tgv(argc, argv, 4, P)) return 2;
    P.N = atoi(argv[2]);
    int nsteps = atoi(argv[3]);
    LBM lb(P.N, P.N, P.N, P.nu);
    setup_tgv(lb, P);
    lb.Init();
    auto t0 = std::chrono::steady_clock::now();
    for (int step = 1; step <= nsteps; ++step) {
        lb.Collide();
        lb.Streaming();
        lb.TMS_BC();
        lb.calculate_moment();
    }
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    double cells = (double)P.N * P.N * P.N;
    printf("{\"seconds\": %.6f, \"steps\": %d, \"cells\": %.0f, \"mlups\": %.6f, \"threads\": %d}\n", sec, nsteps,
           cells, cells * nsteps / sec / 1e6, omp_get_max_threads());
    return 0;
}

// Conjugate heat-transfer box, for the build compiled with -DCONJUGATE:
// periodic ghosts on every face; solid slabs of `ns` cells at the bottom
// (T_hot) and the top (T_cold) of the y range, gas between them at T0 and p0
// moving with u = U plus small transverse waves.  The two slabs meet through
// the periodic y wrap, so a solid cell has only solid, periodic or gas
// neighbours (the reference's solid loops index i +- 1 without bounds).
struct ConjParams { int nx, ny, nz, ns; double U, nu, R, gamma, Pr, T0, p0, Thot, Tcold; };

static void setup_conj(LBM& lb, const ConjParams& P)
{
    lb.set_gasconst(P.R);
    lb.set_gamma(P.gamma);
    lb.set_prtl(P.Pr);
    int Nx = lb.get_Nx(), Ny = lb.get_Ny(), Nz = lb.get_Nz();
    double H = Ny - 2 - 2 * P.ns;  // gas layer height
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                MIXTURE& m = lb.mixture[i][j][k];
                bool ghost = i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1;
                m.u = 0.0; m.v = 0.0; m.w = 0.0;
                m.temp = P.T0;
                m.p = P.p0;
                if (ghost) {
                    m.type = TYPE_P;
                } else if (j <= P.ns) {
                    m.type = TYPE_S;
                    m.temp = P.Thot;
                } else if (j >= Ny - 1 - P.ns) {
                    m.type = TYPE_S;
                    m.temp = P.Tcold;
                } else {
                    double y = (j - P.ns - 0.5) / H;
                    m.u = P.U;
                    m.v = 0.1 * P.U * sin(2.0 * M_PI * i / (Nx - 2)) * sin(M_PI * y);
                    m.w = 0.05 * P.U * cos(2.0 * M_PI * (k - 1) / (Nz - 2)) * sin(M_PI * y);
                }
                m.rho = m.p / (P.R * m.temp);
            }
}

// conj <nx> <ny> <nz> <ns> <nsteps> <snap csv> <outdir> <U> <nu> <R> <gamma> <Pr> <T0> <p0> <Thot> <Tcold>
static int mode_conj(int argc, char** argv)
{
    if (argc < 18) return 2;
    ConjParams P;
    P.nx = atoi(argv[2]); P.ny = atoi(argv[3]); P.nz = atoi(argv[4]); P.ns = atoi(argv[5]);
    int nsteps = atoi(argv[6]);
    std::vector<int> snaps;
    {
        std::string s(argv[7]);
        size_t pos = 0;
        while (pos < s.size()) {
            size_t e = s.find(',', pos);
            if (e == std::string::npos) e = s.size();
            snaps.push_back(atoi(s.substr(pos, e - pos).c_str()));
            pos = e + 1;
        }
    }
    std::string outdir(argv[8]);
    P.U = atof(argv[9]); P.nu = atof(argv[10]); P.R = atof(argv[11]); P.gamma = atof(argv[12]);
    P.Pr = atof(argv[13]); P.T0 = atof(argv[14]); P.p0 = atof(argv[15]); P.Thot = atof(argv[16]);
    P.Tcold = atof(argv[17]);
    LBM lb(P.nx, P.ny, P.nz, P.nu);
    setup_conj(lb, P);
    lb.Init();
    for (int s : snaps) if (s == 0) dump_snapshot(lb, outdir + "/snap_0.bin");
    for (int step = 1; step <= nsteps; ++step) {
        lb.Collide();
        if (std::find(snaps.begin(), snaps.end(), -step) != snaps.end())  // post-collision state
            dump_snapshot(lb, outdir + "/post_" + std::to_string(step) + ".bin");
        lb.Streaming();
        lb.TMS_BC();
        lb.calculate_moment();
        for (int s : snaps)
            if (s == step) dump_snapshot(lb, outdir + "/snap_" + std::to_string(step) + ".bin");
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 2) { fprintf(stderr, "usage: ref_driver fn|tgv|time ...\n"); return 2; }
    if (!strcmp(argv[1], "fn")) return mode_fn(argc, argv);
    if (!strcmp(argv[1], "tgv")) return mode_tgv(argc, argv);
    if (!strcmp(argv[1], "time")) return mode_time(argc, argv);
    if (!strcmp(argv[1], "channel")) return mode_channel(argc, argv);
    if (!strcmp(argv[1], "io")) return mode_io(argc, argv);
    if (!strcmp(argv[1], "conj")) return mode_conj(argc, argv);
    return 2;
}
