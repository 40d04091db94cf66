// Host-side LBM class with the reference's interface (lbm.hpp:104-274) whose
// state lives on the GPU behind the C ABI of lbm_amd.h.  A set-up written for
// the reference keeps its shape: construct, set types and primitive fields,
// run(nstep, tout); LBM::loop keeps the reference's five step calls
// (lbm-main.cpp:149-196).  Per-cell access lb.mixture[i][j][k].x becomes
// bulk field access (field(LBM_FIELD_U)[idx(i, j, k)]) followed by upload();
// lb.species[a][i][j][k].X of the MULTICOMP build becomes
// species_field(LBM_SP_X)[a * n + idx(i, j, k)].
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_amd.h"

class LBM {
public:
    LBM(int Nx, int Ny, int Nz, double nu, int mode = LBM_MODE_FUSED, int precision = LBM_FP64, int device = 0)
    {
        cfg_ = {};
        cfg_.nx = Nx;
        cfg_.ny = Ny;
        cfg_.nz = Nz;
        cfg_.n_species = 0;
        cfg_.precision = precision;
        cfg_.mode = mode;
        cfg_.device = device;
        cfg_.nu = nu;
        cfg_.gas_const = 1.0;  // LBM defaults, lbm.hpp:117-121
        cfg_.gamma = 1.4;
        cfg_.prtl = 0.5;
        cfg_.ref_rho = 1.0;
        cfg_.ref_T = 1.0;
        this->Nx = Nx + 2;
        this->Ny = Ny + 2;
        this->Nz = Nz + 2;
        const size_t n = (size_t)this->Nx * this->Ny * this->Nz;
        type.assign(n, 0);
        fields_.assign(LBM_FIELD_COUNT, std::vector<double>(n, 0.0));
        fields_[LBM_FIELD_RHO].assign(n, 1.0);
        fields_[LBM_FIELD_P].assign(n, 1.0 / 3.0);
        fields_[LBM_FIELD_T].assign(n, 1.0 / 3.0);
    }
    // LBM(Nx, Ny, Nz, species) of the MULTICOMP build (lbm-main.cpp:38-118)
    // with units.set_m_kg_s(dx_si, dt_si) (units.hpp:32-44) and the REACTION
    // switch; the species are the library mechanism's, in its order
    // (lbm_mechanism, lbm_species_name).  Mixture defaults as LBM() (lbm.hpp:70-83).
    LBM(int Nx, int Ny, int Nz, double dx_si, double dt_si, bool reaction, int mode = LBM_MODE_FUSED,
        int precision = LBM_FP64, int device = 0)
        : LBM(Nx, Ny, Nz, 0.0, mode, precision, device)
    {
        int ns = 0;
        lbm_mechanism(&ns);
        cfg_.n_species = ns;
        cfg_.dx_si = dx_si;
        cfg_.dt_si = dt_si;
        cfg_.reaction = reaction ? 1 : 0;
        for (int a = 0; a < ns; ++a) species_names.push_back(lbm_species_name(a));
        const size_t n = (size_t)this->Nx * this->Ny * this->Nz;
        species_.assign(5, std::vector<double>(n * ns, 0.0));
    }
    ~LBM()
    {
        if (ctx_) lbm_destroy(ctx_);
    }
    LBM(const LBM&) = delete;
    LBM& operator=(const LBM&) = delete;

    // ---- the reference's getters and setters (lbm.hpp:200-271)
    int get_Nx() const { return Nx; }
    int get_Ny() const { return Ny; }
    int get_Nz() const { return Nz; }
    int get_NX() const { return Nx - 2; }
    int get_NY() const { return Ny - 2; }
    int get_NZ() const { return Nz - 2; }
    double get_nu() const { return cfg_.nu; }
    double get_gasconst() const { return cfg_.gas_const; }
    double get_gamma() const { return cfg_.gamma; }
    double get_prtl() const { return cfg_.prtl; }
    double get_soundspeed(double temp) const { return std::sqrt(cfg_.gamma * cfg_.gas_const * temp); }
    void set_nu(double v) { cfg_.nu = v; }
    void set_gasconst(double v) { cfg_.gas_const = v; }
    void set_gamma(double v) { cfg_.gamma = v; }
    void set_prtl(double v) { cfg_.prtl = v; }
    // rest state of the FP32 shifted populations (no reference counterpart:
    // the reference stores FP64 values)
    void set_reference_state(double rho, double T)
    {
        cfg_.ref_rho = rho;
        cfg_.ref_T = T;
    }
    // the reference's #define CONJUGATE (defines.hpp:62-63) as a run-time switch;
    // takes effect at the next upload()
    void set_conjugate(bool on) { conjugate_ = on; }
    int get_step() const { return step; }

    size_t idx(int i, int j, int k) const { return ((size_t)i * Ny + j) * Nz + k; }
    size_t ncell() const { return (size_t)Nx * Ny * Nz; }
    std::vector<double>& field(int f) { return fields_[f]; }
    // species planes [NS][n] of enum lbm_species_field (MULTICOMP only)
    std::vector<double>& species_field(int f) { return species_.at(f); }
    int n_species() const { return cfg_.n_species; }
    std::vector<std::string> species_names;
    std::vector<uint8_t> type;

    // ---- host <-> device
    void upload()
    {
        ensure_ctx();
        check(lbm_set_conjugate(ctx_, conjugate_ ? 1 : 0), "lbm_set_conjugate");
        check(lbm_set_types(ctx_, type.data()), "lbm_set_types");
        const int prim[] = {LBM_FIELD_RHO, LBM_FIELD_U, LBM_FIELD_V, LBM_FIELD_W, LBM_FIELD_P, LBM_FIELD_T};
        for (int f : prim) check(lbm_set_field(ctx_, f, fields_[f].data()), "lbm_set_field");
        if (cfg_.n_species) check(lbm_set_species_field(ctx_, LBM_SP_X, species_[LBM_SP_X].data()), "lbm_set_species_field");
    }
    void download()
    {
        for (int f = 0; f < LBM_FIELD_COUNT; ++f) check(lbm_get_field(ctx_, f, fields_[f].data()), "lbm_get_field");
        for (int f = 0; f < (cfg_.n_species ? 5 : 0); ++f)
            check(lbm_get_species_field(ctx_, f, species_[f].data()), "lbm_get_species_field");
    }
    // a population set of enum lbm_pop, [27][n] (with species [NS][27][n])
    std::vector<double> populations(int which)
    {
        const bool sp = cfg_.n_species && (which == LBM_POP_F || which == LBM_POP_FPC);
        std::vector<double> out(ncell() * 27 * (sp ? cfg_.n_species : 1));
        check(lbm_get_populations(ctx_, which, out.data()), "lbm_get_populations");
        return out;
    }
    lbm_diag diagnostics()
    {
        lbm_diag d{};
        check(lbm_diagnostics(ctx_, &d), "lbm_diagnostics");
        return d;
    }
    lbm_ctx* context() { return ctx_; }

    // ---- the step calls of LBM::loop
    void Init() { check(lbm_init(ctx_), "Init"); }
    void Collide_Species() { check(lbm_collide_species(ctx_), "Collide_Species"); }
    void Collide() { check(lbm_collide(ctx_), "Collide"); }
    void Streaming() { check(lbm_stream(ctx_), "Streaming"); }
    void TMS_BC() { check(lbm_tms_bc(ctx_), "TMS_BC"); }
    void calculate_moment() { check(lbm_moments(ctx_), "calculate_moment"); }

    // OutputVTK (output.cpp:9-210) and write_restart / read_restart
    // (restart_file.cpp:274-440) from the device state, same file names and formats
    void OutputVTK(int nout, int flags = 0)
    {
        char name[64];
        snprintf(name, sizeof(name), "./field%06d.vtr", nout);
        check(lbm_write_vtk(ctx_, name, flags), "OutputVTK");
    }
    void write_restart(int s)
    {
        char name[64];
        snprintf(name, sizeof(name), "./restart%06d.dat", s);
        check(lbm_write_restart(ctx_, name, s), "write_restart");
    }
    void read_restart(const std::string& file)
    {
        ensure_ctx();
        check(lbm_read_restart(ctx_, file.c_str(), &step), "read_restart");
    }

    // LBM::run (lbm-main.cpp:126-147): Init, output at t = 0, loop.
    void run(int nstep, int tout, const std::function<void(int, LBM&)>& output)
    {
        upload();
        step = 0;
        Init();
        if (output) output(step, *this);
        loop(nstep, tout, output);
    }

    // LBM::loop (lbm-main.cpp:149-196).
    void loop(int nstep, int tout, const std::function<void(int, LBM&)>& output)
    {
        for (int s = step + 1; s <= nstep; ++s) {
            Collide_Species();
            Collide();
            Streaming();
            TMS_BC();
            calculate_moment();
            step = s;
            if (output && tout > 0 && s % tout == 0) output(s, *this);
        }
    }

private:
    void ensure_ctx()
    {
        if (ctx_) return;
        if (int rc = lbm_create(&cfg_, &ctx_))
            throw std::runtime_error("lbm_create failed (" + std::to_string(rc) + "): " + lbm_create_error());
    }
    void check(int rc, const char* what)
    {
        if (rc) throw std::runtime_error(std::string(what) + ": " + lbm_last_error(ctx_));
    }
    lbm_config cfg_;
    lbm_ctx* ctx_ = nullptr;
    bool conjugate_ = false;
    std::vector<std::vector<double>> fields_;
    std::vector<std::vector<double>> species_;
    int Nx, Ny, Nz;
    int step = 0;
};
