// Case driver in the reference's main -> run -> loop shape (main.cpp:9-25,
// lbm-main.cpp:126-196) on the GPU path, for set-ups prepared elsewhere:
//   lbm_run CASE_FILE OUT_FILE
// CASE_FILE (little-endian): "LBMCASE1"; int32 nx ny nz n_species mode
// precision reaction nstep; float64 nu R gamma Pr ref_rho ref_T dx_si dt_si;
// uint8 type[n]; float64 rho u v w p T [n each]; float64 X[n_species][n].
// n = (nx + 2)(ny + 2)(nz + 2), cell index (i (ny + 2) + j)(nz + 2) + k.
// OUT_FILE: "LBMOUT01"; int32 n_species; float64 fields[LBM_FIELD_COUNT][n],
// then with species: X, rho, u, v, w [n_species][n]; then the populations f
// ([n_species or 1][27][n]) and g ([27][n]) after nstep steps.
// The step loop's wall time is printed.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_host.hpp"

namespace {

struct Reader {
    FILE* fp;
    template <typename T>
    void get(T* p, size_t count)
    {
        if (fread(p, sizeof(T), count, fp) != count) throw std::runtime_error("truncated case file");
    }
};

void put(FILE* fp, const void* p, size_t bytes)
{
    if (fwrite(p, 1, bytes, fp) != bytes) throw std::runtime_error("write error");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: lbm_run CASE_FILE OUT_FILE\n");
        return 2;
    }
    try {
        FILE* in = std::fopen(argv[1], "rb");
        if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
        Reader r{in};
        char magic[8];
        r.get(magic, 8);
        if (std::memcmp(magic, "LBMCASE1", 8)) throw std::runtime_error("not a case file");
        int32_t h[8];
        r.get(h, 8);
        double g[8];
        r.get(g, 8);
        const int nx = h[0], ny = h[1], nz = h[2], ns = h[3], mode = h[4], precision = h[5], nstep = h[7];
        const bool reaction = h[6] != 0;
        std::unique_ptr<LBM> lb;
        if (ns > 0) {
            lb.reset(new LBM(nx, ny, nz, g[6], g[7], reaction, mode, precision));
            if (lb->n_species() != ns) throw std::runtime_error("species count differs from the library's mechanism");
        } else {
            lb.reset(new LBM(nx, ny, nz, g[0], mode, precision));
            lb->set_gasconst(g[1]);
            lb->set_gamma(g[2]);
            lb->set_prtl(g[3]);
        }
        lb->set_reference_state(g[4], g[5]);
        const size_t n = lb->ncell();
        r.get(lb->type.data(), n);
        const int prim[6] = {LBM_FIELD_RHO, LBM_FIELD_U, LBM_FIELD_V, LBM_FIELD_W, LBM_FIELD_P, LBM_FIELD_T};
        for (int f : prim) r.get(lb->field(f).data(), n);
        if (ns > 0) r.get(lb->species_field(LBM_SP_X).data(), n * ns);
        std::fclose(in);

        const auto t0 = std::chrono::steady_clock::now();
        lb->run(nstep, 0, nullptr);
        lbm_sync(lb->context());
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        lb->download();

        FILE* out = std::fopen(argv[2], "wb");
        if (!out) throw std::runtime_error(std::string("cannot open ") + argv[2]);
        put(out, "LBMOUT01", 8);
        const int32_t nsp = ns;
        put(out, &nsp, 4);
        for (int f = 0; f < LBM_FIELD_COUNT; ++f) put(out, lb->field(f).data(), n * 8);
        for (int f = 0; f < (ns ? 5 : 0); ++f) put(out, lb->species_field(f).data(), n * ns * 8);
        for (int which : {LBM_POP_F, LBM_POP_G}) {
            const std::vector<double> p = lb->populations(which);
            put(out, p.data(), p.size() * 8);
        }
        std::fclose(out);
        std::printf("lbm_run: %dx%dx%d, %d species, %d steps, %.3f s\n", nx, ny, nz, ns, nstep, sec);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "lbm_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
