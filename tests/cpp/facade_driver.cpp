// Drives the BioFVM-compatible façade (biofvm_amd/facade/BioFVM.h) the way
// the reference's example programs drive BioFVM (examples/convergence_test5.cpp:
// 152-241): build the microenvironment through its API, add Dirichlet nodes,
// bulk callbacks and agents, then loop bulk -> cells -> diffusion.
//
//   facade_driver <case.bin> <out.bin> [<grad.bin> <snapshot.mat>]
//   facade_driver <case.bin> <out.bin> --time W [--bulk-every-step]
//
// case.bin (little endian): int32 nx ny nz S steps phases n_dir n_agents
// has_bulk precision; float64 dx dt; D[S] lambda[S]; rho0[nvox*S];
// int32 dir_idx[n_dir]; float64 dir_val[n_dir*S]; if has_bulk: supply,
// target, uptake [nvox*S each]; agents: pos[3n] volume[n] secretion[nS]
// saturation[nS] uptake[nS].  out.bin: float64 rho[nvox*S].  With the two
// optional paths, the final gradient vectors (compute_all_gradient_vectors,
// float64 [nvox][S][3]) and a write_to_matlab snapshot are written too.
// Halfway through, every density is read through density_vector() (which
// synchronises the host mirror) and written back unchanged, so the result
// also covers the mirror's download/upload path.
//
// With --time W the driver runs W steps untimed (the first LOD call picks its
// launch shapes), then the case's steps timed with a steady clock until the
// device has finished them (no density is read in between), and prints one
// JSON line: seconds, ms per step, host seconds spent on bulk callbacks, and
// how often the handle built its Thomas tables.  --bulk-every-step evaluates
// the bulk callbacks at every step, as the reference does, instead of once.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "BioFVM.h"
#include "biofvm_amd.h"

using namespace BioFVM;

namespace {
int S_ = 0;
std::vector<double> g_sup, g_tgt, g_upt;
void sup_fn(Microenvironment*, int v, std::vector<double>* d) { for (int s = 0; s < S_; ++s) (*d)[s] = g_sup[(size_t)v * S_ + s]; }
void tgt_fn(Microenvironment*, int v, std::vector<double>* d) { for (int s = 0; s < S_; ++s) (*d)[s] = g_tgt[(size_t)v * S_ + s]; }
void upt_fn(Microenvironment*, int v, std::vector<double>* d) { for (int s = 0; s < S_; ++s) (*d)[s] = g_upt[(size_t)v * S_ + s]; }

template <typename T>
void rd(FILE* f, T* p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}
}  // namespace

int main(int argc, char** argv)
{
    int warmup = -1;  // --time W: timing mode
    bool every_step = false;
    {
        std::vector<char*> pos;
        for (int i = 0; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--time" && i + 1 < argc) warmup = std::atoi(argv[++i]);
            else if (a == "--bulk-every-step") every_step = true;
            else pos.push_back(argv[i]);
        }
        argc = (int)pos.size();
        for (int i = 0; i < argc; ++i) argv[i] = pos[i];
    }
    if (argc != 3 && argc != 5 && argc != 7) {
        fprintf(stderr, "usage: facade_driver case.bin out.bin [grad.bin snapshot.mat [mcds_base flags]]\n"
                        "  flags: m = mesh as MAT, x = data as CSV lines, - = neither\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[10];
    rd(f, hdr, 10);
    const int nx = hdr[0], ny = hdr[1], nz = hdr[2], S = hdr[3], phases = hdr[5],
              n_dir = hdr[6], n_agents = hdr[7], has_bulk = hdr[8], precision = hdr[9];
    int steps = hdr[4];
    double dd[2];
    rd(f, dd, 2);
    const double dx = dd[0], dt = dd[1];
    const size_t nvox = (size_t)nx * ny * nz;
    std::vector<double> D(S), lam(S), rho((size_t)nvox * S), dval((size_t)n_dir * S);
    std::vector<int> didx(n_dir);
    rd(f, D.data(), S);
    rd(f, lam.data(), S);
    rd(f, rho.data(), rho.size());
    rd(f, didx.data(), n_dir);
    rd(f, dval.data(), dval.size());
    S_ = S;
    if (has_bulk) {
        g_sup.resize(nvox * S); g_tgt.resize(nvox * S); g_upt.resize(nvox * S);
        rd(f, g_sup.data(), g_sup.size());
        rd(f, g_tgt.data(), g_tgt.size());
        rd(f, g_upt.data(), g_upt.size());
    }
    std::vector<double> pos(3 * (size_t)n_agents), vol(n_agents), sec((size_t)n_agents * S),
        sat((size_t)n_agents * S), upt((size_t)n_agents * S);
    rd(f, pos.data(), pos.size());
    rd(f, vol.data(), vol.size());
    rd(f, sec.data(), sec.size());
    rd(f, sat.data(), sat.size());
    rd(f, upt.data(), upt.size());
    fclose(f);

    Microenvironment M;
    M.device_precision = precision;
    M.bulk_rates_every_step = every_step;
    M.name = "microenvironment";
    M.set_density(0, "s0", "dimensionless");
    for (int s = 1; s < S; ++s) M.add_density("s" + std::to_string(s), "mmHg");
    M.resize_space_uniform(0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx);
    M.spatial_units = "micron";
    M.mesh.units = "micron";
    // phase bit 8: the 2-D LOD solver (a one-layer mesh)
    M.diffusion_decay_solver = (phases & 8) ? diffusion_decay_solver__constant_coefficients_LOD_2D
                                            : diffusion_decay_solver__constant_coefficients_LOD_3D;
    for (int s = 0; s < S; ++s) {
        M.diffusion_coefficients[s] = D[s];
        M.decay_rates[s] = lam[s];
    }
    for (size_t v = 0; v < nvox; ++v)
        for (int s = 0; s < S; ++s) M.density_vector((int)v)[s] = rho[v * S + s];
    for (int d = 0; d < n_dir; ++d) {
        std::vector<double> value(dval.begin() + (size_t)d * S, dval.begin() + (size_t)(d + 1) * S);
        M.add_dirichlet_node(didx[d], value);
    }
    if (has_bulk) {
        M.bulk_supply_rate_function = sup_fn;
        M.bulk_supply_target_densities_function = tgt_fn;
        M.bulk_uptake_rate_function = upt_fn;
    }
    for (int a = 0; a < n_agents; ++a) {
        Basic_Agent* p = create_basic_agent();
        p->register_microenvironment(&M);
        if (!p->assign_position(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2])) {
            fprintf(stderr, "agent %d outside the domain\n", a);
            return 3;
        }
        p->volume = vol[a];
        for (int s = 0; s < S; ++s) {
            (*p->secretion_rates)[s] = sec[(size_t)a * S + s];
            (*p->saturation_densities)[s] = sat[(size_t)a * S + s];
            (*p->uptake_rates)[s] = upt[(size_t)a * S + s];
        }
        p->set_internal_uptake_constants(dt);
    }
    auto step = [&] {
        if ((phases & 1) && has_bulk) M.simulate_bulk_sources_and_sinks(dt);
        if (phases & 2) M.simulate_cell_sources_and_sinks(dt);
        if (phases & 4) M.simulate_diffusion_decay(dt);
    };
    if (warmup >= 0) {
        for (int n = 0; n < warmup; ++n) step();
        M.wait_for_device();
        const double host0 = M.bulk_rates_host_seconds;
        const auto t0 = std::chrono::steady_clock::now();
        for (int n = 0; n < steps; ++n) step();
        M.wait_for_device();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        unsigned builds = 0;
        bfvm_get_coefficient_builds(M.device_handle(), &builds);
        printf("{\"steps\": %d, \"warmup\": %d, \"seconds\": %.6f, \"ms_per_step\": %.6f, "
               "\"bulk_rates_host_seconds\": %.6f, \"bulk_every_step\": %s, \"coefficient_builds\": %u}\n",
               steps, warmup, secs, 1e3 * secs / steps, M.bulk_rates_host_seconds - host0,
               every_step ? "true" : "false", builds);
        fflush(stdout);
        steps = 0;  // the timed steps are done; write the result below
    }
    for (int n = 0; n < steps; ++n) {
        if (n == steps / 2) {
            for (size_t v = 0; v < nvox; ++v) {
                std::vector<double>& d = M.density_vector((int)v);
                for (int s = 0; s < S; ++s) d[s] = d[s] + 0.0;
            }
        }
        step();
    }
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    for (size_t v = 0; v < nvox; ++v) fwrite(M.density_vector((int)v).data(), sizeof(double), S, o);
    fclose(o);
    if (argc == 5) {
        M.compute_all_gradient_vectors();
        FILE* g = fopen(argv[3], "wb");
        if (!g) return 2;
        for (size_t v = 0; v < nvox; ++v)
            for (int s = 0; s < S; ++s) fwrite(M.gradient_vector((int)v)[s].data(), sizeof(double), 3, g);
        fclose(g);
        M.write_to_matlab(argv[4]);
    }
    if (argc == 7) {  // two MultiCellDS snapshots: the document, then its update
        const std::string base = argv[5], flags = argv[6];
        set_save_biovm_mesh_as_matlab(flags.find('m') != std::string::npos);
        set_save_biovm_data_as_matlab(flags.find('x') == std::string::npos);
        default_metadata.current_time = steps * dt;
        default_metadata.current_runtime = 1.25;
        save_microenvironment_to_MultiCellDS_xml_pugi(base + "_a", M);
        default_metadata.current_time += 1.0;
        default_metadata.current_runtime = 2.5;
        save_microenvironment_to_MultiCellDS_xml_pugi(base + "_b", M);
    }
    while (!all_basic_agents.empty()) delete_basic_agent((int)all_basic_agents.size() - 1);
    return 0;
}
