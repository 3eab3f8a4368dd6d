/*
 * TEST INFRASTRUCTURE ONLY — not part of the product.
 *
 * A C-ABI driver around the compiled BioFVM 1.0.4 reference (built from the
 * unmodified sources under the reference tree by oracle/Makefile into
 * oracle/_ref/).  It sets up one Microenvironment exactly as the reference's
 * example programs do, feeds it caller-supplied arrays, runs the reference's
 * own step methods and copies the densities back out.  It is linked only
 * into oracle/_ref/libbfvm_ref.so and is loaded only by tests/, by the
 * fixture generator under tests/golden/ and by bench.py's CPU-baseline leg.
 *
 * Reference call sequence followed (file:line in the reference tree):
 *   Microenvironment ctor + set_density/add_density    BioFVM_microenvironment.cpp:88-139, 325-458
 *   resize_space_uniform                                BioFVM_microenvironment.cpp:287-290
 *   bulk callbacks (S, T, U)                            BioFVM_microenvironment.h:130-132
 *   add_dirichlet_node                                  BioFVM_microenvironment.cpp:146-151
 *   create_basic_agent / register / assign_position     BioFVM_basic_agent.cpp:61-81, 172-181
 *   set_internal_uptake_constants                       BioFVM_basic_agent.cpp:114-148
 *   step order bulk -> cells -> diffusion               examples/convergence_test5.cpp:236-238
 */
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

#include "BioFVM.h"

namespace BioFVM {
void set_default_microenvironment(Microenvironment* M);
}

using namespace BioFVM;

namespace {

// Per-voxel source/sink fields handed to the reference's bulk callbacks.
const double* g_bulk_supply = nullptr;
const double* g_bulk_target = nullptr;
const double* g_bulk_uptake = nullptr;
int g_S = 0;

void copy_field(const double* field, int voxel, std::vector<double>* dst)
{
    for (int s = 0; s < g_S; ++s) (*dst)[s] = field[(size_t)voxel * g_S + s];
}
void supply_cb(Microenvironment*, int voxel, std::vector<double>* dst) { copy_field(g_bulk_supply, voxel, dst); }
void target_cb(Microenvironment*, int voxel, std::vector<double>* dst) { copy_field(g_bulk_target, voxel, dst); }
void uptake_cb(Microenvironment*, int voxel, std::vector<double>* dst) { copy_field(g_bulk_uptake, voxel, dst); }

}  // namespace

extern "C" {

/* Flags for ref_run's `phases`: which of the three step methods run each step. */
enum { REF_BULK = 1, REF_CELLS = 2, REF_DIFFUSION = 4, REF_LOD_2D = 8 };

void ref_destroy(void* h);

/* One reference Microenvironment with its agents, set up once and stepped
 * any number of times (ref_create / ref_steps / ref_read / ref_destroy). */
struct RefSession {
    Microenvironment* M = nullptr;
    std::vector<Basic_Agent*> agents;
    int nvox = 0, S = 0;
};

/*
 * Sets up the reference on an nx*ny*nz grid of S substrates.
 *   rho        nx*ny*nz*S doubles, voxel-major / substrate-innermost
 *   dir_*      n_dir Dirichlet nodes (voxel index, S values each)
 *   bulk_*     per-voxel supply rate S, target T, uptake U (nvox*S each), or
 *              NULL; read by the callbacks at every step, so they must stay
 *              valid until ref_destroy
 *   agent_*    n_agents agents: position (3 each), volume, and S rates each
 *   out_voxel  optional, n_agents ints: the voxel each agent was assigned
 * Returns the session, or NULL with *rc < 0 if the reference rejected the
 * input.  One session at a time (the reference keeps its agents and the
 * default microenvironment in globals).
 */
void* ref_create(int nx, int ny, int nz, int S, double dx, const double* D, const double* lambda,
                 double dt, int phases, const double* rho, int n_dir, const int* dir_idx,
                 const double* dir_val, const double* bulk_supply, const double* bulk_target,
                 const double* bulk_uptake, int n_agents, const double* agent_pos,
                 const double* agent_volume, const double* agent_secretion,
                 const double* agent_saturation, const double* agent_uptake, int* out_voxel, int* rc)
{
    *rc = 0;
    RefSession* R = new RefSession;
    Microenvironment* M = R->M = new Microenvironment;
    set_default_microenvironment(M);
    M->set_density(0, "s0", "dimensionless");
    for (int s = 1; s < S; ++s) M->add_density("s", "dimensionless");
    M->resize_space_uniform(0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx);
    if ((int)M->mesh.x_coordinates.size() != nx || (int)M->mesh.y_coordinates.size() != ny ||
        (int)M->mesh.z_coordinates.size() != nz || M->number_of_densities() != S) {
        *rc = -1;
    }
    M->diffusion_decay_solver = (phases & REF_LOD_2D) ? diffusion_decay_solver__constant_coefficients_LOD_2D
                                                      : diffusion_decay_solver__constant_coefficients_LOD_3D;
    R->nvox = nx * ny * nz;
    R->S = S;
    if (*rc == 0) {
        for (int s = 0; s < S; ++s) {
            M->diffusion_coefficients[s] = D[s];
            M->decay_rates[s] = lambda[s];
        }
        for (int v = 0; v < R->nvox; ++v) {
            std::vector<double>& d = M->density_vector(v);
            for (int s = 0; s < S; ++s) d[s] = rho[(size_t)v * S + s];
        }
        for (int d = 0; d < n_dir; ++d) {
            std::vector<double> value(dir_val + (size_t)d * S, dir_val + (size_t)(d + 1) * S);
            M->add_dirichlet_node(dir_idx[d], value);
        }
        g_S = S;
        g_bulk_supply = bulk_supply;
        g_bulk_target = bulk_target;
        g_bulk_uptake = bulk_uptake;
        if (bulk_supply && bulk_target && bulk_uptake) {
            M->bulk_supply_rate_function = supply_cb;
            M->bulk_supply_target_densities_function = target_cb;
            M->bulk_uptake_rate_function = uptake_cb;
        }
        R->agents.reserve(n_agents);
        for (int a = 0; a < n_agents; ++a) {
            Basic_Agent* p = create_basic_agent();
            R->agents.push_back(p);
            p->register_microenvironment(M);
            if (!p->assign_position(agent_pos[3 * a], agent_pos[3 * a + 1], agent_pos[3 * a + 2])) {
                *rc = -2;  // the reference would leave this agent's voxel uninitialised
                break;
            }
            p->volume = agent_volume[a];
            for (int s = 0; s < S; ++s) {
                (*p->secretion_rates)[s] = agent_secretion[(size_t)a * S + s];
                (*p->saturation_densities)[s] = agent_saturation[(size_t)a * S + s];
                (*p->uptake_rates)[s] = agent_uptake[(size_t)a * S + s];
            }
            p->set_internal_uptake_constants(dt);
            if (out_voxel) out_voxel[a] = p->nearest_voxel_index();
        }
    }
    if (*rc != 0) {
        ref_destroy(R);
        return nullptr;
    }
    return R;
}

/* `steps` steps in the reference's order (bulk -> cells -> diffusion, the
 * phases flagged) on `nthreads` OpenMP threads (0: the runtime's default);
 * *seconds = wall-clock seconds of the step loop only. */
int ref_steps(void* h, double dt, int steps, int phases, int nthreads, double* seconds)
{
    RefSession* R = (RefSession*)h;
    if (!R) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Microenvironment* M = R->M;
    auto t0 = std::chrono::steady_clock::now();
    for (int n = 0; n < steps; ++n) {
        if (phases & REF_BULK) M->simulate_bulk_sources_and_sinks(dt);
        if (phases & REF_CELLS) M->simulate_cell_sources_and_sinks(R->agents, dt);
        if (phases & REF_DIFFUSION) M->simulate_diffusion_decay(dt);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return 0;
}

/* The densities, voxel-major / substrate-innermost. */
void ref_read(void* h, double* rho)
{
    const RefSession* R = (const RefSession*)h;
    for (int v = 0; v < R->nvox; ++v) {
        const std::vector<double>& d = R->M->density_vector(v);
        for (int s = 0; s < R->S; ++s) rho[(size_t)v * R->S + s] = d[s];
    }
}

/* Tears the session down: its agents must not leak into the next one (the
 * reference keeps them in the global all_basic_agents list). */
void ref_destroy(void* h)
{
    RefSession* R = (RefSession*)h;
    if (!R) return;
    for (Basic_Agent* p : R->agents) delete p;
    all_basic_agents.clear();
    set_default_microenvironment(nullptr);
    delete R->M;
    delete R;
    g_bulk_supply = g_bulk_target = g_bulk_uptake = nullptr;
}

/*
 * Runs `steps` steps of the reference on one fresh session (ref_create's
 * arguments; rho in/out) and tears it down.  seconds: optional, wall-clock
 * seconds of the step loop only.  Returns 0 on success, a negative code if
 * the reference rejected the input.
 */
int ref_run(int nx, int ny, int nz, int S, double dx,
            const double* D, const double* lambda, double dt, int steps, int phases,
            double* rho,
            int n_dir, const int* dir_idx, const double* dir_val,
            const double* bulk_supply, const double* bulk_target, const double* bulk_uptake,
            int n_agents, const double* agent_pos, const double* agent_volume,
            const double* agent_secretion, const double* agent_saturation, const double* agent_uptake,
            int* out_voxel, int nthreads, double* seconds)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int rc = 0;
    void* R = ref_create(nx, ny, nz, S, dx, D, lambda, dt, phases, rho, n_dir, dir_idx, dir_val, bulk_supply,
                         bulk_target, bulk_uptake, n_agents, agent_pos, agent_volume, agent_secretion,
                         agent_saturation, agent_uptake, out_voxel, &rc);
    if (!R) return rc;
    ref_steps(R, dt, steps, phases, nthreads, seconds);
    ref_read(R, rho);
    ref_destroy(R);
    return 0;
}

/*
 * Gradient vectors of the reference (Microenvironment::compute_all_gradient_vectors,
 * BioFVM_microenvironment.cpp:699-779) of the densities rho on a uniform grid:
 * out[(n*S + s)*3 + axis].  The reference keeps 2*dx in a function-static
 * set on its first call in a process (:701-710): every call in one process
 * must use the same dx.
 */
int ref_gradients(int nx, int ny, int nz, int S, double dx, const double* rho, double* out)
{
    Microenvironment* M = new Microenvironment;
    set_default_microenvironment(M);
    M->set_density(0, "s0", "dimensionless");
    for (int s = 1; s < S; ++s) M->add_density("s", "dimensionless");
    M->resize_space_uniform(0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx);
    if ((int)M->mesh.x_coordinates.size() != nx || (int)M->mesh.y_coordinates.size() != ny ||
        (int)M->mesh.z_coordinates.size() != nz || M->number_of_densities() != S) {
        set_default_microenvironment(nullptr);
        delete M;
        return -1;
    }
    const int nvox = nx * ny * nz;
    for (int v = 0; v < nvox; ++v) {
        std::vector<double>& d = M->density_vector(v);
        for (int s = 0; s < S; ++s) d[s] = rho[(size_t)v * S + s];
    }
    // The neighbour strides the gradient uses (thomas_*_jump) are set only by
    // the LOD solver's setup (BioFVM_solvers.cpp:100-103).  One step with
    // D = lambda = 0 sets them and leaves every density unchanged (denom = 1,
    // c = 0: each sweep is the identity).
    M->diffusion_decay_solver = diffusion_decay_solver__constant_coefficients_LOD_3D;
    for (int s = 0; s < S; ++s) {
        M->diffusion_coefficients[s] = 0.0;
        M->decay_rates[s] = 0.0;
    }
    M->simulate_diffusion_decay(1.0);
    M->compute_all_gradient_vectors();
    for (int v = 0; v < nvox; ++v)
        for (int s = 0; s < S; ++s)
            for (int a = 0; a < 3; ++a) out[((size_t)v * S + s) * 3 + a] = M->gradient_vector(v)[s][a];
    set_default_microenvironment(nullptr);
    delete M;
    return 0;
}

/*
 * The reference's own MAT v4 snapshot (Microenvironment::write_to_matlab,
 * BioFVM_microenvironment.cpp:557-580) of densities rho on a uniform grid
 * with origin 0.
 */
int ref_write_matlab(int nx, int ny, int nz, int S, double dx, const double* rho, const char* filename)
{
    Microenvironment* M = new Microenvironment;
    set_default_microenvironment(M);
    M->set_density(0, "s0", "dimensionless");
    for (int s = 1; s < S; ++s) M->add_density("s", "dimensionless");
    M->resize_space_uniform(0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx);
    const int nvox = nx * ny * nz;
    int rc = -1;
    if ((int)M->mesh.x_coordinates.size() == nx && (int)M->mesh.y_coordinates.size() == ny &&
        (int)M->mesh.z_coordinates.size() == nz && M->number_of_densities() == S) {
        for (int v = 0; v < nvox; ++v) {
            std::vector<double>& d = M->density_vector(v);
            for (int s = 0; s < S; ++s) d[s] = rho[(size_t)v * S + s];
        }
        M->write_to_matlab(filename);
        rc = 0;
    }
    set_default_microenvironment(nullptr);
    delete M;
    return rc;
}

/*
 * MultiCellDS snapshot (save_microenvironment_to_MultiCellDS_xml_pugi,
 * BioFVM_MultiCellDS.cpp:394-720) of densities rho with coefficients D,
 * lambda on a uniform grid with origin 0: writes filename_base.xml and the
 * MAT files it names.  The reference keeps one document per process and only
 * updates it on later calls; `fresh` != 0 starts a new one (as a first call).
 */
int ref_write_multicellds(int nx, int ny, int nz, int S, double dx, const double* D, const double* lam,
                          const double* rho, const char* filename_base, double current_time,
                          double current_runtime, int mesh_as_matlab, int data_as_matlab, int fresh)
{
    Microenvironment* M = new Microenvironment;
    set_default_microenvironment(M);
    M->name = "microenvironment";
    M->set_density(0, "s0", "dimensionless");
    for (int s = 1; s < S; ++s) M->add_density("s" + std::to_string(s), "mmHg");
    M->resize_space_uniform(0.0, nx * dx, 0.0, ny * dx, 0.0, nz * dx, dx);
    M->spatial_units = "micron";
    M->mesh.units = "micron";
    const int nvox = nx * ny * nz;
    int rc = -1;
    if ((int)M->mesh.x_coordinates.size() == nx && (int)M->mesh.y_coordinates.size() == ny &&
        (int)M->mesh.z_coordinates.size() == nz && M->number_of_densities() == S) {
        for (int s = 0; s < S; ++s) {
            M->diffusion_coefficients[s] = D[s];
            M->decay_rates[s] = lam[s];
        }
        for (int v = 0; v < nvox; ++v) {
            std::vector<double>& d = M->density_vector(v);
            for (int s = 0; s < S; ++s) d[s] = rho[(size_t)v * S + s];
        }
        if (fresh) biofvm_doc.reset();
        default_metadata.current_time = current_time;
        default_metadata.current_runtime = current_runtime;
        set_save_biovm_mesh_as_matlab(mesh_as_matlab != 0);
        set_save_biovm_data_as_matlab(data_as_matlab != 0);
        save_microenvironment_to_MultiCellDS_xml_pugi(filename_base, *M);
        rc = 0;
    }
    set_default_microenvironment(nullptr);
    delete M;
    return rc;
}

}  // extern "C"
