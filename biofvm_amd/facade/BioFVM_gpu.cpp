// BioFVM-compatible host façade over the biofvm_amd C ABI (see BioFVM.h).
//
// Reference semantics kept (file:line in the reference tree):
//   mesh sizes / centres / volumes / uniform test   BioFVM_mesh.cpp:663-922
//   nearest_voxel_index (floor + clamp)             BioFVM_mesh.cpp:930-948
//   is_position_valid (closed box)                  BioFVM_mesh.cpp:208-217
//   add_density pushes onto every voxel             BioFVM_microenvironment.cpp:325-441
//   Dirichlet whole-vector overwrite before each sweep  microenvironment.cpp:192-197,
//                                                   BioFVM_solvers.cpp:177,207,241
//   bulk callbacks into persistent temp vectors     microenvironment.cpp:584-613
//   agent constants / update                        BioFVM_basic_agent.cpp:114-148, 229-238
//   solver refuses a non-uniform mesh (print, no-op) BioFVM_solvers.cpp:79-83
//
// Differences, by design:
//   - the Thomas factorisation is rebuilt when dt, D or lambda change (the
//     reference computes it once per Microenvironment, solvers.cpp:87,172);
//   - agents sharing a voxel are applied in list order on the device, i.e. the
//     reference's single-threaded result (its OpenMP loop races,
//     microenvironment.cpp:628-631).
#include "BioFVM.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "biofvm_amd.h"

namespace BioFVM {

std::vector<Basic_Agent*> all_basic_agents(0);

namespace {
Microenvironment* g_default_microenvironment = nullptr;
unsigned long long g_agent_revision = 1;

void report(const char* where, int rc, bfvm_handle h)
{
    std::cout << "biofvm_amd: " << where << " failed (" << rc << "): " << bfvm_last_error(h)
              << std::endl;
}
}  // namespace

void set_default_microenvironment(Microenvironment* M) { g_default_microenvironment = M; }
Microenvironment* get_default_microenvironment(void) { return g_default_microenvironment; }

void zero_function(Microenvironment*, int, std::vector<double>* write_destination)
{
    for (double& v : *write_destination) v = 0.0;
}

void one_function(Microenvironment*, int, std::vector<double>* write_destination)
{
    for (double& v : *write_destination) v = 1.0;
}

// ---------------------------------------------------------------- mesh

Voxel::Voxel() : mesh_index(0), volume(10 * 10 * 10), center(3, 0.0), is_Dirichlet(false) {}

Cartesian_Mesh::Cartesian_Mesh()
    : bounding_box(6, 0.0), dx(1), dy(1), dz(1), dV(1), dS(1), dS_xy(1), dS_yz(1), dS_xz(1),
      Cartesian_mesh(true), uniform_mesh(true), regular_mesh(true), use_voxel_faces(false),
      units("none")
{
    bounding_box[3] = bounding_box[4] = bounding_box[5] = 1.0;
    x_coordinates.assign(1, 0.5);
    y_coordinates.assign(1, 0.5);
    z_coordinates.assign(1, 0.5);
    fill_voxels();
}

int Cartesian_Mesh::voxel_index(int i, int j, int k)
{
    return (k * (int)y_coordinates.size() + j) * (int)x_coordinates.size() + i;
}

std::vector<int> Cartesian_Mesh::cartesian_indices(int n)
{
    std::vector<int> out(3, -1);
    int nx = (int)x_coordinates.size(), XY = nx * (int)y_coordinates.size();
    out[2] = n / XY;
    out[1] = (n - out[2] * XY) / nx;
    out[0] = n - nx * (out[1] + (int)y_coordinates.size() * out[2]);
    return out;
}

void Cartesian_Mesh::fill_voxels()
{
    dV = dx * dy * dz;
    dS = dx * dy;
    dS_xy = dx * dy;
    dS_yz = dy * dz;
    dS_xz = dx * dz;
    Voxel tv;
    tv.volume = dV;
    voxels.assign(x_coordinates.size() * y_coordinates.size() * z_coordinates.size(), tv);
    int n = 0;
    for (size_t k = 0; k < z_coordinates.size(); ++k)
        for (size_t j = 0; j < y_coordinates.size(); ++j)
            for (size_t i = 0; i < x_coordinates.size(); ++i, ++n) {
                voxels[n].center[0] = x_coordinates[i];
                voxels[n].center[1] = y_coordinates[j];
                voxels[n].center[2] = z_coordinates[k];
                voxels[n].mesh_index = n;
                voxels[n].volume = dV;
            }
}

void Cartesian_Mesh::resize(double x_start, double x_end, double y_start, double y_end,
                            double z_start, double z_end, int x_nodes, int y_nodes, int z_nodes)
{
    x_coordinates.assign(x_nodes, 0.0);
    y_coordinates.assign(y_nodes, 0.0);
    z_coordinates.assign(z_nodes, 0.0);
    dx = (x_end - x_start) / (double)x_nodes;
    if (x_nodes < 2) dx = 1;
    dy = (y_end - y_start) / (double)y_nodes;
    if (y_nodes < 2) dy = 1;
    dz = (z_end - z_start) / (double)z_nodes;
    if (z_nodes < 2) dz = 1;
    uniform_mesh = true;
    regular_mesh = true;
    const double tol = 1e-16;
    if (std::fabs(dx - dy) > tol && x_nodes > 1 && y_nodes > 1) uniform_mesh = false;
    if (std::fabs(dy - dz) > tol && y_nodes > 1 && z_nodes > 1) uniform_mesh = false;
    if (std::fabs(dx - dz) > tol && x_nodes > 1 && z_nodes > 1) uniform_mesh = false;
    for (int i = 0; i < x_nodes; ++i) x_coordinates[i] = x_start + (i + 0.5) * dx;
    for (int i = 0; i < y_nodes; ++i) y_coordinates[i] = y_start + (i + 0.5) * dy;
    for (int i = 0; i < z_nodes; ++i) z_coordinates[i] = z_start + (i + 0.5) * dz;
    bounding_box = {x_start, y_start, z_start, x_end, y_end, z_end};
    fill_voxels();
}

void Cartesian_Mesh::resize(int x_nodes, int y_nodes, int z_nodes)
{
    resize(0 - .5, x_nodes - 1 + .5, 0 - .5, y_nodes - 1 + .5, 0 - .5, z_nodes - 1 + .5, x_nodes,
           y_nodes, z_nodes);
}

void Cartesian_Mesh::resize(double x_start, double x_end, double y_start, double y_end,
                            double z_start, double z_end, double dx_new, double dy_new,
                            double dz_new)
{
    dx = dx_new;
    dy = dy_new;
    dz = dz_new;
    const double eps = 1e-16;
    int x_nodes = (int)std::ceil(eps + (x_end - x_start) / dx);
    int y_nodes = (int)std::ceil(eps + (y_end - y_start) / dy);
    int z_nodes = (int)std::ceil(eps + (z_end - z_start) / dz);
    x_coordinates.assign(x_nodes, 0.0);
    y_coordinates.assign(y_nodes, 0.0);
    z_coordinates.assign(z_nodes, 0.0);
    uniform_mesh = true;
    regular_mesh = true;
    const double tol = 1e-16;
    if (std::fabs(dx - dy) > tol || std::fabs(dy - dz) > tol || std::fabs(dx - dz) > tol)
        uniform_mesh = false;
    for (int i = 0; i < x_nodes; ++i) x_coordinates[i] = x_start + (i + 0.5) * dx;
    for (int i = 0; i < y_nodes; ++i) y_coordinates[i] = y_start + (i + 0.5) * dy;
    for (int i = 0; i < z_nodes; ++i) z_coordinates[i] = z_start + (i + 0.5) * dz;
    bounding_box = {x_start, y_start, z_start, x_end, y_end, z_end};
    fill_voxels();
}

void Cartesian_Mesh::resize_uniform(double x_start, double x_end, double y_start, double y_end,
                                    double z_start, double z_end, double dx_new)
{
    resize(x_start, x_end, y_start, y_end, z_start, z_end, dx_new, dx_new, dx_new);
}

bool Cartesian_Mesh::is_position_valid(double x, double y, double z)
{
    if (x < bounding_box[0] || x > bounding_box[3]) return false;
    if (y < bounding_box[1] || y > bounding_box[4]) return false;
    if (z < bounding_box[2] || z > bounding_box[5]) return false;
    return true;
}

std::vector<int> Cartesian_Mesh::nearest_cartesian_indices(std::vector<double>& position)
{
    std::vector<int> out(3, 0);
    out[0] = (int)std::floor((position[0] - bounding_box[0]) / dx);
    out[1] = (int)std::floor((position[1] - bounding_box[1]) / dy);
    out[2] = (int)std::floor((position[2] - bounding_box[2]) / dz);
    const int n[3] = {(int)x_coordinates.size(), (int)y_coordinates.size(),
                      (int)z_coordinates.size()};
    for (int d = 0; d < 3; ++d) {
        if (out[d] >= n[d]) out[d] = n[d] - 1;
        if (out[d] < 0) out[d] = 0;
    }
    return out;
}

int Cartesian_Mesh::nearest_voxel_index(std::vector<double>& position)
{
    std::vector<int> c = nearest_cartesian_indices(position);
    return voxel_index(c[0], c[1], c[2]);
}

Voxel& Cartesian_Mesh::nearest_voxel(std::vector<double>& position)
{
    return voxels[nearest_voxel_index(position)];
}

void Cartesian_Mesh::display_information(std::ostream& os)
{
    os << "Mesh information: " << std::endl
       << "type: uniform Cartesian" << std::endl
       << "Domain: [" << bounding_box[0] << "," << bounding_box[3] << "] " << units << " x ["
       << bounding_box[1] << "," << bounding_box[4] << "] " << units << " x [" << bounding_box[2]
       << "," << bounding_box[5] << "] " << units << std::endl
       << "   resolution: dx = " << dx << " " << units << std::endl
       << "   voxels: " << voxels.size() << std::endl;
}

// ---------------------------------------------------------------- device state

struct DeviceState {
    bfvm_handle h = nullptr;
    int lod_dims = 3;           // the LOD solver the handle is set for
    bool host_newer = true;     // the mirror holds changes the device lacks
    bool device_newer = false;  // the device holds changes the mirror lacks
    bool dirichlet_dirty = true;
    // the agents last uploaded: the global agent revision then, and the list
    // (its address, storage and length); any change to an agent's voxel or
    // constants, or any agent created or deleted, moves the revision
    unsigned long long agents_revision = 0;
    const void* agents_list = nullptr;
    const void* agents_data = nullptr;
    size_t agents_count = (size_t)-1;
    bool tuned = false;  // launch shapes picked (bfvm_autotune) for this handle
    // bulk (BFVM_BULK) and cell (BFVM_CELLS) phases requested but not yet
    // enqueued: they go with the next diffusion step, or alone before any
    // density is read (pending_dt: the bulk phase's dt)
    int pending = 0;
    double pending_dt = 0;
    // the device holds the bulk fields of these three callbacks
    bool bulk_loaded = false;
    void* bulk_callbacks[3] = {nullptr, nullptr, nullptr};
    // the callbacks leave some entries unwritten, which the reference's
    // per-step update carries over from its previous step (see
    // load_bulk_fields): evaluated at every step
    bool bulk_carry = false;
    std::vector<double> flat;  // staging for host<->device copies
};

// ---------------------------------------------------------------- Microenvironment

Microenvironment::Microenvironment()
    : spatial_units("none"), time_units("none"), name("unnamed"), device_precision(0),
      bulk_rates_every_step(false), bulk_rates_host_seconds(0.0), dev(nullptr)
{
    diffusion_decay_solver = diffusion_decay_solver__constant_coefficients_LOD_3D;
    bulk_supply_rate_function = zero_function;
    bulk_supply_target_densities_function = zero_function;
    bulk_uptake_rate_function = zero_function;
    one.assign(1, 1.0);
    zero.assign(1, 0.0);
    density_names.assign(1, "unnamed");
    density_units.assign(1, "none");
    diffusion_coefficients.assign(1, 0.0);
    decay_rates.assign(1, 0.0);
    mesh.resize(1, 1, 1);
    densities.assign(mesh.voxels.size(), zero);
    if (!g_default_microenvironment) g_default_microenvironment = this;
}

Microenvironment::Microenvironment(std::string name_) : Microenvironment() { name = name_; }

Microenvironment::~Microenvironment()
{
    if (dev) {
        if (dev->h) bfvm_destroy(dev->h);
        delete dev;
    }
    if (g_default_microenvironment == this) g_default_microenvironment = nullptr;
}

int Microenvironment::number_of_densities(void) { return (int)densities[0].size(); }
int Microenvironment::number_of_voxels(void) { return (int)mesh.voxels.size(); }

// Any change of grid or substrate count drops the device copy; it is rebuilt
// from the mirror on the next simulate_* call.
void Microenvironment::reset_space(void)
{
    if (dev) {
        if (dev->h) bfvm_destroy(dev->h);
        delete dev;
        dev = nullptr;
    }
    bulk_temp_S.clear();
    bulk_temp_T.clear();
    bulk_temp_U.clear();
}

void Microenvironment::resize_space(int x_nodes, int y_nodes, int z_nodes)
{
    host_view();
    mesh.resize(x_nodes, y_nodes, z_nodes);
    densities.assign(mesh.voxels.size(), zero);
    reset_space();
}

void Microenvironment::resize_space(double x_start, double x_end, double y_start, double y_end,
                                    double z_start, double z_end, int x_nodes, int y_nodes,
                                    int z_nodes)
{
    host_view();
    mesh.resize(x_start, x_end, y_start, y_end, z_start, z_end, x_nodes, y_nodes, z_nodes);
    densities.assign(mesh.voxels.size(), zero);
    reset_space();
}

void Microenvironment::resize_space(double x_start, double x_end, double y_start, double y_end,
                                    double z_start, double z_end, double dx_new, double dy_new,
                                    double dz_new)
{
    host_view();
    mesh.resize(x_start, x_end, y_start, y_end, z_start, z_end, dx_new, dy_new, dz_new);
    densities.assign(mesh.voxels.size(), zero);
    reset_space();
}

void Microenvironment::resize_space_uniform(double x_start, double x_end, double y_start,
                                            double y_end, double z_start, double z_end,
                                            double dx_new)
{
    resize_space(x_start, x_end, y_start, y_end, z_start, z_end, dx_new, dx_new, dx_new);
}

void Microenvironment::resize_densities(int new_size)
{
    host_view();
    zero.assign(new_size, 0.0);
    one.assign(new_size, 1.0);
    densities.assign(mesh.voxels.size(), zero);
    diffusion_coefficients.assign(new_size, 0.0);
    decay_rates.assign(new_size, 0.0);
    density_names.assign(new_size, "unnamed");
    density_units.assign(new_size, "none");
    reset_space();
}

void Microenvironment::add_density(std::string name_, std::string units, double diffusion_constant,
                                   double decay_rate)
{
    host_view();
    zero.push_back(0.0);
    one.push_back(1.0);
    density_names.push_back(name_);
    density_units.push_back(units);
    diffusion_coefficients.push_back(diffusion_constant);
    decay_rates.push_back(decay_rate);
    for (auto& d : densities) d.push_back(0.0);
    reset_space();
}

void Microenvironment::add_density(void) { add_density("unnamed", "none", 0.0, 0.0); }

void Microenvironment::add_density(std::string name_, std::string units)
{
    add_density(name_, units, 0.0, 0.0);
}

void Microenvironment::set_density(int index, std::string name_, std::string units)
{
    density_names[index] = name_;
    density_units[index] = units;
}

void Microenvironment::set_density(int index, std::string name_, std::string units,
                                   double diffusion_constant, double decay_rate)
{
    density_names[index] = name_;
    density_units[index] = units;
    diffusion_coefficients[index] = diffusion_constant;
    decay_rates[index] = decay_rate;
}

int Microenvironment::voxel_index(int i, int j, int k) { return mesh.voxel_index(i, j, k); }
std::vector<int> Microenvironment::cartesian_indices(int n) { return mesh.cartesian_indices(n); }
int Microenvironment::nearest_voxel_index(std::vector<double>& position)
{
    return mesh.nearest_voxel_index(position);
}
Voxel& Microenvironment::voxels(int voxel_index) { return mesh.voxels[voxel_index]; }
Voxel& Microenvironment::nearest_voxel(std::vector<double>& position)
{
    return mesh.nearest_voxel(position);
}

// Density access: bring the mirror up to date and mark it the newer copy
// (the caller may write through the returned reference).
void Microenvironment::host_view(void)
{
    refresh_mirror();
    if (dev) dev->host_newer = true;
}

void Microenvironment::refresh_mirror(void)
{
    if (dev && dev->pending) flush_pending();
    if (dev && dev->device_newer && dev->h) {
        const size_t S = densities[0].size();
        dev->flat.resize(densities.size() * S);
        int rc = bfvm_download_density(dev->h, dev->flat.data());
        if (rc) report("download", rc, dev->h);
        for (size_t v = 0; v < densities.size(); ++v)
            std::memcpy(densities[v].data(), &dev->flat[v * S], S * sizeof(double));
        dev->device_newer = false;
    }
}

void Microenvironment::synchronize_host(void) { host_view(); }

// Every member but the device copy; the source's mirror is brought up to date
// without marking it newer, so the source keeps its device copy.
void Microenvironment::copy_from(const Microenvironment& o)
{
    const_cast<Microenvironment&>(o).refresh_mirror();
    mesh = o.mesh;
    spatial_units = o.spatial_units;
    time_units = o.time_units;
    name = o.name;
    density_names = o.density_names;
    density_units = o.density_units;
    diffusion_coefficients = o.diffusion_coefficients;
    decay_rates = o.decay_rates;
    diffusion_decay_solver = o.diffusion_decay_solver;
    bulk_supply_rate_function = o.bulk_supply_rate_function;
    bulk_supply_target_densities_function = o.bulk_supply_target_densities_function;
    bulk_uptake_rate_function = o.bulk_uptake_rate_function;
    gradient_vectors = o.gradient_vectors;
    gradient_vector_computed = o.gradient_vector_computed;
    supply_target_densities_times_supply_rates = o.supply_target_densities_times_supply_rates;
    supply_rates = o.supply_rates;
    uptake_rates = o.uptake_rates;
    dirichlet_substrate_mask = o.dirichlet_substrate_mask;
    device_precision = o.device_precision;
    bulk_rates_every_step = o.bulk_rates_every_step;
    bulk_rates_host_seconds = o.bulk_rates_host_seconds;
    densities = o.densities;
    dirichlet_indices = o.dirichlet_indices;
    dirichlet_value_vectors = o.dirichlet_value_vectors;
    bulk_temp_S = o.bulk_temp_S;
    bulk_temp_T = o.bulk_temp_T;
    bulk_temp_U = o.bulk_temp_U;
    zero = o.zero;
    one = o.one;
}

Microenvironment::Microenvironment(const Microenvironment& other) : dev(nullptr) { copy_from(other); }

Microenvironment& Microenvironment::operator=(const Microenvironment& other)
{
    if (this != &other) {
        reset_space();
        copy_from(other);
    }
    return *this;
}

// The bulk and cell phases requested since the last device step, as one step.
int Microenvironment::flush_pending(void)
{
    if (!dev || !dev->h || !dev->pending) return BFVM_OK;
    const int flags = dev->pending;
    dev->pending = 0;
    int rc = bfvm_step(dev->h, dev->pending_dt, flags);
    if (rc) report("simulate_bulk/cell_sources_and_sinks", rc, dev->h);
    dev->device_newer = true;
    return rc;
}

void Microenvironment::wait_for_device(void)
{
    if (!dev || !dev->h) return;
    int rc = flush_pending();
    if (!rc) rc = bfvm_sync(dev->h);
    if (rc) report("wait_for_device", rc, dev->h);
}

bfvm_env* Microenvironment::device_handle(void) { return dev ? dev->h : nullptr; }

std::vector<double>& Microenvironment::nearest_density_vector(std::vector<double>& position)
{
    host_view();
    return densities[mesh.nearest_voxel_index(position)];
}
std::vector<double>& Microenvironment::nearest_density_vector(int voxel_index)
{
    host_view();
    return densities[voxel_index];
}
std::vector<double>& Microenvironment::operator()(int i, int j, int k)
{
    host_view();
    return densities[voxel_index(i, j, k)];
}
std::vector<double>& Microenvironment::operator()(int i, int j)
{
    host_view();
    return densities[voxel_index(i, j, 0)];
}
std::vector<double>& Microenvironment::operator()(int n)
{
    host_view();
    return densities[n];
}
std::vector<double>& Microenvironment::density_vector(int i, int j, int k)
{
    host_view();
    return densities[voxel_index(i, j, k)];
}
std::vector<double>& Microenvironment::density_vector(int i, int j)
{
    host_view();
    return densities[voxel_index(i, j, 0)];
}
std::vector<double>& Microenvironment::density_vector(int n)
{
    host_view();
    return densities[n];
}

DeviceState* Microenvironment::device(void)
{
    if (!dev) dev = new DeviceState;
    if (!dev->h) {
        const int nx = (int)mesh.x_coordinates.size(), ny = (int)mesh.y_coordinates.size(),
                  nz = (int)mesh.z_coordinates.size();
        // the LOD solver uses mesh.dx on every axis (BioFVM_solvers.cpp:111-113)
        // and refuses a mesh the reference's own test calls non-uniform
        double dy = mesh.uniform_mesh ? mesh.dx : mesh.dy;
        double dz = mesh.uniform_mesh ? mesh.dx : mesh.dz;
        int rc = bfvm_create_mesh(nx, ny, nz, number_of_densities(), mesh.dx, dy, dz,
                                  device_precision, &dev->h);
        if (rc) {
            report("bfvm_create_mesh", rc, nullptr);
            dev->h = nullptr;
            return dev;
        }
        dev->host_newer = true;
        dev->device_newer = false;
        dev->dirichlet_dirty = true;
        dev->agents_count = (size_t)-1;
        dev->tuned = false;
    }
    return dev;
}

int Microenvironment::push_to_device(void)
{
    DeviceState* d = device();
    if (!d->h) return BFVM_E_DEVICE;
    int rc;
    if (d->host_newer) {
        const size_t S = densities[0].size();
        d->flat.resize(densities.size() * S);
        for (size_t v = 0; v < densities.size(); ++v)
            std::memcpy(&d->flat[v * S], densities[v].data(), S * sizeof(double));
        if ((rc = bfvm_upload_density(d->h, d->flat.data()))) return rc;
        d->host_newer = false;
    }
    if ((rc = bfvm_set_coefficients(d->h, diffusion_coefficients.data(), decay_rates.data())))
        return rc;
    if (d->dirichlet_dirty) {
        const size_t S = densities[0].size();
        std::vector<double> vals(dirichlet_indices.size() * S);
        for (size_t q = 0; q < dirichlet_indices.size(); ++q)
            std::memcpy(&vals[q * S], dirichlet_value_vectors[q].data(), S * sizeof(double));
        const unsigned char* mask =
            dirichlet_substrate_mask.size() == S ? dirichlet_substrate_mask.data() : nullptr;
        if ((rc = bfvm_set_dirichlet(d->h, (int)dirichlet_indices.size(), dirichlet_indices.data(),
                                     vals.data(), mask)))
            return rc;
        d->dirichlet_dirty = false;
    }
    return BFVM_OK;
}

// ---------------------------------------------------------------- gradients, output

void Microenvironment::reset_all_gradient_vectors(void)
{
    const size_t S = densities[0].size();
    gradient_vectors.assign(densities.size(), std::vector<gradient>(S, gradient(3, 0.0)));
    gradient_vector_computed.assign(densities.size(), false);
}

// On the GPU from the device copy (pushed first if the mirror is newer); the
// result is copied into gradient_vectors and every voxel marked computed.
void Microenvironment::compute_all_gradient_vectors(void)
{
    const size_t nvox = densities.size(), S = densities[0].size();
    if (gradient_vectors.size() != nvox || gradient_vectors[0].size() != S) reset_all_gradient_vectors();
    DeviceState* d = device();
    int rc = d->h ? push_to_device() : BFVM_E_DEVICE;
    if (!rc) rc = flush_pending();
    if (!rc) rc = bfvm_compute_gradients(d->h);
    std::vector<double> g(nvox * S * 3);
    if (!rc) rc = bfvm_download_gradients(d->h, g.data());
    if (rc) {
        report("compute_all_gradient_vectors", rc, d->h);
        return;
    }
    for (size_t v = 0; v < nvox; ++v) {
        for (size_t s = 0; s < S; ++s)
            for (int a = 0; a < 3; ++a) gradient_vectors[v][s][a] = g[(v * S + s) * 3 + a];
        gradient_vector_computed[v] = true;
    }
}

// One voxel on the host from the mirror, as the reference
// (BioFVM_microenvironment.cpp:781-824), with this mesh's spacings.
void Microenvironment::compute_gradient_vector(int n)
{
    const size_t S = densities[0].size();
    if (gradient_vectors.size() != densities.size() || gradient_vectors[0].size() != S)
        reset_all_gradient_vectors();
    host_view();
    const std::vector<int> ijk = cartesian_indices(n);
    const int counts[3] = {(int)mesh.x_coordinates.size(), (int)mesh.y_coordinates.size(),
                           (int)mesh.z_coordinates.size()};
    const int jumps[3] = {1, counts[0], counts[0] * counts[1]};
    double two_d[3] = {mesh.dx, mesh.dy, mesh.dz};
    for (int a = 0; a < 3; ++a) {
        two_d[a] *= 2.0;
        if (ijk[a] <= 0 || ijk[a] >= counts[a] - 1) continue;
        for (size_t q = 0; q < S; ++q) {
            double g = densities[n + jumps[a]][q];
            g -= densities[n - jumps[a]][q];
            g /= two_d[a];
            gradient_vectors[n][q][a] = g;
        }
        gradient_vector_computed[n] = true;
    }
}

std::vector<gradient>& Microenvironment::gradient_vector(int n)
{
    if (gradient_vector_computed.size() != densities.size() || !gradient_vector_computed[n])
        compute_gradient_vector(n);
    return gradient_vectors[n];
}
std::vector<gradient>& Microenvironment::gradient_vector(int i, int j, int k)
{
    return gradient_vector(voxel_index(i, j, k));
}
std::vector<gradient>& Microenvironment::gradient_vector(int i, int j)
{
    return gradient_vector(voxel_index(i, j, 0));
}
std::vector<gradient>& Microenvironment::nearest_gradient_vector(std::vector<double>& position)
{
    return gradient_vector(nearest_voxel_index(position));
}

void Microenvironment::write_to_matlab(std::string filename)
{
    DeviceState* d = device();
    int rc = d->h ? push_to_device() : BFVM_E_DEVICE;
    if (!rc) rc = flush_pending();
    if (!rc)
        rc = bfvm_write_matlab(d->h, filename.c_str(), mesh.bounding_box[0], mesh.bounding_box[1],
                               mesh.bounding_box[2]);
    if (rc) report("write_to_matlab", rc, d->h);
}

void Microenvironment::add_dirichlet_node(int voxel_index, std::vector<double>& value)
{
    mesh.voxels[voxel_index].is_Dirichlet = true;
    dirichlet_indices.push_back(voxel_index);
    dirichlet_value_vectors.push_back(value);
    if (dev) dev->dirichlet_dirty = true;
}

void Microenvironment::update_dirichlet_node(int voxel_index, std::vector<double>& new_value)
{
    for (size_t n = 0; n < dirichlet_indices.size(); ++n)
        if (dirichlet_indices[n] == voxel_index) {
            dirichlet_value_vectors[n] = new_value;
            if (dev) dev->dirichlet_dirty = true;
            return;
        }
    std::cout << "BioFVM Warning: No Dirichlet condition previously specified at voxel "
              << voxel_index << "! Creating a new one now ... " << std::endl;
    add_dirichlet_node(voxel_index, new_value);
}

// The reference swaps with the last node but pops the value list twice and
// never the index list (BioFVM_microenvironment.cpp:183-187); here the node is
// removed from both lists.
void Microenvironment::remove_dirichlet_node(int voxel_index)
{
    mesh.voxels[voxel_index].is_Dirichlet = false;
    for (size_t n = 0; n < dirichlet_indices.size(); ++n)
        if (dirichlet_indices[n] == voxel_index) {
            dirichlet_indices[n] = dirichlet_indices.back();
            dirichlet_value_vectors[n] = dirichlet_value_vectors.back();
            dirichlet_indices.pop_back();
            dirichlet_value_vectors.pop_back();
            if (dev) dev->dirichlet_dirty = true;
            return;
        }
    std::cout << "BioFVM Warning: No Dirichlet condition previously specified at voxel "
              << voxel_index << "! Nothing to remove!" << std::endl;
}

void Microenvironment::apply_dirichlet_conditions(void)
{
    host_view();
    const size_t S = densities[0].size();
    for (size_t n = 0; n < dirichlet_indices.size(); ++n) {
        std::vector<double>& d = densities[dirichlet_indices[n]];
        for (size_t s = 0; s < S; ++s)
            if (dirichlet_substrate_mask.size() != S || dirichlet_substrate_mask[s])
                d[s] = dirichlet_value_vectors[n][s];
    }
}

void Microenvironment::simulate_diffusion_decay(double dt)
{
    if (diffusion_decay_solver) {
        diffusion_decay_solver(*this, dt);
    } else {
        std::cout << "Warning: diffusion-reaction-source/sink solver not set for Microenvironment "
                     "object at "
                  << this << ". Nothing happened!" << std::endl;
    }
}

// The three bulk callbacks evaluated per voxel into the persistent vectors S,
// T, U (the reference's bulk_source_sink_solver_temp1..3), S, T and U handed
// to the device, which builds dt*(T*S) and (U+S)*dt + 1 in the reference's
// operation order.  The reference updates T and U in place each step
// (temp2 *= temp1; temp3 += temp1; temp3 *= dt; temp3 += one:
// BioFVM_microenvironment.cpp:603-609), so an entry a callback does not write
// at the next step carries that updated value; evaluated at every step
// (bulk_rates_every_step, or callbacks found to leave entries unwritten),
// the host vectors are updated the same way after each hand-over.  Whether
// the callbacks write every entry is found at the first evaluation (vectors
// filled with NaN first); if one does not, every later step evaluates them,
// from zero vectors as the reference starts.
int Microenvironment::load_bulk_fields(double dt)
{
    const auto t0 = std::chrono::steady_clock::now();
    const size_t nvox = densities.size(), S = densities[0].size();
    const bool fresh = bulk_temp_S.size() != nvox;
    if (fresh) {
        bulk_temp_S.assign(nvox, zero);
        bulk_temp_T.assign(nvox, zero);
        bulk_temp_U.assign(nvox, zero);
    }
    auto evaluate = [&] {
        for (size_t i = 0; i < nvox; ++i) {
            bulk_supply_rate_function(this, (int)i, &bulk_temp_S[i]);
            bulk_supply_target_densities_function(this, (int)i, &bulk_temp_T[i]);
            bulk_uptake_rate_function(this, (int)i, &bulk_temp_U[i]);
        }
    };
    bool carry = bulk_rates_every_step || dev->bulk_carry || !fresh;
    if (!carry) {
        const std::vector<double> nan(S, std::nan(""));
        for (size_t i = 0; i < nvox; ++i) bulk_temp_S[i] = bulk_temp_T[i] = bulk_temp_U[i] = nan;
        evaluate();
        bool all = true;
        for (size_t i = 0; i < nvox && all; ++i)
            for (size_t q = 0; q < S; ++q)
                if (std::isnan(bulk_temp_S[i][q]) || std::isnan(bulk_temp_T[i][q]) || std::isnan(bulk_temp_U[i][q])) {
                    all = false;
                    break;
                }
        if (!all) {
            carry = true;
            bulk_temp_S.assign(nvox, zero);
            bulk_temp_T.assign(nvox, zero);
            bulk_temp_U.assign(nvox, zero);
        }
    }
    if (carry) {
        dev->bulk_carry = true;
        evaluate();
    }
    std::vector<double> fs(nvox * S), ft(nvox * S), fu(nvox * S);
    for (size_t i = 0; i < nvox; ++i) {
        std::memcpy(&fs[i * S], bulk_temp_S[i].data(), S * sizeof(double));
        std::memcpy(&ft[i * S], bulk_temp_T[i].data(), S * sizeof(double));
        std::memcpy(&fu[i * S], bulk_temp_U[i].data(), S * sizeof(double));
    }
    int rc = bfvm_set_bulk(dev->h, fs.data(), ft.data(), fu.data());
    if (!rc && carry) {
        for (size_t i = 0; i < nvox; ++i)
            for (size_t q = 0; q < S; ++q) {
                bulk_temp_T[i][q] *= bulk_temp_S[i][q];
                double u = bulk_temp_U[i][q];
                u += bulk_temp_S[i][q];
                u *= dt;
                u += 1.0;
                bulk_temp_U[i][q] = u;
            }
    }
    if (!rc) {
        dev->bulk_loaded = true;
        dev->bulk_callbacks[0] = (void*)bulk_supply_rate_function;
        dev->bulk_callbacks[1] = (void*)bulk_supply_target_densities_function;
        dev->bulk_callbacks[2] = (void*)bulk_uptake_rate_function;
    }
    bulk_rates_host_seconds +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// update_rates (BioFVM_microenvironment.cpp:635-656): the three fields per
// voxel, supply_target_densities_times_supply_rates = T*S; the device's
// bulk fields are evaluated again from the same callbacks.
void Microenvironment::update_rates(void)
{
    const size_t nvox = densities.size();
    if (supply_target_densities_times_supply_rates.size() != nvox)
        supply_target_densities_times_supply_rates.assign(nvox, zero);
    if (supply_rates.size() != nvox) supply_rates.assign(nvox, zero);
    if (uptake_rates.size() != nvox) uptake_rates.assign(nvox, zero);
    for (size_t i = 0; i < nvox; ++i) {
        bulk_uptake_rate_function(this, (int)i, &uptake_rates[i]);
        bulk_supply_rate_function(this, (int)i, &supply_rates[i]);
        bulk_supply_target_densities_function(this, (int)i, &supply_target_densities_times_supply_rates[i]);
        std::vector<double>& ts = supply_target_densities_times_supply_rates[i];
        for (size_t s = 0; s < ts.size(); ++s) ts[s] *= supply_rates[i][s];
    }
    if (dev) dev->bulk_loaded = false;  // the next bulk step hands them to the device
}

// rho = (rho + dt*(T*S)) / ((U+S)*dt + 1) (BioFVM_microenvironment.cpp:584-613)
// on the device.  The callbacks are evaluated at every call with
// bulk_rates_every_step or when they leave entries unwritten, otherwise only
// when the device lacks the fields of the callbacks set now (see BioFVM.h
// and load_bulk_fields).  The phase itself is enqueued with the next
// diffusion step (or before any density is read).
void Microenvironment::simulate_bulk_sources_and_sinks(double dt)
{
    int rc = push_to_device();
    // phases run in the order they were requested: a bulk or cell phase
    // already waiting goes first
    if (!rc && (dev->pending & (BFVM_BULK | BFVM_CELLS))) rc = flush_pending();
    if (rc) {
        report("simulate_bulk_sources_and_sinks", rc, dev ? dev->h : nullptr);
        return;
    }
    // with every callback zero_function the update is rho = (rho + 0)/1
    if (bulk_supply_rate_function == zero_function && bulk_supply_target_densities_function == zero_function &&
        bulk_uptake_rate_function == zero_function)
        return;
    const bool same = dev->bulk_loaded && dev->bulk_callbacks[0] == (void*)bulk_supply_rate_function &&
                      dev->bulk_callbacks[1] == (void*)bulk_supply_target_densities_function &&
                      dev->bulk_callbacks[2] == (void*)bulk_uptake_rate_function;
    if (bulk_rates_every_step || dev->bulk_carry || !same) rc = load_bulk_fields(dt);
    if (rc) {
        report("simulate_bulk_sources_and_sinks", rc, dev->h);
        return;
    }
    dev->pending |= BFVM_BULK;
    dev->pending_dt = dt;
}

void Microenvironment::simulate_cell_sources_and_sinks(std::vector<Basic_Agent*>& list, double dt)
{
    int rc = push_to_device();
    if (!rc && (dev->pending & BFVM_CELLS)) rc = flush_pending();
    if (rc) {
        report("simulate_cell_sources_and_sinks", rc, dev ? dev->h : nullptr);
        return;
    }
    if (g_agent_revision != dev->agents_revision || (const void*)&list != dev->agents_list ||
        (const void*)list.data() != dev->agents_data || list.size() != dev->agents_count) {
        const size_t S = densities[0].size();
        std::vector<int> vox(list.size());
        std::vector<double> t1(list.size() * S), t2(list.size() * S);
        for (size_t a = 0; a < list.size(); ++a) {
            vox[a] = list[a]->nearest_voxel_index();
            std::memcpy(&t1[a * S], list[a]->temp1().data(), S * sizeof(double));
            std::memcpy(&t2[a * S], list[a]->temp2().data(), S * sizeof(double));
        }
        if ((rc = bfvm_set_agents(dev->h, (int)list.size(), vox.data(), t1.data(), t2.data()))) {
            report("bfvm_set_agents", rc, dev->h);
            return;
        }
        dev->agents_revision = g_agent_revision;
        dev->agents_list = &list;
        dev->agents_data = list.data();
        dev->agents_count = list.size();
    }
    if (!(dev->pending & BFVM_BULK)) dev->pending_dt = dt;
    dev->pending |= BFVM_CELLS;
}

// The reference's global overload walks every agent of every microenvironment
// (BioFVM_microenvironment.cpp:626-633); here only the agents registered with
// this microenvironment are applied to it.
void Microenvironment::simulate_cell_sources_and_sinks(double dt)
{
    std::vector<Basic_Agent*> mine;
    mine.reserve(all_basic_agents.size());
    for (Basic_Agent* a : all_basic_agents)
        if (a->get_microenvironment() == this) mine.push_back(a);
    simulate_cell_sources_and_sinks(mine, dt);
}

void Microenvironment::display_information(std::ostream& os)
{
    os << std::endl << "Microenvironment summary: " << name << ": " << std::endl;
    mesh.display_information(os);
    os << "Densities: (" << number_of_densities() << " total)" << std::endl;
    for (size_t i = 0; i < density_names.size(); ++i)
        os << "   " << density_names[i] << ":" << std::endl
           << "     units: " << density_units[i] << std::endl
           << "     diffusion coefficient: " << diffusion_coefficients[i] << " " << spatial_units
           << "^2 / " << time_units << std::endl
           << "     decay rate: " << decay_rates[i] << " " << time_units << "^-1" << std::endl
           << "     diffusion length scale: "
           << std::sqrt(diffusion_coefficients[i] / (1e-12 + decay_rates[i])) << " "
           << spatial_units << std::endl
           << std::endl;
    os << "Device: biofvm_amd (MI355X), store "
       << (device_precision == BFVM_FP32 ? "FP32" : "FP64") << std::endl
       << std::endl;
}

// The first diffusion step of a handle picks its launch shapes (bfvm_autotune:
// every candidate timed on a scratch copy of the store, results unchanged but
// for the last bits of the segmented solve's rounding), as the reference's
// solver sets itself up on its first call (BioFVM_solvers.cpp:87-173).  The
// environment variable BIOFVM_AMD_AUTOTUNE=0 keeps the default shapes.
static int tune_once(DeviceState* d, double dt)
{
    if (d->tuned) return 0;
    d->tuned = true;
    const char* env = std::getenv("BIOFVM_AMD_AUTOTUNE");
    if (env && std::string(env) == "0") return 0;
    return bfvm_autotune(d->h, dt);
}

// One device step: the bulk and cell phases waiting (in the reference's order,
// bulk -> cells -> diffusion) and the diffusion step, so that the sources run
// fused into the x sweep; a bulk phase requested with another dt goes alone.
static int step_with_pending(DeviceState* d, double dt)
{
    int flags = BFVM_DIFFUSION;
    if (d->pending) {
        if ((d->pending & BFVM_BULK) && d->pending_dt != dt) {
            const int rc = bfvm_step(d->h, d->pending_dt, d->pending);
            d->pending = 0;
            if (rc) return rc;
        } else {
            flags |= d->pending;
            d->pending = 0;
        }
    }
    return bfvm_step(d->h, dt, flags);
}

// The GPU LOD_3D solver (replaces BioFVM_solvers.cpp:77-276).
void diffusion_decay_solver__constant_coefficients_LOD_3D(Microenvironment& M, double dt)
{
    if (M.mesh.uniform_mesh == false || M.mesh.Cartesian_mesh == false) {
        std::cout << "Error: This algorithm is written for uniform Cartesian meshes. Try: other "
                     "solvers!"
                  << std::endl
                  << std::endl;
        return;
    }
    int rc = M.push_to_device();
    if (!rc && M.dev->lod_dims != 3) {
        rc = bfvm_set_lod_dimensions(M.dev->h, 3);
        if (!rc) M.dev->lod_dims = 3;
    }
    if (!rc) rc = tune_once(M.dev, dt);
    if (!rc) rc = step_with_pending(M.dev, dt);
    if (rc) {
        report("diffusion_decay_solver__constant_coefficients_LOD_3D", rc, M.dev ? M.dev->h : nullptr);
        return;
    }
    M.dev->device_newer = true;
}

// BioFVM_solvers.cpp:278-427: x and y sweeps, decay split in halves.  The
// device solver needs a single z layer (nz = 1); the reference's would sweep
// the x and y lines of its first layer only.
void diffusion_decay_solver__constant_coefficients_LOD_2D(Microenvironment& M, double dt)
{
    if (M.mesh.uniform_mesh == false) {
        std::cout << "Error: This algorithm is written for uniform Cartesian meshes. Try: something else."
                  << std::endl
                  << std::endl;
        return;
    }
    int rc = M.push_to_device();
    if (!rc && M.dev->lod_dims != 2) {
        rc = bfvm_set_lod_dimensions(M.dev->h, 2);
        if (!rc) M.dev->lod_dims = 2;
    }
    if (!rc) rc = tune_once(M.dev, dt);
    if (!rc) rc = step_with_pending(M.dev, dt);
    if (rc) {
        report("diffusion_decay_solver__constant_coefficients_LOD_2D", rc, M.dev ? M.dev->h : nullptr);
        return;
    }
    M.dev->device_newer = true;
}

// ---------------------------------------------------------------- agents

Basic_Agent::Basic_Agent()
    : microenvironment(nullptr), current_voxel_index(0), ID(0), index(0), type(0), volume(0),
      position(3, 0.0), velocity(3, 0.0), revision(0)
{
    static int max_basic_agent_ID = 0;
    ID = max_basic_agent_ID++;
    secretion_rates = new std::vector<double>(0);
    uptake_rates = new std::vector<double>(0);
    saturation_densities = new std::vector<double>(0);
    if (g_default_microenvironment) register_microenvironment(g_default_microenvironment);
}

Basic_Agent::~Basic_Agent()
{
    delete secretion_rates;
    delete uptake_rates;
    delete saturation_densities;
}

void Basic_Agent::update_position(double) {}

bool Basic_Agent::assign_position(std::vector<double> p) { return assign_position(p[0], p[1], p[2]); }

bool Basic_Agent::assign_position(double x, double y, double z)
{
    if (!microenvironment || !microenvironment->mesh.is_position_valid(x, y, z)) return false;
    position = {x, y, z};
    current_voxel_index = microenvironment->mesh.nearest_voxel_index(position);
    revision = g_agent_revision++;
    return true;
}

#pragma GCC push_options
#pragma GCC optimize("fp-contract=off")
// temp1 = ((0 + S)*T)*c, temp2 = (1 + c*S) + c*U, c = dt*V/V_voxel
// (BioFVM_basic_agent.cpp:114-148)
void Basic_Agent::set_internal_uptake_constants(double dt)
{
    double c = dt * volume / microenvironment->voxels(current_voxel_index).volume;
    const size_t S = secretion_rates->size();
    cell_source_sink_solver_temp1.assign(S, 0.0);
    cell_source_sink_solver_temp2.assign(S, 1.0);
    for (size_t s = 0; s < S; ++s) {
        double a = cell_source_sink_solver_temp1[s];
        a += (*secretion_rates)[s];
        a *= (*saturation_densities)[s];
        a *= c;
        cell_source_sink_solver_temp1[s] = a;
        double b = cell_source_sink_solver_temp2[s];
        double cs = c * (*secretion_rates)[s];
        b += cs;
        double cu = c * (*uptake_rates)[s];
        b += cu;
        cell_source_sink_solver_temp2[s] = b;
    }
    revision = g_agent_revision++;
}
#pragma GCC pop_options

void Basic_Agent::register_microenvironment(Microenvironment* m)
{
    microenvironment = m;
    const size_t S = m->number_of_densities();
    secretion_rates->resize(S, 0.0);
    saturation_densities->resize(S, 0.0);
    uptake_rates->resize(S, 0.0);
    cell_source_sink_solver_temp1.resize(S, 0.0);
    cell_source_sink_solver_temp2.resize(S, 1.0);
    revision = g_agent_revision++;
}

Microenvironment* Basic_Agent::get_microenvironment(void) { return microenvironment; }

int Basic_Agent::nearest_voxel_index(void) { return current_voxel_index; }

std::vector<double>& Basic_Agent::nearest_density_vector(void)
{
    return microenvironment->nearest_density_vector(current_voxel_index);
}

// Host fallback of one agent's update (BioFVM_basic_agent.cpp:229-238), for
// code that calls it directly; the step methods run it on the device.
void Basic_Agent::simulate_secretion_and_uptake(Microenvironment* pS, double)
{
    std::vector<double>& d = pS->nearest_density_vector(current_voxel_index);
    for (size_t s = 0; s < d.size(); ++s) {
        d[s] += cell_source_sink_solver_temp1[s];
        d[s] /= cell_source_sink_solver_temp2[s];
    }
}

Basic_Agent* create_basic_agent(void)
{
    Basic_Agent* p = new Basic_Agent;
    all_basic_agents.push_back(p);
    p->index = (int)all_basic_agents.size() - 1;
    ++g_agent_revision;
    return p;
}

// Swap-with-last removal, as the reference (BioFVM_basic_agent.cpp:183-216).
void delete_basic_agent(int index)
{
    Basic_Agent* victim = all_basic_agents[index];
    Basic_Agent* last = all_basic_agents.back();
    last->index = index;
    all_basic_agents[index] = last;
    all_basic_agents.pop_back();
    delete victim;
    ++g_agent_revision;
}

void delete_basic_agent(Basic_Agent* p) { delete_basic_agent(p->index); }

}  // namespace BioFVM
