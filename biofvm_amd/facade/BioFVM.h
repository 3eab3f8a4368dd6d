/*
 * BioFVM-compatible host façade over the biofvm_amd C ABI.
 *
 * The classes and functions below keep the names, signatures and argument
 * meaning of BioFVM 1.0.4's hot-path API, so a program written against the
 * reference's Microenvironment / Basic_Agent (e.g. its examples/ drivers)
 * compiles against this header instead and runs its time loop on the GPU:
 *
 *   Microenvironment                 BioFVM_microenvironment.h:45-228
 *     diffusion_decay_solver         function pointer, h:129 (default: the GPU LOD_3D solver)
 *     bulk_*_function callbacks      h:130-132
 *     simulate_diffusion_decay       h:194, cpp:502-513
 *     simulate_bulk_sources_and_sinks h:197, cpp:584-613
 *     simulate_cell_sources_and_sinks h:200,202, cpp:615-633
 *     add/update/remove_dirichlet_node, apply_dirichlet_conditions  h:206-209
 *     density_vector / operator()    h:170-191
 *   Cartesian_Mesh (geometry + indexing)  BioFVM_mesh.h:146-188
 *   Basic_Agent, all_basic_agents, create/delete_basic_agent  BioFVM_basic_agent.h:36-94
 *   diffusion_decay_solver__constant_coefficients_LOD_3D       BioFVM_solvers.h:38
 *
 * Storage: the densities live on the GPU (substrate-innermost HBM array);
 * the std::vector<double> returned by density_vector() is a host mirror
 * synchronised lazily — any density access first copies the device state
 * down if it is newer, and marks the mirror as the newer copy, so the next
 * simulate_* call uploads it.  Also here: the gradient vectors
 * (compute_all_gradient_vectors on the GPU, BioFVM_microenvironment.cpp:658-842)
 * and write_to_matlab (:557-580, packed on the GPU); MultiCellDS snapshots
 * (save_microenvironment_to_MultiCellDS_xml_pugi, BioFVM_MultiCellDS.cpp:
 * 394-720) with the densities' MAT file packed on the GPU.  The general
 * (non-Cartesian) mesh and reading MultiCellDS files are not part of this
 * façade.
 *
 * Launch shapes: a microenvironment's first LOD solver call picks the sweep
 * shapes for its grid (bfvm_autotune: each candidate timed on a scratch copy
 * of the store), as the reference's solver sets itself up on its first call;
 * BIOFVM_AMD_AUTOTUNE=0 in the environment keeps the default shapes.
 */
#ifndef BIOFVM_AMD_FACADE_H
#define BIOFVM_AMD_FACADE_H

// the standard headers the reference's BioFVM.h brings in (through its
// BioFVM_utilities.h, BioFVM_matlab.h and BioFVM_mesh.h), which programs
// written against it may rely on
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

struct bfvm_env;  // the handle type of the biofvm_amd C ABI (biofvm_amd.h)

namespace BioFVM {

class Microenvironment;
class Basic_Agent;

class Voxel {
 public:
    Voxel();
    int mesh_index;
    double volume;
    std::vector<double> center;
    bool is_Dirichlet;
};

/* Uniform Cartesian mesh: node coordinates, bounding box, voxel volumes and
 * the index maps of BioFVM_mesh.cpp:641-661, 663-928, 930-973. */
class Cartesian_Mesh {
 public:
    Cartesian_Mesh();
    std::vector<double> x_coordinates, y_coordinates, z_coordinates;
    std::vector<double> bounding_box;  // [xmin ymin zmin xmax ymax zmax]
    std::vector<Voxel> voxels;
    double dx, dy, dz, dV, dS, dS_xy, dS_yz, dS_xz;
    bool Cartesian_mesh, uniform_mesh, regular_mesh, use_voxel_faces;
    // MAT v4: a 4 x voxels matrix "mesh", one column (centre x, y, z, volume)
    // per voxel (General_Mesh::write_to_matlab, BioFVM_mesh.cpp:350-367)
    void write_to_matlab(std::string filename);
    std::string units;

    int voxel_index(int i, int j, int k);
    std::vector<int> cartesian_indices(int n);
    void resize(int x_nodes, int y_nodes, int z_nodes);
    void resize(double x_start, double x_end, double y_start, double y_end, double z_start,
                double z_end, int x_nodes, int y_nodes, int z_nodes);
    void resize(double x_start, double x_end, double y_start, double y_end, double z_start,
                double z_end, double dx, double dy, double dz);
    void resize_uniform(double x_start, double x_end, double y_start, double y_end, double z_start,
                        double z_end, double dx);
    int nearest_voxel_index(std::vector<double>& position);
    std::vector<int> nearest_cartesian_indices(std::vector<double>& position);
    Voxel& nearest_voxel(std::vector<double>& position);
    bool is_position_valid(double x, double y, double z);
    void display_information(std::ostream& os);

 private:
    void fill_voxels();
};

struct DeviceState;  // the biofvm_amd handle and its synchronisation flags

typedef std::vector<double> gradient;  // BioFVM_microenvironment.h:39

class Microenvironment {
 public:
    Microenvironment();
    explicit Microenvironment(std::string name);
    ~Microenvironment();
    // A copy holds the densities and settings at the time of the copy (the
    // source's device state is read back first) and gets a device copy of
    // its own at its first device use; agents stay with the source.
    Microenvironment(const Microenvironment& other);
    Microenvironment& operator=(const Microenvironment& other);

    Cartesian_Mesh mesh;
    std::string spatial_units, time_units, name;
    std::vector<std::string> density_names, density_units;
    std::vector<double> diffusion_coefficients, decay_rates;

    void (*diffusion_decay_solver)(Microenvironment&, double);
    void (*bulk_supply_rate_function)(Microenvironment* pMicroenvironment, int voxel_index,
                                      std::vector<double>* write_destination);
    void (*bulk_supply_target_densities_function)(Microenvironment* pMicroenvironment,
                                                  int voxel_index,
                                                  std::vector<double>* write_destination);
    void (*bulk_uptake_rate_function)(Microenvironment* pMicroenvironment, int voxel_index,
                                      std::vector<double>* write_destination);

    int number_of_densities(void);
    int number_of_voxels(void);

    void resize_space(int x_nodes, int y_nodes, int z_nodes);
    void resize_space(double x_start, double x_end, double y_start, double y_end, double z_start,
                      double z_end, int x_nodes, int y_nodes, int z_nodes);
    void resize_space(double x_start, double x_end, double y_start, double y_end, double z_start,
                      double z_end, double dx_new, double dy_new, double dz_new);
    void resize_space_uniform(double x_start, double x_end, double y_start, double y_end,
                              double z_start, double z_end, double dx_new);
    void resize_densities(int new_size);
    void add_density(void);
    void add_density(std::string name, std::string units);
    void add_density(std::string name, std::string units, double diffusion_constant,
                     double decay_rate);
    void set_density(int index, std::string name, std::string units);
    void set_density(int index, std::string name, std::string units, double diffusion_constant,
                     double decay_rate);

    int voxel_index(int i, int j, int k);
    std::vector<int> cartesian_indices(int n);
    int nearest_voxel_index(std::vector<double>& position);
    Voxel& voxels(int voxel_index);
    Voxel& nearest_voxel(std::vector<double>& position);
    std::vector<double>& nearest_density_vector(std::vector<double>& position);
    std::vector<double>& nearest_density_vector(int voxel_index);
    std::vector<double>& operator()(int i, int j, int k);
    std::vector<double>& operator()(int i, int j);
    std::vector<double>& operator()(int n);
    std::vector<double>& density_vector(int i, int j, int k);
    std::vector<double>& density_vector(int i, int j);
    std::vector<double>& density_vector(int n);

    void simulate_diffusion_decay(double dt);
    void simulate_bulk_sources_and_sinks(double dt);
    void simulate_cell_sources_and_sinks(std::vector<Basic_Agent*>& basic_agent_list, double dt);
    void simulate_cell_sources_and_sinks(double dt);

    void add_dirichlet_node(int voxel_index, std::vector<double>& value);
    void update_dirichlet_node(int voxel_index, std::vector<double>& new_value);
    void remove_dirichlet_node(int voxel_index);
    void apply_dirichlet_conditions(void);

    void display_information(std::ostream& os);

    // Gradients (BioFVM_microenvironment.h:65-66, 176-184): vectors of S
    // 3-vectors per voxel; compute_all_gradient_vectors runs on the GPU.
    std::vector<std::vector<gradient> > gradient_vectors;
    std::vector<bool> gradient_vector_computed;
    std::vector<gradient>& gradient_vector(int i, int j, int k);
    std::vector<gradient>& gradient_vector(int i, int j);
    std::vector<gradient>& gradient_vector(int n);
    std::vector<gradient>& nearest_gradient_vector(std::vector<double>& position);
    void compute_all_gradient_vectors(void);
    void compute_gradient_vector(int n);
    void reset_all_gradient_vectors(void);

    // MAT v4 snapshot (BioFVM_microenvironment.cpp:557-580), packed on the GPU.
    void write_to_matlab(std::string filename);

    // Bulk source/sink fields (BioFVM_microenvironment.h:121-124, filled by
    // update_rates, BioFVM_microenvironment.cpp:635-656): the three callbacks
    // evaluated per voxel, supply_target_densities_times_supply_rates = T*S.
    // Here update_rates also hands S, T and U to the device, where the bulk
    // step reads them (see bulk_rates_every_step).
    std::vector<std::vector<double> > supply_target_densities_times_supply_rates;
    std::vector<std::vector<double> > supply_rates;
    std::vector<std::vector<double> > uptake_rates;
    void update_rates(void);

    /* Extensions (no counterpart in the reference). */
    // Restrict Dirichlet overwrites to substrates with mask[s] != 0
    // (empty = every substrate, the reference's behaviour).
    std::vector<unsigned char> dirichlet_substrate_mask;
    // Store precision of the device copy: 0 = FP64 (default), 1 = FP32 store
    // with FP64 arithmetic.  Takes effect at the next resize.
    int device_precision;
    // Copy the device state into the host mirror now.
    void synchronize_host(void);
    // Bulk callbacks: false (the default) evaluates them once, at the first
    // simulate_bulk_sources_and_sinks, and keeps the fields on the device,
    // evaluating them again only after update_rates(), after any of the three
    // callback pointers changes, or after a resize; true evaluates them at
    // every call, as the reference does (BioFVM_microenvironment.cpp:595-600),
    // for callbacks that depend on time or on the densities.  Callbacks that
    // leave some entries unwritten (found at the first evaluation), and
    // callbacks set after a first evaluation, are evaluated at every call
    // too: the reference then carries its in-place updated values of those
    // entries from step to step (:603-609), and so does this façade.  With
    // all three callbacks zero_function the bulk step is the identity and is
    // skipped.
    bool bulk_rates_every_step;
    // Host seconds spent evaluating the bulk callbacks and handing their
    // fields to the device, summed over the calls.
    double bulk_rates_host_seconds;
    // Waits until every step requested so far has run on the device.  The
    // bulk and cell phases are enqueued with the next diffusion step (one
    // device step, sources fused into its x sweep) or, before any density is
    // read, on their own.
    void wait_for_device(void);
    // The C-ABI handle of the device copy (null before the first device use),
    // for calls such as bfvm_mass_balance on the same store.
    bfvm_env* device_handle(void);

    // internal: the device handle (created on first use)
    DeviceState* device(void);
    friend void diffusion_decay_solver__constant_coefficients_LOD_3D(Microenvironment& M, double dt);
    friend void diffusion_decay_solver__constant_coefficients_LOD_2D(Microenvironment& M, double dt);

 private:
    std::vector<std::vector<double> > densities;  // host mirror
    std::vector<int> dirichlet_indices;
    std::vector<std::vector<double> > dirichlet_value_vectors;
    std::vector<std::vector<double> > bulk_temp_S, bulk_temp_T, bulk_temp_U;
    std::vector<double> zero, one;
    DeviceState* dev;
    void reset_space(void);
    void refresh_mirror(void);      // mirror current (the device copy stays valid)
    void host_view(void);           // mirror current + mark it the newer copy
    void copy_from(const Microenvironment& other);
    int push_to_device(void);       // upload whatever the device lacks
    int flush_pending(void);        // run the bulk/cell phases not yet enqueued
    int load_bulk_fields(double dt);  // evaluate the callbacks, hand S, T, U to the device
};

class Basic_Agent {
 private:
    Microenvironment* microenvironment;
    std::vector<double> cell_source_sink_solver_temp1;
    std::vector<double> cell_source_sink_solver_temp2;

 protected:
    int current_voxel_index;

 public:
    std::vector<double>* secretion_rates;
    std::vector<double>* saturation_densities;
    std::vector<double>* uptake_rates;
    void set_internal_uptake_constants(double dt);
    void register_microenvironment(Microenvironment*);
    Microenvironment* get_microenvironment(void);

    int ID;
    int index;
    int type;
    double volume;
    bool assign_position(double x, double y, double z);
    bool assign_position(std::vector<double> new_position);
    std::vector<double> position;
    std::vector<double> velocity;
    void update_position(double dt);

    Basic_Agent();
    ~Basic_Agent();
    void simulate_secretion_and_uptake(Microenvironment* M, double dt);
    int nearest_voxel_index(void);
    std::vector<double>& nearest_density_vector(void);

    // internal: monotone counter bumped whenever this agent's voxel or
    // implicit constants change, so the device copy knows to refresh
    unsigned long long revision;
    const std::vector<double>& temp1(void) const { return cell_source_sink_solver_temp1; }
    const std::vector<double>& temp2(void) const { return cell_source_sink_solver_temp2; }
};

extern std::vector<Basic_Agent*> all_basic_agents;
Basic_Agent* create_basic_agent(void);
void delete_basic_agent(int);
void delete_basic_agent(Basic_Agent*);

/* MultiCellDS metadata (BioFVM_MultiCellDS.h:55-77, defaults
 * BioFVM_MultiCellDS.cpp:58-74). */
class MultiCellDS_Metadata {
 public:
    std::string MultiCellDS_type;
    std::string user;
    std::string program;
    std::string program_version;
    std::vector<double> bounding_box;
    std::string spatial_units;
    std::string time_units;
    std::string runtime_time_units;
    double current_time;
    double current_runtime;
    MultiCellDS_Metadata();
};
extern MultiCellDS_Metadata default_metadata;

/* MultiCellDS snapshot: filename_base.xml, the data as
 * filename_base_microenvironment0.mat (default) or as CSV lines, the voxels
 * as XML (default) or filename_base_mesh0.mat (BioFVM_MultiCellDS.cpp:79-83,
 * 394-720).  One document per process: later calls update the time, the run
 * time and the data only. */
void set_save_biovm_mesh_as_matlab(bool newvalue);
void set_save_biovm_data_as_matlab(bool newvalue);
void save_microenvironment_to_MultiCellDS_xml_pugi(std::string filename_base, Microenvironment& M);

void set_default_microenvironment(Microenvironment* M);
Microenvironment* get_default_microenvironment(void);

/* The GPU LOD solver: same signature as BioFVM_solvers.h:38, assignable to
 * Microenvironment::diffusion_decay_solver. */
void diffusion_decay_solver__constant_coefficients_LOD_3D(Microenvironment& M, double dt);
/* The 2-D LOD solver (BioFVM_solvers.h:40, BioFVM_solvers.cpp:278-427) on the
 * GPU, for meshes of one z layer. */
void diffusion_decay_solver__constant_coefficients_LOD_2D(Microenvironment& M, double dt);

void zero_function(Microenvironment* pMicroenvironment, int voxel_index,
                   std::vector<double>* write_destination);
void one_function(Microenvironment* pMicroenvironment, int voxel_index,
                  std::vector<double>* write_destination);

/* Host helpers of the reference's public API used by its example programs:
 * element-wise vector arithmetic and norms (BioFVM_vector.h:57-107) and the
 * stopwatch (BioFVM_utilities.h:56-73). */
std::vector<double> operator-(const std::vector<double>& v1, const std::vector<double>& v2);
std::vector<double> operator+(const std::vector<double>& v1, const std::vector<double>& v2);
std::vector<double> operator*(const std::vector<double>& v1, const std::vector<double>& v2);
std::vector<double> operator/(const std::vector<double>& v1, const std::vector<double>& v2);
std::vector<double> operator*(double d, const std::vector<double>& v1);
std::vector<double> operator+(double d, const std::vector<double>& v1);
std::vector<double> operator+(const std::vector<double>& v1, double d);
std::vector<double> operator-(double d, const std::vector<double>& v1);
std::vector<double> operator-(const std::vector<double>& v1, double d);
void operator+=(std::vector<double>& v1, const std::vector<double>& v2);
void operator-=(std::vector<double>& v1, const std::vector<double>& v2);
void operator/=(std::vector<double>& v1, const std::vector<double>& v2);
void operator*=(std::vector<double>& v1, const double& a);
void operator*=(std::vector<double>& v1, const std::vector<double>& v2);
void operator/=(std::vector<double>& v1, const double& a);
std::ostream& operator<<(std::ostream& os, const std::vector<double>& v);
std::vector<double> normalize(std::vector<double>& v);
void normalize(std::vector<double>* v);
double norm_squared(const std::vector<double>& v);
double norm(const std::vector<double>& v);
double maxabs(const std::vector<double>& v);
double max_abs_difference(const std::vector<double>& v1, const std::vector<double>& v2);
std::vector<double> exponentiate(const std::vector<double>& exponent);
void randomize(std::vector<double>* v);
void axpy(std::vector<double>* y, double& a, std::vector<double>& x);
void axpy(std::vector<double>* y, std::vector<double>& a, std::vector<double>& x);
void naxpy(std::vector<double>* y, double& a, std::vector<double>& x);
void naxpy(std::vector<double>* y, std::vector<double>& a, std::vector<double>& x);
void csv_to_vector(const char* buffer, std::vector<double>& vect);
char* vector_to_csv(const std::vector<double>& vect);
void vector_to_csv_safe(const std::vector<double>& vect, char*& buffer);
void vector_to_csv(const std::vector<double>& vect, char*& buffer);

/* MAT v4 files on the host (BioFVM_matlab.h:57-72): read_matlab returns
 * rows x columns (output[row][column]); write_matlab stores input[i] as
 * column i. */
std::vector<std::vector<double> > read_matlab(std::string filename);
FILE* read_matlab_header(int* rows, int* cols, std::string filename);
FILE* write_matlab_header(int rows, int cols, std::string filename, std::string variable_name);
bool write_matlab(std::vector<std::vector<double> >& input, std::string filename);

void TIC(void);
void TOC(void);
void RUNTIME_TIC(void);
void RUNTIME_TOC(void);
double stopwatch_value(void);
double runtime_stopwatch_value(void);
double total_stopwatch_time(void);
void display_stopwatch_value(std::ostream& os, double dIn);
std::string format_stopwatch_value(double dIn);

}  // namespace BioFVM

#endif
