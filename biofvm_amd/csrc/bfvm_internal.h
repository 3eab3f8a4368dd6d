// Internal state of a biofvm_amd handle and the kernel launchers shared by
// the C-ABI (capi.hip) and the kernel files (sources.hip, lod.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "biofvm_amd.h"

namespace bfvm {

// Thomas factorisation of one axis for all substrates, in device precision.
// Element (i, s) of an axis table is stored at [i*S + s].
struct AxisCoef {
    int n = 0;
    void* denom = nullptr;      // denom_i                  (BioFVM_solvers.cpp:130-141)
    void* inv_denom = nullptr;  // 1/denom_i, rounded once from FP64
    void* c = nullptr;          // c_i = -c1/denom_i        (BioFVM_solvers.cpp:137-141)
    int converged_from = 0;     // first i with (denom_i, c_i) equal to their interior
                                // fixed point for every i in [converged_from, n-1)
};

struct Env {
    int nx = 0, ny = 0, nz = 0, S = 0;
    double dx = 0, dy = 0, dz = 0;
    bool uniform = true;
    int precision = BFVM_FP64;
    size_t nvox = 0, nvs = 0;  // voxels, voxel-substrate values
    size_t elem = 8;           // bytes per device value
    hipStream_t stream = nullptr;
    int device = 0;

    void* rho = nullptr;  // density store, nvs values

    // Coefficients: D, lambda as set by the caller; the factorisation is
    // rebuilt when they or dt change.
    std::vector<double> D, lambda;
    bool coef_valid = false;
    double coef_dt = 0;
    void* c1 = nullptr;  // dt*D/dx^2 per substrate (device precision)
    AxisCoef ax[3];

    // Dirichlet nodes.
    int n_dir = 0;
    int* dir_idx = nullptr;
    void* dir_val = nullptr;          // n_dir*S, device precision
    unsigned char* dir_mask = nullptr;  // S bytes

    // Bulk fields (FP64 on the device) and the implicit-step terms built
    // from them for one dt: A = dt*(T*S), B = (U+S)*dt + 1 (device precision).
    bool have_bulk = false;
    double* bulk_supply = nullptr;
    double* bulk_target = nullptr;
    double* bulk_uptake = nullptr;
    void* bulk_A = nullptr;
    void* bulk_B = nullptr;
    double bulk_dt = 0;
    bool bulk_valid = false;

    // Agents grouped by voxel (list order kept within a voxel).
    int n_agents = 0;
    int n_agent_voxels = 0;
    int* agent_seg_voxel = nullptr;  // n_agent_voxels
    int* agent_seg_start = nullptr;  // n_agent_voxels + 1
    void* agent_t1 = nullptr;        // n_agents*S, grouped order
    void* agent_t2 = nullptr;

    // Per-launch timing (bfvm_timed_run with per-kernel events).
    bool kernel_timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_marks;
    size_t ev_next = 0;

    std::string err;
};

// Launchers (device precision chosen from env.precision).  Each returns the
// hipError_t of the launch.
hipError_t launch_dirichlet(Env& e);
hipError_t launch_bulk_prepare(Env& e, double dt);
hipError_t launch_bulk(Env& e);
hipError_t launch_agents(Env& e);
hipError_t launch_sweep(Env& e, int axis);

}  // namespace bfvm

// The opaque handle type of the C ABI is the internal state.
struct bfvm_env : public bfvm::Env {};
