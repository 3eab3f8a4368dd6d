// Agents set up on the device from device arrays: nearest voxel
// (Cartesian_Mesh::nearest_voxel_index, BioFVM_mesh.cpp:930-948, with the
// domain check of Basic_Agent::assign_position, BioFVM_basic_agent.cpp:66-71),
// the implicit constants (Basic_Agent::set_internal_uptake_constants,
// BioFVM_basic_agent.cpp:114-148) and the grouping by voxel that k_agents
// uses, with a stable radix sort of (voxel, list index) so that list order is
// kept inside a voxel (sort.hip).  The result equals bfvm_set_agent_rates on
// the same values; agents that move are re-binned by calling it again,
// without a host round trip of their arrays.
#include <algorithm>
#include <string>
#include <vector>

#include "bfvm_internal.h"

using bfvm::Env;

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n)
{
    return (unsigned)((n + kBlock - 1) / kBlock);
}

int afail(Env& e, int code, const std::string& msg)
{
    e.err = msg;
    return code;
}

#define GHIP(e, call)                                                                             \
    do {                                                                                          \
        hipError_t st_ = (call);                                                                  \
        if (st_ != hipSuccess) return afail(e, BFVM_E_DEVICE, std::string(#call) + ": " +         \
                                                                   bfvm::hip_error_text(st_));       \
    } while (0)

// voxel of each agent, key = voxel, value = list index; bad[0] counts agents
// outside [0, n*d] on some axis (assign_position refuses them)
__global__ void k_agent_voxels(const double* __restrict__ pos, int n, int nx, int ny, int nz, double dx,
                               double dy, double dz, unsigned* __restrict__ key, int* __restrict__ idx,
                               int* __restrict__ bad)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int cnt[3] = {nx, ny, nz};
    const double d[3] = {dx, dy, dz};
    int ijk[3];
    bool ok = true;
    for (int q = 0; q < 3; ++q) {
        const double p = pos[3 * a + q];
        const double ext = (double)cnt[q] * d[q];
        if (!(p >= 0.0 && p <= ext)) ok = false;
        int v = (int)floor(p / d[q]);  // BioFVM_mesh.cpp:932-946, origin 0
        ijk[q] = min(max(v, 0), cnt[q] - 1);
    }
    if (!ok) atomicAdd(bad, 1);
    key[a] = (unsigned)((ijk[2] * ny + ijk[1]) * nx + ijk[0]);
    idx[a] = a;
}

// group heads of the sorted keys
__global__ void k_group_heads(const unsigned* __restrict__ key, int n, int* __restrict__ head)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    head[q] = (q == 0 || key[q] != key[q - 1]) ? 1 : 0;
}

// gvox[g], gstart[g] for every head (g = exclusive sum of heads), gstart[ng] = n
__global__ void k_group_tables(const unsigned* __restrict__ key, const int* __restrict__ head,
                               const int* __restrict__ gid, int n, int* __restrict__ gvox,
                               int* __restrict__ gstart, int* __restrict__ ng)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (head[q]) {
        gvox[gid[q]] = (int)key[q];
        gstart[gid[q]] = q;
    }
    if (q == n - 1) {
        const int groups = gid[q] + head[q];
        gstart[groups] = n;
        *ng = groups;
    }
}

// constants of grouped agent q (list index order[q]), FP64 with the
// reference's operation order and no contraction (BioFVM_basic_agent.cpp:122-147)
__global__ void k_agent_constants(const int* __restrict__ order, int n, int S, double dt, double dV,
                                  const double* __restrict__ vol, const double* __restrict__ sec,
                                  const double* __restrict__ sat, const double* __restrict__ upt,
                                  double* __restrict__ t1, double* __restrict__ t2,
                                  double* __restrict__ cg, double* __restrict__ satg)
{
#pragma clang fp contract(off)  // the reference's separate roundings, no FMA
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int a = order[q];
    double c = dt * vol[a];
    c = c / dV;
    cg[q] = c;
    for (int s = 0; s < S; ++s) {
        const size_t i = (size_t)a * S + s, o = (size_t)q * S + s;
        double x = 0.0;
        x += sec[i];
        x *= sat[i];
        x *= c;
        double y = 1.0;
        const double cs = c * sec[i];
        y += cs;
        const double cu = c * upt[i];
        y += cu;
        t1[o] = x;
        t2[o] = y;
        satg[o] = sat[i];
    }
}

template <typename P>
void release(P*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

}  // namespace

extern "C" {

int bfvm_set_agent_rates_device(bfvm_handle h, int n, const double* position, const double* volume,
                                const double* secretion, const double* saturation, const double* uptake,
                                double dt, int* out_voxel)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    if (n < 0) return afail(e, BFVM_E_ARG, "n < 0");
    if (n > 0 && (!position || !volume || !secretion || !saturation || !uptake))
        return afail(e, BFVM_E_ARG, "agent arrays must be non-NULL");
    if (e.ensemble_stream) GHIP(e, hipStreamSynchronize(e.ensemble_stream));
    GHIP(e, hipStreamSynchronize(e.stream));
    // drop the current agents (the arrays are rebuilt)
    release(e.agent_seg_voxel);
    release(e.agent_seg_start);
    release(e.agent_bits);
    release(e.agent_t1);
    release(e.agent_t2);
    release(e.agent_c);
    release(e.agent_sat);
    e.agent_rates = false;
    e.agent_order.clear();
    e.n_agents = 0;
    e.n_agent_voxels = 0;
    ++e.agent_epoch;
    if (n == 0) return BFVM_OK;
    const int S = e.S;
    const size_t ns = (size_t)n * S;
    void *key = nullptr, *key_s = nullptr, *idx = nullptr, *idx_s = nullptr, *head = nullptr, *gid = nullptr,
         *bad = nullptr, *ng = nullptr, *tmp = nullptr;
    int rc = BFVM_OK;
    auto cleanup = [&] {
        for (void** p : {&key, &key_s, &idx, &idx_s, &head, &gid, &bad, &ng, &tmp}) release(*p);
    };
    hipError_t st = hipSuccess;
    auto chk = [&](hipError_t x) {
        if (st == hipSuccess) st = x;
        return st == hipSuccess;
    };
    chk(hipMalloc(&key, sizeof(unsigned) * n));
    chk(hipMalloc(&key_s, sizeof(unsigned) * n));
    chk(hipMalloc(&idx, sizeof(int) * n));
    chk(hipMalloc(&idx_s, sizeof(int) * n));
    chk(hipMalloc(&head, sizeof(int) * n));
    chk(hipMalloc(&gid, sizeof(int) * n));
    chk(hipMalloc(&bad, sizeof(int)));
    chk(hipMalloc(&ng, sizeof(int)));
    chk(hipMalloc((void**)&e.agent_seg_voxel, sizeof(int) * n));
    chk(hipMalloc((void**)&e.agent_seg_start, sizeof(int) * (n + 1)));
    chk(hipMalloc(&e.agent_t1, sizeof(double) * ns));
    chk(hipMalloc(&e.agent_t2, sizeof(double) * ns));
    chk(hipMalloc((void**)&e.agent_c, sizeof(double) * n));
    chk(hipMalloc((void**)&e.agent_sat, sizeof(double) * ns));
    if (st != hipSuccess) {
        cleanup();
        return afail(e, BFVM_E_DEVICE, std::string("agent arrays: ") + bfvm::hip_error_text(st));
    }
    // key bits: enough for the largest voxel index
    int bits = 1;
    while (bits < 32 && ((size_t)1 << bits) < e.nvox) ++bits;
    const size_t sort_bytes = bfvm::radix_sort_tmp_bytes(n), scan_bytes = bfvm::exclusive_scan_tmp_bytes(n);
    chk(hipMalloc(&tmp, std::max(sort_bytes, scan_bytes)));
    chk(hipMemsetAsync(bad, 0, sizeof(int), e.stream));
    if (st == hipSuccess) {
        k_agent_voxels<<<grid_for(n), kBlock, 0, e.stream>>>(position, n, e.nx, e.ny, e.nz, e.dx, e.dy, e.dz,
                                                            (unsigned*)key, (int*)idx, (int*)bad);
        chk(hipGetLastError());
    }
    if (st == hipSuccess)
        chk(bfvm::radix_sort_pairs((const unsigned*)key, (const int*)idx, (unsigned*)key_s, (int*)idx_s, n, bits,
                                   tmp, sort_bytes, e.stream));
    if (st == hipSuccess) {
        k_group_heads<<<grid_for(n), kBlock, 0, e.stream>>>((const unsigned*)key_s, n, (int*)head);
        chk(hipGetLastError());
    }
    if (st == hipSuccess) chk(bfvm::exclusive_scan((const int*)head, (int*)gid, n, tmp, scan_bytes, e.stream));
    if (st == hipSuccess) {
        k_group_tables<<<grid_for(n), kBlock, 0, e.stream>>>((const unsigned*)key_s, (const int*)head,
                                                            (const int*)gid, n, e.agent_seg_voxel,
                                                            e.agent_seg_start, (int*)ng);
        chk(hipGetLastError());
    }
    const double dV = e.dx * e.dy * e.dz;  // BioFVM_mesh.cpp:839
    if (st == hipSuccess) {
        k_agent_constants<<<grid_for(n), kBlock, 0, e.stream>>>((const int*)idx_s, n, S, dt, dV, volume,
                                                               secretion, saturation, uptake,
                                                               (double*)e.agent_t1, (double*)e.agent_t2,
                                                               e.agent_c, e.agent_sat);
        chk(hipGetLastError());
    }
    int nbad = 0, groups = 0;
    std::vector<int> order(n);
    chk(hipMemcpyAsync(&nbad, bad, sizeof(int), hipMemcpyDeviceToHost, e.stream));
    chk(hipMemcpyAsync(&groups, ng, sizeof(int), hipMemcpyDeviceToHost, e.stream));
    chk(hipMemcpyAsync(order.data(), idx_s, sizeof(int) * n, hipMemcpyDeviceToHost, e.stream));
    if (out_voxel) chk(hipMemcpyAsync(out_voxel, key, sizeof(int) * n, hipMemcpyDeviceToDevice, e.stream));
    chk(hipStreamSynchronize(e.stream));
    cleanup();
    if (st != hipSuccess) rc = afail(e, BFVM_E_DEVICE, std::string("agent setup: ") + bfvm::hip_error_text(st));
    else if (nbad > 0) rc = afail(e, BFVM_E_ARG, std::to_string(nbad) + " agent(s) outside the domain");
    if (rc != BFVM_OK) {
        release(e.agent_seg_voxel);
        release(e.agent_seg_start);
        release(e.agent_t1);
        release(e.agent_t2);
        release(e.agent_c);
        release(e.agent_sat);
        return rc;
    }
    e.n_agents = n;
    e.n_agent_voxels = groups;
    e.agent_order = std::move(order);
    e.agent_rates = true;
    e.agent_dt = dt;
    GHIP(e, bfvm::build_agent_bits(e));
    return BFVM_OK;
}

}  // extern "C"
