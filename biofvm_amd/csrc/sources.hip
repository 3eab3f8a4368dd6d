// Source/sink and boundary kernels: Dirichlet overwrite, bulk implicit
// source/sink, agent secretion/uptake.  All are HBM-bound elementwise work.
#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block = kBlock)
{
    return (unsigned)((n + block - 1) / block);
}

// apply_dirichlet_conditions (BioFVM_microenvironment.cpp:192-197): one
// thread per (node, substrate); substrates with mask[s] == 0 are left alone.
template <typename T>
__global__ void k_dirichlet(T* __restrict__ rho, int S, int n_dir, const int* __restrict__ idx,
                            const T* __restrict__ val, const unsigned char* __restrict__ mask)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_dir * S) return;
    int d = (int)(t / S);
    int s = (int)(t - (size_t)d * S);
    if (mask[s]) rho[(size_t)idx[d] * S + s] = val[t];
}

// Implicit bulk terms for one dt (BioFVM_microenvironment.cpp:603-607),
// in FP64 with contraction off so that no FMA changes the reference's
// rounding (written as plain operators: the pragma covers only expressions in
// its own scope, not those inside the _rn helper functions):
//   A = dt*(T*S),  B = ((U+S)*dt) + 1
// A and B stay FP64 for either store precision: rounding them to FP32 would
// bias every step's source term the same way.
__global__ void k_bulk_prepare(size_t nvs, double dt, const double* __restrict__ sup,
                               const double* __restrict__ tgt, const double* __restrict__ upt,
                               double* __restrict__ A, double* __restrict__ B)
{
#pragma clang fp contract(off)  // the reference's separate roundings, no FMA
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvs) return;
    const double st = tgt[e] * sup[e];
    A[e] = dt * st;
    const double us = upt[e] + sup[e];
    B[e] = us * dt + 1.0;
}

// rho = (rho + A) / B  (BioFVM_microenvironment.cpp:604,609), evaluated in
// FP64 and rounded once to the store.  Four values per thread.
template <typename T>
__global__ void k_bulk(T* __restrict__ rho, const double* __restrict__ A,
                       const double* __restrict__ B, size_t nvs)
{
    size_t e = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (e + 4 <= nvs) {
#pragma unroll
        for (int q = 0; q < 4; ++q) rho[e + q] = (T)(((double)rho[e + q] + A[e + q]) / B[e + q]);
    } else {
        for (; e < nvs; ++e) rho[e] = (T)(((double)rho[e] + A[e]) / B[e]);
    }
}

// Basic_Agent::simulate_secretion_and_uptake (BioFVM_basic_agent.cpp:229-238)
// for every agent, grouped by voxel: one thread per (voxel group, substrate)
// applies the group's agents in list order, rho = (rho + t1)/t2, so the
// result is the reference's single-threaded result, with no atomics.  The
// constants are FP64 and the chain runs in FP64 for either store precision.
// SAVE: the value each thread reads first goes to save[t] as well (the
// backward pass keeps a step's start this way while the step runs in place).
// The store's values are read and written with the non-temporal policy: with
// ordinary accesses the x sweep launched next ran 0.80 ms instead of 0.71 at
// 400^3 x 8 FP32 with 1e6 clustered agents (this kernel 0.070 -> 0.084 ms);
// a non-temporal load or store alone left it at 0.80-0.81 ms.
template <typename T, bool SAVE = false>
__global__ void k_agents(T* __restrict__ rho, int S, int n_groups, const int* __restrict__ gvox,
                         const int* __restrict__ gstart, const double* __restrict__ t1,
                         const double* __restrict__ t2, const double* __restrict__ bulk_a,
                         const double* __restrict__ bulk_b, T* __restrict__ save = nullptr)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_groups * S) return;
    int g = (int)(t / S);
    int s = (int)(t - (size_t)g * S);
    const size_t i = (size_t)gvox[g] * S + s;
    T* r = rho + i;
    const T r0 = __builtin_nontemporal_load(r);
    if (SAVE) save[t] = r0;
    double v = (double)r0;
    // the bulk update of this voxel first, as k_bulk would have left it (the
    // x sweep applies it to the voxels without agents)
    if (bulk_a) v = (double)(T)((v + bulk_a[i]) / bulk_b[i]);
    int a1 = gstart[g + 1];
    for (int a = gstart[g]; a < a1; ++a) {
        size_t q = (size_t)a * S + s;
        v = (v + t1[q]) / t2[q];
    }
    __builtin_nontemporal_store((T)v, r);
}

// k_agents<T, true> on rho, after writing prev_save back into prev_rho at the
// same agent voxels (the previous step's start, which its own saving agent
// kernel changed): one launch for the restore of one step and the agents of
// the next.  prev_rho and rho are different stores.
template <typename T>
__global__ void k_agents_restore_then_save(T* __restrict__ rho, T* __restrict__ save, T* __restrict__ prev_rho,
                                           const T* __restrict__ prev_save, int S, int n_groups,
                                           const int* __restrict__ gvox, const int* __restrict__ gstart,
                                           const double* __restrict__ t1, const double* __restrict__ t2)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_groups * S) return;
    const int g = (int)(t / S);
    const int s = (int)(t - (size_t)g * S);
    const size_t i = (size_t)gvox[g] * S + s;
    prev_rho[i] = prev_save[t];
    T* r = rho + i;
    const T r0 = __builtin_nontemporal_load(r);
    save[t] = r0;
    double v = (double)r0;
    const int a1 = gstart[g + 1];
    for (int a = gstart[g]; a < a1; ++a) {
        const size_t q = (size_t)a * S + s;
        v = (v + t1[q]) / t2[q];
    }
    __builtin_nontemporal_store((T)v, r);
}

// Writes back what k_agents<T, true> saved: rho at every agent voxel as it
// was before the agents ran.
template <typename T>
__global__ void k_agents_restore(T* __restrict__ rho, int S, int n_groups, const int* __restrict__ gvox,
                                 const T* __restrict__ save)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_groups * S) return;
    const int g = (int)(t / S);
    const int s = (int)(t - (size_t)g * S);
    rho[(size_t)gvox[g] * S + s] = save[t];
}

// k_agents for several grids: block b works for member block_member[b],
// whose threads start on a block boundary (tab[m].first), one (voxel group,
// substrate) per thread as in k_agents.  (A binary search of the member table
// per thread, six dependent loads before any work, made this kernel 2.5x
// slower per item than k_agents.)
template <typename T>
__global__ void __launch_bounds__(kAgentBatchBlock)
k_agents_batch(const AgentBatch* __restrict__ tab, const int* __restrict__ block_member)
{
    const int mi = __builtin_amdgcn_readfirstlane(block_member[blockIdx.x]);
    const AgentBatch& m = tab[mi];
    const long long local = (long long)blockIdx.x * kAgentBatchBlock + threadIdx.x - m.first;
    const int S = m.S;
    if (local >= (long long)m.n_groups * S) return;
    const int g = (int)(local / S);
    const int s = (int)(local - (long long)g * S);
    T* r = (T*)m.rho + (size_t)m.gvox[g] * S + s;
    double v = (double)__builtin_nontemporal_load(r);  // the policy of k_agents
    const int a1 = m.gstart[g + 1];
    for (int a = m.gstart[g]; a < a1; ++a) {
        const size_t q = (size_t)a * S + s;
        v = (v + m.t1[q]) / m.t2[q];
    }
    __builtin_nontemporal_store((T)v, r);
}

// Microenvironment::compute_all_gradient_vectors (BioFVM_microenvironment.cpp:699-779):
// per voxel and substrate, for each axis with the voxel strictly inside,
// g = (rho[n+jump] - rho[n-jump]) / (2*d) in FP64; components on an axis's two
// boundary layers are 0 (the value reset_all_gradient_vectors leaves there,
// :826-842).  out[(n*S + s)*3 + axis].  One thread per (voxel, substrate).
template <typename T>
__global__ void k_gradient(const T* __restrict__ rho, T* __restrict__ out, int nx, int ny, int nz, int S,
                           double two_dx, double two_dy, double two_dz)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvs = (size_t)nx * ny * nz * S;
    if (t >= nvs) return;
    const size_t v = t / S;
    const int i = (int)(v % nx);
    const int j = (int)((v / nx) % ny);
    const int k = (int)(v / ((size_t)nx * ny));
    const size_t sx = (size_t)S, sy = (size_t)nx * S, sz = (size_t)nx * ny * S;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (i > 0 && i < nx - 1) gx = ((double)rho[t + sx] - (double)rho[t - sx]) / two_dx;
    if (j > 0 && j < ny - 1) gy = ((double)rho[t + sy] - (double)rho[t - sy]) / two_dy;
    if (k > 0 && k < nz - 1) gz = ((double)rho[t + sz] - (double)rho[t - sz]) / two_dz;
    out[t * 3 + 0] = (T)gx;
    out[t * 3 + 1] = (T)gy;
    out[t * 3 + 2] = (T)gz;
}

// Bit v%32 of word v/32 for every voxel v that holds agents.
__global__ void k_agent_bits(const int* __restrict__ gvox, int ng, unsigned* __restrict__ bits)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int v = gvox[g];
    atomicOr(bits + (v >> 5), 1u << (v & 31));
}

}  // namespace

hipError_t build_agent_bits(Env& e)
{
    if (e.agent_bits) (void)hipFree(e.agent_bits);
    e.agent_bits = nullptr;
    if (e.n_agent_voxels <= 0) return hipSuccess;
    const size_t words = e.nvox / 32 + 2;
    hipError_t st = hipMalloc((void**)&e.agent_bits, sizeof(unsigned) * words);
    if (st == hipSuccess) st = hipMemsetAsync(e.agent_bits, 0, sizeof(unsigned) * words, e.stream);
    if (st != hipSuccess) return st;
    k_agent_bits<<<grid_for((size_t)e.n_agent_voxels), kBlock, 0, e.stream>>>(e.agent_seg_voxel,
                                                                             e.n_agent_voxels, e.agent_bits);
    return hipGetLastError();
}

hipError_t launch_gradient(Env& e, void* out)
{
    const size_t n = e.nvs;
    // 2*d as the reference forms it (d *= 2.0, :701-710)
    const double tx = e.dx * 2.0, ty = e.dy * 2.0, tz = e.dz * 2.0;
    if (e.precision == BFVM_FP64)
        k_gradient<double><<<grid_for(n), kBlock, 0, e.stream>>>((const double*)e.rho, (double*)out, e.nx,
                                                                 e.ny, e.nz, e.S, tx, ty, tz);
    else
        k_gradient<float><<<grid_for(n), kBlock, 0, e.stream>>>((const float*)e.rho, (float*)out, e.nx, e.ny,
                                                                e.nz, e.S, tx, ty, tz);
    return hipGetLastError();
}

hipError_t launch_agents_batch(int precision, const AgentBatch* d_tab, const int* d_block_member, int blocks,
                               hipStream_t st)
{
    if (blocks == 0) return hipSuccess;
    if (precision == BFVM_FP64)
        k_agents_batch<double><<<blocks, kAgentBatchBlock, 0, st>>>(d_tab, d_block_member);
    else
        k_agents_batch<float><<<blocks, kAgentBatchBlock, 0, st>>>(d_tab, d_block_member);
    return hipGetLastError();
}

hipError_t launch_dirichlet(Env& e)
{
    if (e.n_dir == 0) return hipSuccess;
    size_t n = (size_t)e.n_dir * e.S;
    if (e.precision == BFVM_FP64)
        k_dirichlet<double><<<grid_for(n), kBlock, 0, e.stream>>>(
            (double*)e.rho, e.S, e.n_dir, e.dir_idx, (const double*)e.dir_val, e.dir_mask);
    else
        k_dirichlet<float><<<grid_for(n), kBlock, 0, e.stream>>>(
            (float*)e.rho, e.S, e.n_dir, e.dir_idx, (const float*)e.dir_val, e.dir_mask);
    return hipGetLastError();
}

hipError_t launch_bulk_prepare(Env& e, double dt)
{
    k_bulk_prepare<<<grid_for(e.nvs), kBlock, 0, e.stream>>>(e.nvs, dt, e.bulk_supply, e.bulk_target,
                                                          e.bulk_uptake, (double*)e.bulk_A,
                                                          (double*)e.bulk_B);
    return hipGetLastError();
}

hipError_t launch_bulk(Env& e)
{
    unsigned grid = grid_for((e.nvs + 3) / 4);
    if (e.precision == BFVM_FP64)
        k_bulk<double><<<grid, kBlock, 0, e.stream>>>((double*)e.rho, (const double*)e.bulk_A,
                                                      (const double*)e.bulk_B, e.nvs);
    else
        k_bulk<float><<<grid, kBlock, 0, e.stream>>>((float*)e.rho, (const double*)e.bulk_A,
                                                     (const double*)e.bulk_B, e.nvs);
    return hipGetLastError();
}

hipError_t launch_agents(Env& e, bool with_bulk)
{
    if (e.n_agent_voxels == 0) return hipSuccess;
    size_t n = (size_t)e.n_agent_voxels * e.S;
    const double* ba = with_bulk ? (const double*)e.bulk_A : nullptr;
    const double* bb = with_bulk ? (const double*)e.bulk_B : nullptr;
    if (e.precision == BFVM_FP64)
        k_agents<double><<<grid_for(n), kBlock, 0, e.stream>>>(
            (double*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, e.agent_seg_start,
            (const double*)e.agent_t1, (const double*)e.agent_t2, ba, bb);
    else
        k_agents<float><<<grid_for(n), kBlock, 0, e.stream>>>(
            (float*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, e.agent_seg_start,
            (const double*)e.agent_t1, (const double*)e.agent_t2, ba, bb);
    return hipGetLastError();
}

hipError_t launch_agents_saving(Env& e, void* save)
{
    if (e.n_agent_voxels == 0) return hipSuccess;
    const size_t n = (size_t)e.n_agent_voxels * e.S;
    if (e.precision == BFVM_FP64)
        k_agents<double, true><<<grid_for(n), kBlock, 0, e.stream>>>(
            (double*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, e.agent_seg_start,
            (const double*)e.agent_t1, (const double*)e.agent_t2, nullptr, nullptr, (double*)save);
    else
        k_agents<float, true><<<grid_for(n), kBlock, 0, e.stream>>>(
            (float*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, e.agent_seg_start,
            (const double*)e.agent_t1, (const double*)e.agent_t2, nullptr, nullptr, (float*)save);
    return hipGetLastError();
}

hipError_t launch_agents_restore_then_save(Env& e, void* save, void* prev_rho, const void* prev_save)
{
    if (e.n_agent_voxels == 0) return hipSuccess;
    const size_t n = (size_t)e.n_agent_voxels * e.S;
    if (e.precision == BFVM_FP64)
        k_agents_restore_then_save<double><<<grid_for(n), kBlock, 0, e.stream>>>(
            (double*)e.rho, (double*)save, (double*)prev_rho, (const double*)prev_save, e.S, e.n_agent_voxels,
            e.agent_seg_voxel, e.agent_seg_start, (const double*)e.agent_t1, (const double*)e.agent_t2);
    else
        k_agents_restore_then_save<float><<<grid_for(n), kBlock, 0, e.stream>>>(
            (float*)e.rho, (float*)save, (float*)prev_rho, (const float*)prev_save, e.S, e.n_agent_voxels,
            e.agent_seg_voxel, e.agent_seg_start, (const double*)e.agent_t1, (const double*)e.agent_t2);
    return hipGetLastError();
}

hipError_t launch_agents_restore(Env& e, const void* save)
{
    if (e.n_agent_voxels == 0) return hipSuccess;
    const size_t n = (size_t)e.n_agent_voxels * e.S;
    if (e.precision == BFVM_FP64)
        k_agents_restore<double><<<grid_for(n), kBlock, 0, e.stream>>>(
            (double*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, (const double*)save);
    else
        k_agents_restore<float><<<grid_for(n), kBlock, 0, e.stream>>>(
            (float*)e.rho, e.S, e.n_agent_voxels, e.agent_seg_voxel, (const float*)save);
    return hipGetLastError();
}

}  // namespace bfvm
