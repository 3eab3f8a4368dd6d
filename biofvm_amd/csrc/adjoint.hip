// Differentiation of the time-stepped model: a hand-written adjoint of the
// step (bulk -> cells -> diffusion) with segment checkpointing.
//
// The step's operators and their transposes:
//   bulk     y = (x + A) / B                      x' = y' / B
//   agents   per agent in list order, per voxel:  y = (x + t1) / t2
//            x' = y' / t2, t1' = y' / t2, t2' = -y' y / t2, taken in
//            reverse list order with the intermediate values recomputed;
//            t1 = c S T, t2 = 1 + c S + c U (c = dt V / V_voxel) give the
//            rate gradients S' += c (T t1' + t2'), U' += c t2'
//   Dirichlet y = x except at the nodes (masked substrates)   x' = y' there zeroed
//   sweep    y = A^-1 x, A = (1 + c2) I + c1 L (L the line's Neumann Laplacian,
//            BioFVM_solvers.cpp:104-142) is symmetric, so x' = A^-1 y' — the
//            forward sweep kernels run on the adjoint unchanged — and
//            c1' -= sum_edges (x'_i - x'_j)(y_i - y_j),  c2' -= sum x' . y
//            with c1 = dt D / dx^2 and c2 = dt lambda / 3 per substrate.
// The loss seed dJ/drho_N comes from the caller or from an L2 misfit.
//
// Tape: the state at every K-th step is kept (checkpoints); the backward
// pass walks the segments from the last, recomputes the segment's step
// starts from its checkpoint into a ring of K states, then for each step in
// reverse recomputes the step once more keeping the states the transposes
// need (after bulk, after each sweep) and applies the transposed step.
// Device memory: (steps/K + K + 4) stores, an FP64 adjoint state, FP64 gradients.
// Reductions (gradients of D, lambda, the loss) are FP64 and deterministic:
// fixed-size partial sums per block, then one block adds them in order.
// The adjoint state is FP64 for either store precision: the adjoint of a
// rough misfit decays fast under the transposed sweeps, and an FP32 copy
// keeps the rounding of its early, large values (an FP32 adjoint state
// moved the lambda gradient by 5e-4 relative after 500 steps at 128^3).  The
// recomputed forward states stay in the store's precision.
#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "bfvm_internal.h"

using bfvm::Env;

namespace bfvm {

struct Tape {
    double dt = 0;
    int steps = 0, flags = 0, K = 0;
    std::vector<double> D, lambda;  // coefficients the tape was recorded with
    unsigned agent_epoch = 0, source_epoch = 0;
    std::vector<void*> cps;   // state at steps 0, K, 2K, ...
    std::vector<void*> ring;  // starts of the steps of one segment (in ring_slab)
    void* ring_slab = nullptr;
    void* s1 = nullptr;       // after bulk (the agents' input)
    const void* after_bulk = nullptr;  // s1, or the step's start when no bulk update ran
    void* s4 = nullptr;       // after the x sweep
    void* s6 = nullptr;       // after the y sweep
    void* s8 = nullptr;       // after the z sweep (the step's result)
    // the agent voxels' values of a step's start (n_agent_voxels x S, store
    // precision) while the agents run in place on it (forward_step_keep):
    // two halves of asave_bytes / 2, for a step and the one after it
    void* asave = nullptr;
    size_t asave_bytes = 0;
    double* lam = nullptr;    // adjoint state (FP64)
    bool seeded = false;
    double* rk = nullptr;      // agents: intermediate densities (n_agents x S)
    double* gsec = nullptr;    // agents: secretion gradient (grouped order)
    double* gupt = nullptr;    // agents: uptake gradient (grouped order)
    size_t agents_n = 0;       // values of rk, gsec, gupt
    double* gc = nullptr;      // 2*S: dJ/dc1, dJ/dc2
    double* partial = nullptr; // reduction partials
    // the D / lambda sums' block partials, per chunk of kChunk substrates,
    // added up over the whole backward pass and reduced once at its end
    double* gpart = nullptr;
    size_t gpart_n = 0;
    // per axis, the tile partials of the adjoint sweep with the sums fused in
    // (launch_adjoint_sums; null: that axis runs the sweep and the sums apart)
    double* spart[3] = {nullptr, nullptr, nullptr};
    size_t spart_n[3] = {0, 0, 0};
    size_t stiles[3] = {0, 0, 0};
    // when device memory allows: the states each ring slot's step keeps (s1
    // if the step has a bulk update, s4, s6; s8 is the next slot, or
    // last_s8 for a segment's last step), filled by one pass over the
    // segment, so that each step is recomputed once rather than twice
    std::vector<void*> kept;  // in kept_slab
    void* kept_slab = nullptr;
    void* last_s8 = nullptr;
    int nkeep = 0;  // kept states per slot (0: recompute each step twice)
    // HIP graphs of one segment's backward pass with kept states (per
    // segment length), and what their launches have baked in
    struct SegmentGraph {
        hipGraphExec_t exec = nullptr;
        int len = 0;
        std::vector<uintptr_t> key;
    };
    std::vector<SegmentGraph> graphs;
};

}  // namespace bfvm

using bfvm::Tape;

namespace {

constexpr int kBlock = 256;
// blocks of the reduction kernels: fixed, so the sums are deterministic
// (2048 blocks ran the axis sums 30 % slower than 512 at 128^3 x 2)
constexpr int kRedBlocks = 512;
constexpr int kChunk = 8;        // substrates per reduction pass

inline unsigned grid_for(size_t n, int block = kBlock)
{
    return (unsigned)((n + block - 1) / block);
}

int afail(Env& e, int code, const std::string& msg)
{
    e.err = msg;
    return code;
}

#define AHIP(e, call)                                                                            \
    do {                                                                                         \
        hipError_t st_ = (call);                                                                 \
        if (st_ != hipSuccess) return afail(e, BFVM_E_DEVICE, std::string(#call) + ": " +        \
                                                                   bfvm::hip_error_text(st_));      \
    } while (0)

// ---------------------------------------------------------------- kernels

template <typename T>
__global__ void k_dirichlet_zero(T* __restrict__ lam, int S, int n_dir, const int* __restrict__ idx,
                                 const unsigned char* __restrict__ mask)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_dir * S) return;
    int d = (int)(t / S);
    int s = (int)(t - (size_t)d * S);
    if (mask[s]) lam[(size_t)idx[d] * S + s] = (T)0;
}

template <typename T>
__global__ void k_bulk_adjoint(T* __restrict__ lam, const double* __restrict__ B, size_t nvs)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nvs) lam[e] = (T)((double)lam[e] / B[e]);
}

// One thread per (voxel group, substrate): the group's agents recomputed in
// list order from the phase input x (intermediate values kept in rk), then
// walked back in reverse order.
template <typename T>
__global__ void k_agents_adjoint(double* __restrict__ lam, const T* __restrict__ x, int S, int n_groups,
                                 const int* __restrict__ gvox, const int* __restrict__ gstart,
                                 const double* __restrict__ t1, const double* __restrict__ t2,
                                 const double* __restrict__ c, const double* __restrict__ sat,
                                 double* __restrict__ rk, double* __restrict__ gsec,
                                 double* __restrict__ gupt)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_groups * S) return;
    const int g = (int)(t / S);
    const int s = (int)(t - (size_t)g * S);
    const size_t v = (size_t)gvox[g] * S + s;
    const int a0 = gstart[g], a1 = gstart[g + 1];
    double r = (double)x[v];
    for (int a = a0; a < a1; ++a) {
        const size_t q = (size_t)a * S + s;
        r = (r + t1[q]) / t2[q];
        rk[q] = r;
    }
    double l = (double)lam[v];
    for (int a = a1 - 1; a >= a0; --a) {
        const size_t q = (size_t)a * S + s;
        const double inv = 1.0 / t2[q];
        if (c) {  // agents set from rates: accumulate the rate gradients
            const double gt1 = l * inv;
            const double gt2 = -l * rk[q] * inv;
            gsec[q] += c[a] * (sat[q] * gt1 + gt2);
            gupt[q] += c[a] * gt2;
        }
        l *= inv;
    }
    lam[v] = l;
}

// out[m] = (or, with accumulate, +=) sum over the block's threads of
// acc[m], m < M, in a fixed order (wave shuffles, then the waves in index
// order).
template <int M>
__device__ void block_sums(double (&acc)[M], double* __restrict__ out, bool accumulate = false)
{
    __shared__ double red[kBlock / 64][M];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double v = acc[m];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave][m] = v;
    }
    __syncthreads();
    if (threadIdx.x < M) {
        double v = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = accumulate ? out[threadIdx.x] + v : v;
    }
}

// Per block: for substrates s0 .. s0+SC-1 (SC <= kChunk), the sums over the
// block's voxels of (l_v - l_w)(y_v - y_w) over edges v -> w = v + 1 along x,
// and of l_v y_v, added to partial[block][0..kChunk) and [kChunk..2 kChunk).
// Consecutive lanes hold consecutive voxels, so a lane takes its neighbour's
// values from the next lane (the wave's last lane loads them).  kAhead voxels
// of the grid-stride walk at a time: their loads are all issued before the
// first is used, and they are added in walk order (the sums of a
// one-at-a-time walk, bit for bit).
template <typename T, int SC>
__global__ void __launch_bounds__(kBlock) k_axis_grad_x(const double* __restrict__ lam, const T* __restrict__ y,
                                                        int nx, int ny, int nz, int S, int s0,
                                                        double* __restrict__ partial)
{
    // fewer voxels ahead for more substrates: at most about 128 VGPRs (the
    // grid's 512 blocks put 8 waves on a compute unit)
    constexpr int kAhead = SC <= 2 ? 4 : SC <= 4 ? 2 : 1;
    double acc[2 * SC];
#pragma unroll
    for (int m = 0; m < 2 * SC; ++m) acc[m] = 0.0;
    const unsigned nxu = (unsigned)nx;
    const unsigned nvox = nxu * (unsigned)ny * (unsigned)nz;
    const unsigned step = gridDim.x * blockDim.x;
    const bool last_lane = (threadIdx.x & 63) == 63;
    for (unsigned v0 = blockIdx.x * blockDim.x + threadIdx.x; v0 - threadIdx.x % 64 < nvox; v0 += kAhead * step) {
        double l[kAhead][SC], yv[kAhead][SC], ln[kAhead][SC], yn[kAhead][SC];
#pragma unroll
        for (int k = 0; k < kAhead; ++k) {
            const unsigned v = v0 + (unsigned)k * step;
            const unsigned vc = v < nvox ? v : nvox - 1;  // lanes past the end load a real value, add nothing
#pragma unroll
            for (int m = 0; m < SC; ++m) {
                const size_t q = (size_t)vc * S + s0 + m;
                l[k][m] = lam[q];
                yv[k][m] = (double)y[q];
                // the wave's last lane loads its neighbour (v + 1 < nvox when it has an edge)
                const size_t p = last_lane && vc + 1 < nvox ? q + S : q;
                ln[k][m] = last_lane ? lam[p] : 0.0;
                yn[k][m] = last_lane ? (double)y[p] : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < kAhead; ++k) {
            const unsigned v = v0 + (unsigned)k * step;
#pragma unroll
            for (int m = 0; m < SC; ++m) {
                const double lnext = __shfl_down(l[k][m], 1, 64), ynext = __shfl_down(yv[k][m], 1, 64);
                if (!last_lane) {
                    ln[k][m] = lnext;
                    yn[k][m] = ynext;
                }
            }
            if (v >= nvox) continue;  // (whole waves past the end left the loop)
            const bool edge = v % nxu + 1 < nxu;
#pragma unroll
            for (int m = 0; m < SC; ++m) {
                acc[SC + m] += l[k][m] * yv[k][m];
                if (edge) acc[m] += (l[k][m] - ln[k][m]) * (yv[k][m] - yn[k][m]);
            }
        }
    }
    double out[2 * kChunk];
#pragma unroll
    for (int m = 0; m < kChunk; ++m) {
        out[m] = m < SC ? acc[m] : 0.0;
        out[kChunk + m] = m < SC ? acc[SC + m] : 0.0;
    }
    block_sums<2 * kChunk>(out, partial + (size_t)blockIdx.x * 2 * kChunk, true);
}

// The same sums for the y or z axis, whose neighbour is nx or nx*ny voxels
// away: a work item is a run of kRun consecutive nodes of one line (all sc
// substrates), whose values are loaded once (plus the next run's first node
// for the last edge), all loads of a substrate in flight together; items are
// (run, column) with the column fastest, so a wave's loads are contiguous.
constexpr int kRun = 8;

template <typename T, int SC>
__global__ void __launch_bounds__(kBlock) k_axis_grad_runs(const double* __restrict__ lam,
                                                           const T* __restrict__ y, int nx, int ny, int nz,
                                                           int S, int axis, int s0,
                                                           double* __restrict__ partial)
{
    // work items of the grid-stride walk loaded together (all their loads
    // issued before the first is used, then added in walk order: the sums of
    // a one-at-a-time walk, bit for bit); fewer for more substrates (at most
    // about 128 VGPRs: the grid's 512 blocks put 8 waves on a compute unit)
    constexpr int kAheadRuns = SC <= 2 ? 2 : 1;
    double acc[2 * SC];
#pragma unroll
    for (int m = 0; m < 2 * SC; ++m) acc[m] = 0.0;
    const unsigned nxu = (unsigned)nx, nxy = (unsigned)nx * (unsigned)ny;
    const unsigned n = (unsigned)(axis == 1 ? ny : nz);
    const unsigned stride = axis == 1 ? nxu : nxy;            // voxels between nodes of a line
    const unsigned ncol = axis == 1 ? nxu * (unsigned)nz : nxy;  // lines
    const unsigned nrun = (n + kRun - 1) / kRun;
    const unsigned items = ncol * nrun;
    const unsigned step = gridDim.x * blockDim.x;
    for (unsigned it0 = blockIdx.x * blockDim.x + threadIdx.x; it0 < items; it0 += kAheadRuns * step) {
        // with up to 2 substrates, the values of kAheadRuns items at once; with
        // more, one substrate's values at a time (as many registers)
        constexpr int kM = SC <= 2 ? SC : 1;
        double l[kAheadRuns][kM][kRun + 1], yv[kAheadRuns][kM][kRun + 1];
        unsigned base[kAheadRuns], p0[kAheadRuns];
        int cnt[kAheadRuns];
        bool more[kAheadRuns];
#pragma unroll
        for (int k = 0; k < kAheadRuns; ++k) {
            const unsigned it = it0 + (unsigned)k * step;
            const bool live = it < items;
            const unsigned run = live ? it / ncol : 0, c = live ? it - run * ncol : 0;
            // line c's node 0: z lines start in plane 0; y line (i, k) at (k, 0, i)
            base[k] = axis == 1 ? (c / nxu) * nxy + c % nxu : c;
            p0[k] = run * kRun;
            cnt[k] = live ? (int)min((unsigned)kRun, n - p0[k]) : 0;  // nodes of this run
            more[k] = live && p0[k] + kRun < n;                        // the next run's node 0 exists
        }
        auto load = [&](int k, int mi, int m) {
#pragma unroll
            for (int u = 0; u <= kRun; ++u) {
                const bool in = u < cnt[k] || (u == kRun && more[k]);
                const size_t q = (size_t)(base[k] + (p0[k] + (unsigned)u) * stride) * S + s0 + m;
                l[k][mi][u] = in ? lam[q] : 0.0;
                yv[k][mi][u] = in ? (double)y[q] : 0.0;
            }
        };
        auto add = [&](int k, int mi, int m) {
#pragma unroll
            for (int u = 0; u < kRun; ++u) {
                if (u < cnt[k]) {
                    acc[SC + m] += l[k][mi][u] * yv[k][mi][u];
                    if (u + 1 < cnt[k] || (u + 1 == kRun && more[k]))
                        acc[m] += (l[k][mi][u] - l[k][mi][u + 1]) * (yv[k][mi][u] - yv[k][mi][u + 1]);
                }
            }
        };
        if constexpr (SC <= 2) {
#pragma unroll
            for (int k = 0; k < kAheadRuns; ++k)
#pragma unroll
                for (int m = 0; m < SC; ++m) load(k, m, m);
#pragma unroll
            for (int k = 0; k < kAheadRuns; ++k)
#pragma unroll
                for (int m = 0; m < SC; ++m) add(k, m, m);
        } else {
#pragma unroll
            for (int m = 0; m < SC; ++m) {
                load(0, 0, m);
                add(0, 0, m);
            }
        }
    }
    double out[2 * kChunk];
#pragma unroll
    for (int m = 0; m < kChunk; ++m) {
        out[m] = m < SC ? acc[m] : 0.0;
        out[kChunk + m] = m < SC ? acc[SC + m] : 0.0;
    }
    block_sums<2 * kChunk>(out, partial + (size_t)blockIdx.x * 2 * kChunk, true);
}

// gc[s] -= sum of the blocks' edge sums, gc[S + s] -= sum of the l.y sums.
__global__ void __launch_bounds__(kBlock) k_axis_grad_finish(const double* __restrict__ partial, int nblocks,
                                                              int S, int s0, int sc, double* __restrict__ gc)
{
    // thread t adds blocks t, t + blockDim, ... in order, then a fixed tree in
    // LDS: the same order on every run (deterministic), log-depth (a serial
    // loop over the blocks waited for one load after another)
    __shared__ double red[kBlock][2 * kChunk + 1];
    const int t = threadIdx.x;
    double v[2 * kChunk];
#pragma unroll
    for (int m = 0; m < 2 * kChunk; ++m) v[m] = 0.0;
    for (int b = t; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int m = 0; m < 2 * kChunk; ++m) v[m] += partial[(size_t)b * 2 * kChunk + m];
#pragma unroll
    for (int m = 0; m < 2 * kChunk; ++m) red[t][m] = v[m];
    __syncthreads();
    for (int half = blockDim.x / 2; half > 0; half >>= 1) {
        if (t < half)
#pragma unroll
            for (int m = 0; m < 2 * kChunk; ++m) red[t][m] += red[t + half][m];
        __syncthreads();
    }
    if (t >= 2 * kChunk) return;
    const int s = t % kChunk;
    if (s >= sc) return;
    if (t < kChunk) gc[s0 + s] -= red[0][t];
    else gc[S + s0 + s] -= red[0][t];
}

// lam = rho - target; partial sums of (rho - target)^2 / 2.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_l2_seed(const T* __restrict__ rho, const double* __restrict__ target,
                                                    double* __restrict__ lam, size_t nvs,
                                                    double* __restrict__ partial)
{
    double acc[1] = {0.0};
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nvs;
         e += (size_t)gridDim.x * blockDim.x) {
        const double d = (double)rho[e] - target[e];
        lam[e] = d;
        acc[0] += 0.5 * d * d;
    }
    block_sums<1>(acc, partial + blockIdx.x);
}

__global__ void k_sum(const double* __restrict__ partial, int n, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    double v = 0.0;
    for (int b = 0; b < n; ++b) v += partial[b];
    *out = v;
}

// ---------------------------------------------------------------- host side

hipError_t dirichlet_zero(Env& e, double* lam)
{
    if (e.n_dir == 0) return hipSuccess;
    k_dirichlet_zero<double><<<grid_for((size_t)e.n_dir * e.S), kBlock, 0, e.stream>>>(
        lam, e.S, e.n_dir, e.dir_idx, e.dir_mask);
    return hipGetLastError();
}

// The axis's D / lambda sums, added to the block partials of tp.gpart (one
// kRedBlocks x 2 kChunk slab per chunk of substrates); axis_grad_finish
// reduces them into gc once the backward pass is done.
template <typename T>
hipError_t axis_grad(Env& e, Tape& tp, const void* y, int axis)
{
    for (int s0 = 0; s0 < e.S; s0 += kChunk) {
        const int sc = std::min(kChunk, e.S - s0);
        double* part = tp.gpart + (size_t)(s0 / kChunk) * kRedBlocks * 2 * kChunk;
        // the neighbour is the next voxel for x (the next lane's), nx or nx*ny
        // voxels away for y and z (runs of nodes along the line)
        auto launch = [&](auto sc_const) {
            constexpr int SC = decltype(sc_const)::value;
            if (axis == 0)
                k_axis_grad_x<T, SC><<<kRedBlocks, kBlock, 0, e.stream>>>(tp.lam, (const T*)y, e.nx, e.ny, e.nz,
                                                                          e.S, s0, part);
            else
                k_axis_grad_runs<T, SC><<<kRedBlocks, kBlock, 0, e.stream>>>(tp.lam, (const T*)y, e.nx, e.ny,
                                                                             e.nz, e.S, axis, s0, part);
        };
        switch (sc) {
        case 1: launch(std::integral_constant<int, 1>{}); break;
        case 2: launch(std::integral_constant<int, 2>{}); break;
        case 3: launch(std::integral_constant<int, 3>{}); break;
        case 4: launch(std::integral_constant<int, 4>{}); break;
        case 5: launch(std::integral_constant<int, 5>{}); break;
        case 6: launch(std::integral_constant<int, 6>{}); break;
        case 7: launch(std::integral_constant<int, 7>{}); break;
        default: launch(std::integral_constant<int, 8>{}); break;
        }
        hipError_t st = hipGetLastError();
        if (st != hipSuccess) return st;
    }
    return hipSuccess;
}

hipError_t axis_grad_finish(Env& e, Tape& tp)
{
    static_assert(kChunk == bfvm::kAdjointSumsChunk, "one partial layout for both sums");
    for (int s0 = 0; s0 < e.S; s0 += kChunk) {
        const int sc = std::min(kChunk, e.S - s0);
        const double* part = tp.gpart + (size_t)(s0 / kChunk) * kRedBlocks * 2 * kChunk;
        k_axis_grad_finish<<<1, kBlock, 0, e.stream>>>(part, kRedBlocks, e.S, s0, sc, tp.gc);
        hipError_t st = hipGetLastError();
        if (st != hipSuccess) return st;
        // the tile partials of the axes whose sums ran inside their sweeps
        for (int axis = 0; axis < 3; ++axis) {
            if (!tp.spart[axis]) continue;
            const double* sp = tp.spart[axis] + (size_t)(s0 / kChunk) * tp.stiles[axis] * 2 * kChunk;
            k_axis_grad_finish<<<1, kBlock, 0, e.stream>>>(sp, (int)tp.stiles[axis], e.S, s0, sc, tp.gc);
            if ((st = hipGetLastError()) != hipSuccess) return st;
        }
    }
    return hipSuccess;
}

// Runs launcher f with the handle's store pointing at `buf`.
template <typename F>
hipError_t on_buffer(Env& e, void* buf, F&& f)
{
    void* keep = e.rho;
    e.rho = buf;
    hipError_t st = f();
    e.rho = keep;
    return st;
}

// The sweep of `axis` on the FP64 adjoint state (the FP64 kernels, whatever
// the store's precision).
hipError_t sweep_adjoint(Env& e, double* lam, int axis)
{
    const int keep = e.precision;
    const size_t keep_elem = e.elem;
    const bfvm::SweepPlan keep_plan = e.plan[axis];
    e.precision = BFVM_FP64;
    e.elem = sizeof(double);  // offsets and 4 GiB limits in FP64 values
    // row tiles tuned for an FP32 store take twice the LDS on the FP64
    // adjoint (22 us against 14 us for the default shape at 128^3 x 2); the
    // whole-cache-line x kernel tuned for an FP32 store is not timed on FP64
    if (keep != BFVM_FP64 && keep_plan.stream >= 2) e.plan[axis] = bfvm::SweepPlan{};
    hipError_t st = on_buffer(e, lam, [&] { return bfvm::launch_sweep(e, axis); });
    e.plan[axis] = keep_plan;
    e.elem = keep_elem;
    e.precision = keep;
    return st;
}

// The sweep of `axis` on the FP64 adjoint state with that axis's D/lambda
// sums over it and the forward output y fused in (tp.spart[axis]).
hipError_t sweep_adjoint_sums(Env& e, Tape& tp, const void* y, int axis)
{
    const int keep = e.precision;
    const size_t keep_elem = e.elem;
    e.precision = BFVM_FP64;
    e.elem = sizeof(double);  // offsets and 4 GiB limits in FP64 values
    const hipError_t st = bfvm::launch_adjoint_sums(e, axis, tp.lam, y, keep, tp.spart[axis]);
    e.elem = keep_elem;
    e.precision = keep;
    return st;
}

hipError_t copy_store(Env& e, void* dst, const void* src)
{
    return hipMemcpyAsync(dst, src, e.nvs * e.elem, hipMemcpyDeviceToDevice, e.stream);
}

// One forward step on buffer `buf` (the phases of bfvm_step).
hipError_t forward_step(Env& e, const Tape& tp, void* buf)
{
    return on_buffer(e, buf, [&]() -> hipError_t {
        hipError_t st;
        if ((tp.flags & BFVM_BULK) && e.have_bulk && (st = bfvm::launch_bulk(e))) return st;
        if ((tp.flags & BFVM_CELLS) && e.n_agents > 0 && (st = bfvm::launch_agents(e))) return st;
        if (tp.flags & BFVM_DIFFUSION)
            for (int axis = 0; axis < e.lod_dims; ++axis) {
                if (e.n_dir > 0 && (st = bfvm::launch_dirichlet(e))) return st;
                if ((st = bfvm::launch_sweep(e, axis))) return st;
            }
        return hipSuccess;
    });
}

// Where the agent voxels' values of a step's start are saved (half k % 2).
void* asave_half(const Tape& tp, int k)
{
    return (char*)tp.asave + (size_t)(k & 1) * (tp.asave_bytes / 2);
}

// A step's start changed at its agent voxels (the agents ran on it in place),
// to be written back from `save` before anything reads the start again.
struct PendingRestore {
    void* start = nullptr;
    const void* save = nullptr;
};

hipError_t restore_start(Env& e, const PendingRestore& p)
{
    if (!p.start) return hipSuccess;
    return on_buffer(e, p.start, [&] { return bfvm::launch_agents_restore(e, p.save); });
}

// The step from `start` again, keeping s1 (after bulk), s4, s6, s8 (after
// each sweep; without diffusion all three hold the state after the agents).
// Without a bulk update or Dirichlet nodes the agents run on `start` itself,
// their voxels' values saved to `save` first, and the x sweep reads `start`
// and writes s4 (no copy of the store); `start` is then written back at once,
// or, with `defer`, left to the caller (*pending): the next step's agent
// kernel writes it back (`prev`) in the same launch.  `prev` is always
// written back here.
hipError_t forward_step_keep(Env& e, Tape& tp, const void* start, void* save, const PendingRestore& prev,
                             bool defer, PendingRestore* pending)
{
    hipError_t st;
    const bool bulk = (tp.flags & BFVM_BULK) && e.have_bulk;
    const bool cells = (tp.flags & BFVM_CELLS) && e.n_agents > 0;
    const bool diff = tp.flags & BFVM_DIFFUSION;
    int axis = 0;
    if (pending) *pending = PendingRestore{};
    if (!bulk && diff && e.lod_dims > 0 && e.n_dir == 0 && (!cells || tp.asave)) {
        tp.after_bulk = start;
        void* s0 = const_cast<void*>(start);
        if (cells) {
            st = on_buffer(e, s0, [&] {
                return prev.start ? bfvm::launch_agents_restore_then_save(e, save, prev.start, prev.save)
                                  : bfvm::launch_agents_saving(e, save);
            });
            if (st) return st;
        } else if ((st = restore_start(e, prev))) {
            return st;
        }
        if ((st = bfvm::launch_sweep_to(e, 0, start, tp.s4))) return st;
        if (cells) {
            const PendingRestore mine{s0, save};
            if (defer && pending) *pending = mine;
            else if ((st = restore_start(e, mine))) return st;
        }
        axis = 1;
    } else {
        if ((st = restore_start(e, prev))) return st;
        // without a bulk update the state after it is the start itself (the
        // ring slot stays untouched until the step's transposes have run)
        if (bulk) {
            if ((st = copy_store(e, tp.s1, start))) return st;
            if ((st = on_buffer(e, tp.s1, [&] { return bfvm::launch_bulk(e); }))) return st;
            tp.after_bulk = tp.s1;
        } else {
            tp.after_bulk = start;
        }
        if ((st = copy_store(e, tp.s4, tp.after_bulk))) return st;
        if (cells && (st = on_buffer(e, tp.s4, [&] { return bfvm::launch_agents(e); }))) return st;
    }
    void* bufs[3] = {tp.s4, tp.s6, tp.s8};
    for (; axis < 3; ++axis) {
        const bool sweep = diff && axis < e.lod_dims;
        if (axis > 0 && sweep && e.n_dir == 0) {
            // the y and z sweeps read the kept state and write the next one
            if ((st = bfvm::launch_sweep_to(e, axis, bufs[axis - 1], bufs[axis]))) return st;
            continue;
        }
        // the Dirichlet overwrite must not touch the kept state: copy first
        if (axis > 0 && (st = copy_store(e, bufs[axis], bufs[axis - 1]))) return st;
        if (!sweep) continue;
        st = on_buffer(e, bufs[axis], [&]() -> hipError_t {
            hipError_t q;
            if (e.n_dir > 0 && (q = bfvm::launch_dirichlet(e))) return q;
            return bfvm::launch_sweep(e, axis);
        });
        if (st) return st;
    }
    return hipSuccess;
}

// The transposed step: tp.lam holds dJ/d(step result) on entry and
// dJ/d(step start) on exit; the D / lambda / rate gradients are accumulated.
template <typename T>
hipError_t adjoint_step(Env& e, Tape& tp)
{
    hipError_t st;
    if (tp.flags & BFVM_DIFFUSION) {
        const void* ys[3] = {tp.s4, tp.s6, tp.s8};
        for (int axis = e.lod_dims - 1; axis >= 0; --axis) {
            if (tp.spart[axis]) {
                if ((st = sweep_adjoint_sums(e, tp, ys[axis], axis))) return st;
            } else {
                if ((st = sweep_adjoint(e, tp.lam, axis))) return st;
                if ((st = axis_grad<T>(e, tp, ys[axis], axis))) return st;
            }
            if ((st = dirichlet_zero(e, tp.lam))) return st;
        }
    }
    if ((tp.flags & BFVM_CELLS) && e.n_agent_voxels > 0) {
        const size_t n = (size_t)e.n_agent_voxels * e.S;
        k_agents_adjoint<T><<<grid_for(n), kBlock, 0, e.stream>>>(
            tp.lam, (const T*)tp.after_bulk, e.S, e.n_agent_voxels, e.agent_seg_voxel, e.agent_seg_start,
            (const double*)e.agent_t1, (const double*)e.agent_t2, e.agent_rates ? e.agent_c : nullptr,
            e.agent_sat, tp.rk, tp.gsec, tp.gupt);
        if ((st = hipGetLastError())) return st;
    }
    if ((tp.flags & BFVM_BULK) && e.have_bulk) {
        k_bulk_adjoint<double><<<grid_for(e.nvs), kBlock, 0, e.stream>>>(tp.lam, (const double*)e.bulk_B, e.nvs);
        if ((st = hipGetLastError())) return st;
    }
    return hipSuccess;
}

void free_ptr(void*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
void free_ptr(double*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

hipError_t alloc_store(Env& e, void** p)
{
    return hipMalloc(p, e.nvs * e.elem);
}

// n stores in one allocation (one hipMalloc rather than n: the backward
// pass's set-up), each starting on a 256-byte boundary.
hipError_t alloc_stores(Env& e, size_t n, void** slab, std::vector<void*>& stores)
{
    const size_t pitch = (e.nvs * e.elem + 255) / 256 * 256;
    stores.assign(n, nullptr);
    const hipError_t st = hipMalloc(slab, pitch * n);
    if (st != hipSuccess) {
        *slab = nullptr;
        stores.clear();
        return st;
    }
    for (size_t i = 0; i < n; ++i) stores[i] = (char*)*slab + i * pitch;
    return hipSuccess;
}

int check_tape_handle(Env& e)
{
    if (e.ensemble_stream) return afail(e, BFVM_E_STATE, "the handle belongs to an ensemble");
    return BFVM_OK;
}

// The tape is valid only for the coefficients, agents and sources it was
// recorded with.
int check_unchanged(Env& e, const Tape& tp)
{
    if (e.D != tp.D || e.lambda != tp.lambda || e.agent_epoch != tp.agent_epoch ||
        e.source_epoch != tp.source_epoch)
        return afail(e, BFVM_E_STATE,
                     "coefficients, agents or sources changed since bfvm_tape_record");
    return BFVM_OK;
}

// One segment's backward pass when every step's states are kept: the steps
// again from ring[0], each keeping its states (its result the next slot),
// then the transposed steps in reverse.  shared: the tape's own s1, s4, s6,
// s8 (the slots' states are pointed at the tape step by step).
hipError_t keep_pass(Env& e, Tape& tp, int len, void* const* shared)
{
    const bool bulk = tp.nkeep == 3;
    auto use_slot = [&](int i) {
        void** k = &tp.kept[(size_t)i * tp.nkeep];
        tp.s1 = bulk ? k[0] : shared[0];
        tp.s4 = k[bulk ? 1 : 0];
        tp.s6 = k[bulk ? 2 : 1];
        tp.s8 = i + 1 < len ? tp.ring[i + 1] : tp.last_s8;
    };
    hipError_t st;
    PendingRestore pend;
    for (int i = 0; i < len; ++i) {
        use_slot(i);
        PendingRestore next;
        if ((st = forward_step_keep(e, tp, tp.ring[i], asave_half(tp, i), pend, true, &next))) return st;
        pend = next;
    }
    if ((st = restore_start(e, pend))) return st;
    for (int i = len - 1; i >= 0; --i) {
        use_slot(i);
        tp.after_bulk = bulk ? tp.s1 : tp.ring[i];
        if ((st = e.precision == BFVM_FP64 ? adjoint_step<double>(e, tp) : adjoint_step<float>(e, tp))) return st;
    }
    return hipSuccess;
}

// What a captured keep_pass has baked in besides its length: the buffers,
// the tables and launch shapes (through their epochs), the phases.
std::vector<uintptr_t> segment_key(const Env& e, const Tape& tp, void* const* shared)
{
    std::vector<uintptr_t> k = {(uintptr_t)e.plan_epoch, (uintptr_t)e.coef_epoch, (uintptr_t)e.agent_epoch,
                                (uintptr_t)e.source_epoch, (uintptr_t)e.xy_chunk, (uintptr_t)e.fused_sources,
                                (uintptr_t)e.lod_dims, (uintptr_t)tp.flags, (uintptr_t)tp.nkeep,
                                (uintptr_t)e.bulk_A, (uintptr_t)e.bulk_B, (uintptr_t)tp.lam,
                                (uintptr_t)tp.gpart, (uintptr_t)tp.rk, (uintptr_t)tp.gsec, (uintptr_t)tp.gupt,
                                (uintptr_t)tp.last_s8, (uintptr_t)tp.asave};
    for (int i = 0; i < 4; ++i) k.push_back((uintptr_t)shared[i]);
    for (double* p : tp.spart) k.push_back((uintptr_t)p);
    for (void* p : tp.ring) k.push_back((uintptr_t)p);
    for (void* p : tp.kept) k.push_back((uintptr_t)p);
    return k;
}

// Captures keep_pass over `len` steps into a graph (its replay launches the
// same kernels with the same arguments in the same order: the same bits).
hipError_t capture_keep_pass(Env& e, Tape& tp, int len, void* const* shared, hipGraphExec_t* out)
{
    *out = nullptr;
    hipError_t st = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal);
    if (st != hipSuccess) return st;
    const hipError_t run = keep_pass(e, tp, len, shared);
    hipGraph_t graph = nullptr;
    st = hipStreamEndCapture(e.stream, &graph);
    if (run != hipSuccess || st != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return run != hipSuccess ? run : st;
    }
    st = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (st != hipSuccess) return st;
    if ((st = hipGraphUpload(*out, e.stream)) != hipSuccess) {
        (void)hipGraphExecDestroy(*out);
        *out = nullptr;
    }
    return st;
}

}  // namespace

namespace bfvm {

void free_tape(Env& e)
{
    Tape* tp = e.tape;
    if (!tp) return;
    (void)hipStreamSynchronize(e.stream);
    for (void*& p : tp->cps) free_ptr(p);
    free_ptr(tp->ring_slab);
    free_ptr(tp->s1);
    free_ptr(tp->s4);
    free_ptr(tp->s6);
    free_ptr(tp->s8);
    free_ptr(tp->asave);
    free_ptr(tp->lam);
    free_ptr(tp->rk);
    free_ptr(tp->gsec);
    free_ptr(tp->gupt);
    free_ptr(tp->gc);
    free_ptr(tp->partial);
    free_ptr(tp->gpart);
    for (double*& p : tp->spart) free_ptr(p);
    free_ptr(tp->kept_slab);
    free_ptr(tp->last_s8);
    for (Tape::SegmentGraph& g : tp->graphs) (void)hipGraphExecDestroy(g.exec);
    delete tp;
    e.tape = nullptr;
}

}  // namespace bfvm

extern "C" {

int bfvm_tape_record(bfvm_handle h, double dt, int steps, int flags, int checkpoint_every)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    int rc;
    if ((rc = check_tape_handle(e))) return rc;
    if (steps < 1) return afail(e, BFVM_E_ARG, "steps must be >= 1");
    if (checkpoint_every < 1) return afail(e, BFVM_E_ARG, "checkpoint_every must be >= 1");
    if ((flags & ~BFVM_ALL) != 0) return afail(e, BFVM_E_ARG, "unknown phase flags");
    if ((flags & BFVM_CELLS) && e.n_agents > 0 && e.agent_rates && e.agent_dt != dt)
        return afail(e, BFVM_E_ARG, "dt differs from the dt the agent rates were set with");
    bfvm::free_tape(e);
    Tape* tp = new Tape;
    e.tape = tp;
    tp->dt = dt;
    tp->steps = steps;
    tp->flags = flags;
    tp->K = std::min(checkpoint_every, steps);
    tp->D = e.D;
    tp->lambda = e.lambda;
    tp->agent_epoch = e.agent_epoch;
    tp->source_epoch = e.source_epoch;
    const int ncp = (steps + tp->K - 1) / tp->K;
    tp->cps.assign(ncp, nullptr);
    for (void*& p : tp->cps) AHIP(e, alloc_store(e, &p));
    // the steps of each segment through bfvm_run (HIP graphs of its steps)
    for (int n = 0; n < steps; n += tp->K) {
        AHIP(e, copy_store(e, tp->cps[n / tp->K], e.rho));
        if ((rc = bfvm_run(h, dt, std::min(tp->K, steps - n), flags))) return rc;
    }
    return BFVM_OK;
}

int bfvm_tape_seed(bfvm_handle h, const double* seed)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    Tape* tp = e.tape;
    if (!tp) return afail(e, BFVM_E_STATE, "no tape (bfvm_tape_record)");
    if (!seed) return afail(e, BFVM_E_ARG, "seed is NULL");
    if (!tp->lam) AHIP(e, hipMalloc((void**)&tp->lam, e.nvs * sizeof(double)));
    AHIP(e, hipStreamSynchronize(e.stream));
    AHIP(e, hipMemcpy(tp->lam, seed, e.nvs * sizeof(double), hipMemcpyHostToDevice));
    tp->seeded = true;
    return BFVM_OK;
}

int bfvm_tape_seed_l2(bfvm_handle h, const double* target, double* loss)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    Tape* tp = e.tape;
    if (!tp) return afail(e, BFVM_E_STATE, "no tape (bfvm_tape_record)");
    if (!target) return afail(e, BFVM_E_ARG, "target is NULL");
    if (!tp->lam) AHIP(e, hipMalloc((void**)&tp->lam, e.nvs * sizeof(double)));
    if (!tp->partial) AHIP(e, hipMalloc((void**)&tp->partial, sizeof(double) * kRedBlocks * 2 * kChunk));
    double* dtarget = nullptr;
    double* dloss = nullptr;
    AHIP(e, hipMalloc((void**)&dtarget, e.nvs * sizeof(double)));
    AHIP(e, hipMalloc((void**)&dloss, sizeof(double)));
    hipError_t st = hipMemcpyAsync(dtarget, target, e.nvs * sizeof(double), hipMemcpyHostToDevice, e.stream);
    if (st == hipSuccess) {
        if (e.precision == BFVM_FP64)
            k_l2_seed<double><<<kRedBlocks, kBlock, 0, e.stream>>>((const double*)e.rho, dtarget, tp->lam,
                                                                  e.nvs, tp->partial);
        else
            k_l2_seed<float><<<kRedBlocks, kBlock, 0, e.stream>>>((const float*)e.rho, dtarget, tp->lam,
                                                                 e.nvs, tp->partial);
        st = hipGetLastError();
    }
    if (st == hipSuccess) {
        k_sum<<<1, 64, 0, e.stream>>>(tp->partial, kRedBlocks, dloss);
        st = hipGetLastError();
    }
    double v = 0;
    if (st == hipSuccess) st = hipMemcpyAsync(&v, dloss, sizeof(double), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
    (void)hipFree(dtarget);
    (void)hipFree(dloss);
    AHIP(e, st);
    if (loss) *loss = v;
    tp->seeded = true;
    return BFVM_OK;
}

int bfvm_tape_backward(bfvm_handle h, double* grad_rho0, double* grad_D, double* grad_lambda,
                       double* grad_secretion, double* grad_uptake)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    Tape* tp = e.tape;
    int rc;
    if ((rc = check_tape_handle(e))) return rc;
    if (!tp) return afail(e, BFVM_E_STATE, "no tape (bfvm_tape_record)");
    if (!tp->seeded) return afail(e, BFVM_E_STATE, "no seed (bfvm_tape_seed / bfvm_tape_seed_l2)");
    if ((rc = check_unchanged(e, *tp))) return rc;
    if ((grad_secretion || grad_uptake) && e.n_agents > 0 && !e.agent_rates)
        return afail(e, BFVM_E_STATE, "rate gradients need agents set from rates");
    const int S = e.S;
    const double dt = tp->dt;
    if (tp->flags & BFVM_DIFFUSION) {
        if (!e.uniform) return afail(e, BFVM_E_MESH, "non-uniform mesh");
        if ((rc = bfvm::ensure_coefficients(e, dt))) return rc;
    }
    if (tp->flags & BFVM_BULK) AHIP(e, bfvm::ensure_bulk(e, dt));
    // work buffers
    if (!tp->ring_slab) AHIP(e, alloc_stores(e, tp->K, &tp->ring_slab, tp->ring));
    for (void** p : {&tp->s1, &tp->s4, &tp->s6, &tp->s8})
        if (!*p) AHIP(e, alloc_store(e, p));
    if (!tp->partial) AHIP(e, hipMalloc((void**)&tp->partial, sizeof(double) * kRedBlocks * 2 * kChunk));
    if (!tp->gc) AHIP(e, hipMalloc((void**)&tp->gc, sizeof(double) * 2 * S));
    AHIP(e, hipMemsetAsync(tp->gc, 0, sizeof(double) * 2 * S, e.stream));
    const size_t gpart_n = (size_t)((S + kChunk - 1) / kChunk) * kRedBlocks * 2 * kChunk;
    if (tp->gpart_n != gpart_n) {
        free_ptr(tp->gpart);
        AHIP(e, hipMalloc((void**)&tp->gpart, sizeof(double) * gpart_n));
        tp->gpart_n = gpart_n;
    }
    AHIP(e, hipMemsetAsync(tp->gpart, 0, sizeof(double) * gpart_n, e.stream));
    // the axes whose adjoint sweeps take their D/lambda sums along
    for (int axis = 0; axis < 3; ++axis) {
        bfvm::AdjointSumsShape sh;
        if ((tp->flags & BFVM_DIFFUSION) && axis < e.lod_dims) sh = bfvm::adjoint_sums_shape(e, axis, e.precision);
        tp->stiles[axis] = sh.cols ? sh.tiles : 0;
        const size_t n = tp->stiles[axis] * (size_t)((S + kChunk - 1) / kChunk) * 2 * kChunk;
        if (tp->spart_n[axis] != n) {
            free_ptr(tp->spart[axis]);
            if (n) AHIP(e, hipMalloc((void**)&tp->spart[axis], sizeof(double) * n));
            tp->spart_n[axis] = n;
        }
        if (n) AHIP(e, hipMemsetAsync(tp->spart[axis], 0, sizeof(double) * n, e.stream));
    }
    const size_t na = (size_t)e.n_agents * S;
    if (na > 0) {
        if (tp->agents_n != na) {
            free_ptr(tp->rk);
            free_ptr(tp->gsec);
            free_ptr(tp->gupt);
            tp->agents_n = 0;
            AHIP(e, hipMalloc((void**)&tp->rk, sizeof(double) * na));
            AHIP(e, hipMalloc((void**)&tp->gsec, sizeof(double) * na));
            AHIP(e, hipMalloc((void**)&tp->gupt, sizeof(double) * na));
            tp->agents_n = na;
        }
        AHIP(e, hipMemsetAsync(tp->gsec, 0, sizeof(double) * na, e.stream));
        AHIP(e, hipMemsetAsync(tp->gupt, 0, sizeof(double) * na, e.stream));
    }
    // the agent voxels' saved values (forward_step_keep)
    const size_t asave_bytes = (tp->flags & BFVM_CELLS) ? 2 * (size_t)e.n_agent_voxels * S * e.elem : 0;
    if (tp->asave_bytes != asave_bytes) {
        free_ptr(tp->asave);
        tp->asave_bytes = 0;
        if (asave_bytes) AHIP(e, hipMalloc(&tp->asave, asave_bytes));
        tp->asave_bytes = asave_bytes;
    }
    // segments from the last
    const int K = tp->K;
    // keep every step's states from one pass over the segment when they fit
    // in half the free device memory, else recompute each step twice
    {
        const bool bulk = (tp->flags & BFVM_BULK) && e.have_bulk;
        const int nkeep = bulk ? 3 : 2;
        const size_t store = e.nvs * e.elem;
        const bool ready = e.tape_keep_states && tp->nkeep == nkeep && tp->kept.size() == (size_t)K * nkeep;
        size_t free_b = 0, total_b = 0;
        const size_t want = ready ? 0 : (size_t)K * nkeep + (tp->last_s8 ? 0 : 1);
        const bool fits = e.tape_keep_states && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                          want * store <= free_b / 2;
        if (!ready) {
            free_ptr(tp->kept_slab);
            tp->kept.clear();
            tp->nkeep = 0;
            // an allocation that fails leaves the pass recomputing instead
            if (fits && (tp->last_s8 || alloc_store(e, &tp->last_s8) == hipSuccess) &&
                alloc_stores(e, (size_t)K * nkeep, &tp->kept_slab, tp->kept) == hipSuccess)
                tp->nkeep = nkeep;
            (void)hipGetLastError();
        }
    }
    // the tape's own state buffers, pointed at again however this returns
    struct SharedStates {
        Tape* tp;
        void* s[4];
        ~SharedStates()
        {
            tp->s1 = s[0];
            tp->s4 = s[1];
            tp->s6 = s[2];
            tp->s8 = s[3];
        }
    } keep_states{tp, {tp->s1, tp->s4, tp->s6, tp->s8}};
    // one segment's pass replayed from a HIP graph (per segment length),
    // captured from the second pass on; graphs of an earlier call holding
    // other launches are dropped
    const bool graphs = tp->nkeep && e.graph_steps >= 2 && !e.kernel_timing;
    const std::vector<uintptr_t> key = graphs ? segment_key(e, *tp, keep_states.s) : std::vector<uintptr_t>();
    for (Tape::SegmentGraph& g : tp->graphs)
        if (!graphs || g.key != key) {
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
    tp->graphs.erase(std::remove_if(tp->graphs.begin(), tp->graphs.end(),
                                    [](const Tape::SegmentGraph& g) { return !g.exec; }),
                     tp->graphs.end());
    bool warm = false;  // a pass ran eagerly: kernel attributes are set
    for (int j = (int)tp->cps.size() - 1; j >= 0; --j) {
        const int n0 = j * K, n1 = std::min(n0 + K, tp->steps);
        AHIP(e, copy_store(e, tp->ring[0], tp->cps[j]));
        if (tp->nkeep) {
            const int len = n1 - n0;
            Tape::SegmentGraph* g = nullptr;
            for (Tape::SegmentGraph& c : tp->graphs)
                if (c.len == len) g = &c;
            if (!g && graphs && warm) {
                hipGraphExec_t exec = nullptr;
                AHIP(e, capture_keep_pass(e, *tp, len, keep_states.s, &exec));
                tp->graphs.push_back({exec, len, key});
                g = &tp->graphs.back();
            }
            if (g) {
                AHIP(e, hipGraphLaunch(g->exec, e.stream));
            } else {
                AHIP(e, keep_pass(e, *tp, len, keep_states.s));
                warm = true;
            }
            continue;
        }
        for (int n = n0; n + 1 < n1; ++n) {
            AHIP(e, copy_store(e, tp->ring[n - n0 + 1], tp->ring[n - n0]));
            AHIP(e, forward_step(e, *tp, tp->ring[n - n0 + 1]));
        }
        for (int n = n1 - 1; n >= n0; --n) {
            AHIP(e, forward_step_keep(e, *tp, tp->ring[n - n0], asave_half(*tp, 0), PendingRestore{}, false,
                                      nullptr));
            AHIP(e, e.precision == BFVM_FP64 ? adjoint_step<double>(e, *tp) : adjoint_step<float>(e, *tp));
        }
    }
    if (tp->flags & BFVM_DIFFUSION) AHIP(e, axis_grad_finish(e, *tp));
    tp->seeded = false;  // the adjoint state now holds dJ/drho_0
    // results
    std::vector<double> gc(2 * S);
    AHIP(e, hipMemcpyAsync(gc.data(), tp->gc, sizeof(double) * 2 * S, hipMemcpyDeviceToHost, e.stream));
    if (grad_rho0)
        AHIP(e, hipMemcpyAsync(grad_rho0, tp->lam, e.nvs * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    std::vector<double> gs(na), gu(na);
    if (na > 0 && (grad_secretion || grad_uptake)) {
        AHIP(e, hipMemcpyAsync(gs.data(), tp->gsec, sizeof(double) * na, hipMemcpyDeviceToHost, e.stream));
        AHIP(e, hipMemcpyAsync(gu.data(), tp->gupt, sizeof(double) * na, hipMemcpyDeviceToHost, e.stream));
    }
    AHIP(e, hipStreamSynchronize(e.stream));
    // c1 = dt D / dx^2 (BioFVM_solvers.cpp:111-113), c2 = dt lambda / 3 (:118-119;
    // / 2 for LOD_2D, :313-314)
    for (int s = 0; s < S; ++s) {
        if (grad_D) grad_D[s] = gc[s] * dt / (e.dx * e.dx);
        if (grad_lambda) grad_lambda[s] = gc[S + s] * dt * (e.lod_dims == 2 ? 0.5 : 1.0 / 3.0);
    }
    // grouped order -> list order
    for (int q = 0; q < e.n_agents; ++q) {
        const int a = e.agent_order[q];
        for (int s = 0; s < S; ++s) {
            if (grad_secretion) grad_secretion[(size_t)a * S + s] = gs[(size_t)q * S + s];
            if (grad_uptake) grad_uptake[(size_t)a * S + s] = gu[(size_t)q * S + s];
        }
    }
    return BFVM_OK;
}

int bfvm_tape_free(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    bfvm::free_tape(*h);
    return BFVM_OK;
}

int bfvm_set_tape_keep_states(bfvm_handle h, int on)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    h->tape_keep_states = on != 0;
    return BFVM_OK;
}

int bfvm_get_tape_keep_states(bfvm_handle h, int* on)
{
    bfvm::RandGuard rand_guard;
    if (!h || !on) return BFVM_E_ARG;
    *on = h->tape_keep_states ? 1 : 0;
    return BFVM_OK;
}

int bfvm_tape_states_kept(bfvm_handle h, int* kept)
{
    bfvm::RandGuard rand_guard;
    if (!h || !kept) return BFVM_E_ARG;
    *kept = h->tape && h->tape->nkeep > 0 ? 1 : 0;
    return BFVM_OK;
}

}  // extern "C"
