// LOD Thomas sweeps (diffusion_decay_solver__constant_coefficients_LOD_3D,
// BioFVM_solvers.cpp:175-273) on the substrate-innermost store.
//
// Segmented Thomas.  Every line of n nodes is cut into nseg segments of at
// most L nodes; one thread holds one segment of one (line, substrate) in
// registers, so each sweep reads and writes the store once.  With
//   forward  y_i = x_i*r_i + a_i*y_{i-1}      backward  z_i = y_i + b_i*z_{i+1}
// a segment [p, q) is first solved with zero carries (y_{p-1} = 0, z_q = 0);
// the true values follow from the carries by linearity,
//   y_i = yhat_i + (prod_{m=p..i} a_m) * Y,   Y = y_{p-1}
//   z_i = zhat_i + (prod_{m=i..q-1} b_m) * Z,  Z = z_q
// and the carries chain across segments through LDS:
//   Y_w = yhat_end(w) + (prod_w a) * Y_{w-1},  Z_w = zhat_start(w) + (prod_w b) * Z_{w+1}.
// This is the reference's forward elimination / back substitution
// (BioFVM_solvers.cpp:186-201) reassociated; the coefficient tables are the
// reference's thomas_denom / thomas_c (:130-141) computed in FP64.
//
// Coefficients: rows [k0, n-1) of the tables are identical (the factorisation
// converges to its interior fixed point), so a segment inside that range
// keeps one row in registers; a segment holding a row < k0 or the last row
// reads the tables.
//
// Kernels
//   k_sweep       all three sweeps: 32 lines per block, up to 32 segments per
//                 line, one thread per (line, segment); see its comment.
//   k_line_sweep  one thread per whole line, for lines longer than the
//                 segmented kernel holds (n > 32*32).
#include <algorithm>

#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxThreads = 1024;  // block = COLS columns x nseg segments

// One thread's segment: nodes [p0, p0+cnt) of one (line, substrate).
// Tables are FP64 for either store precision; an FP32 store is rounded once
// per sweep, on the way out.
struct SegCtx {
    const double* __restrict__ tr;  // r_i = 1/denom_i           [i*S + s]
    const double* __restrict__ ta;  // a_i = c1/denom_i
    const double* __restrict__ tb;  // b_i = -c_i (0 at the last node)
    double r_in, a_in, b_in;        // the converged interior row
    int p0, cnt, S, s;
    bool special;                   // segment holds a row < k0 or the last row
};

__device__ __forceinline__ SegCtx make_ctx(const double* tr, const double* ta, const double* tb,
                                           int k0, int n, int S, int s, int p0, int cnt)
{
    SegCtx c;
    c.tr = tr; c.ta = ta; c.tb = tb;
    c.S = S; c.s = s; c.p0 = p0; c.cnt = cnt;
    const int row = k0 < n - 1 ? k0 : n - 1;
    c.r_in = tr[row * S + s];
    c.a_in = ta[row * S + s];
    c.b_in = tb[row * S + s];
    c.special = cnt > 0 && (p0 < k0 || p0 + cnt >= n);
    return c;
}

// Forward elimination over v[0..cnt) with zero carry-in: yhat, its value at
// the segment end and the product of a over the segment.  Interior segments
// use the register constants; the others read the tables row by row
// (branch-free, so the loads pipeline ahead of the recurrence).
template <int LMAX>
__device__ __forceinline__ void seg_forward(double (&v)[LMAX], const SegCtx& c, double& yend,
                                            double& pa)
{
    double y = 0.0, P = 1.0;
    if (!c.special) {
#pragma unroll
        for (int t = 0; t < LMAX; ++t)
            if (t < c.cnt) {
                y = fma(c.a_in, y, v[t] * c.r_in);
                v[t] = y;
                P *= c.a_in;
            }
    } else {
#pragma unroll
        for (int t = 0; t < LMAX; ++t)
            if (t < c.cnt) {
                const int q = (c.p0 + t) * c.S + c.s;
                const double a = c.ta[q];
                y = fma(a, y, v[t] * c.tr[q]);
                v[t] = y;
                P *= a;
            }
    }
    yend = y;
    pa = P;
}

// y_i += (prod_{m=p0..i} a_m) * Y
template <int LMAX>
__device__ __forceinline__ void seg_fix_forward(double (&v)[LMAX], const SegCtx& c, double Y)
{
    double P = 1.0;
    if (!c.special) {
#pragma unroll
        for (int t = 0; t < LMAX; ++t)
            if (t < c.cnt) {
                P *= c.a_in;
                v[t] = fma(P, Y, v[t]);
            }
    } else {
#pragma unroll
        for (int t = 0; t < LMAX; ++t)
            if (t < c.cnt) {
                P *= c.ta[(c.p0 + t) * c.S + c.s];
                v[t] = fma(P, Y, v[t]);
            }
    }
}

// Back substitution over v[0..cnt) with zero carry-in from above: zhat, its
// value at the segment start and the product of b over the segment.
template <int LMAX>
__device__ __forceinline__ void seg_backward(double (&v)[LMAX], const SegCtx& c, double& zstart,
                                             double& pb)
{
    double z = 0.0, Q = 1.0;
    if (!c.special) {
#pragma unroll
        for (int t = LMAX - 1; t >= 0; --t)
            if (t < c.cnt) {
                z = fma(c.b_in, z, v[t]);
                v[t] = z;
                Q *= c.b_in;
            }
    } else {
#pragma unroll
        for (int t = LMAX - 1; t >= 0; --t)
            if (t < c.cnt) {
                const double b = c.tb[(c.p0 + t) * c.S + c.s];
                z = fma(b, z, v[t]);
                v[t] = z;
                Q *= b;
            }
    }
    zstart = z;
    pb = Q;
}

// z_i += (prod_{m=i..q-1} b_m) * Z
template <int LMAX>
__device__ __forceinline__ void seg_fix_backward(double (&v)[LMAX], const SegCtx& c, double Z)
{
    double Q = 1.0;
    if (!c.special) {
#pragma unroll
        for (int t = LMAX - 1; t >= 0; --t)
            if (t < c.cnt) {
                Q *= c.b_in;
                v[t] = fma(Q, Z, v[t]);
            }
    } else {
#pragma unroll
        for (int t = LMAX - 1; t >= 0; --t)
            if (t < c.cnt) {
                Q *= c.tb[(c.p0 + t) * c.S + c.s];
                v[t] = fma(Q, Z, v[t]);
            }
    }
}

// One kernel for the three sweeps.  The lines of a sweep are "columns": a
// column is one (voxel of a plane, substrate) pair, col = q*S + s, whose line
// starts at rho[plane*plane_stride + q*vstride + s] and advances line_stride
// per node.
//   x: plane = k, q = j, vstride = nx*S, line_stride = S
//   y: plane = k, q = i, vstride = S,    line_stride = nx*S
//   z: one plane, q = j*nx + i, vstride = S, line_stride = nx*ny*S
// Block = COLS columns x nseg segments; thread tid -> column tid % COLS,
// segment tid / COLS, so a wave holds COLS columns of 64/COLS consecutive
// segments.  For y and z the COLS columns of a row are contiguous
// (coalesced); for x they are row pieces of S contiguous values, and
// consecutive nodes of a segment reuse the same cache lines.  Two blocks of
// 1024 threads fit a CU (8 waves per SIMD) up to 16 nodes per segment.
template <typename T, int LMAX, int COLS>
__global__ void __launch_bounds__(kMaxThreads, LMAX <= 16 ? 8 : 4)
k_sweep(T* __restrict__ rho, int n, int L, int nseg, size_t line_stride, int cols_per_plane,
        size_t vstride, size_t plane_stride, int S, const double* __restrict__ tr,
        const double* __restrict__ ta, const double* __restrict__ tb, int k0)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    const int tid = threadIdx.x;
    const int lane = tid % COLS;
    const int g = tid / COLS;
    const int col = blockIdx.x * COLS + lane;
    const bool act = col < cols_per_plane && g < nseg;
    const int cc = act ? col : 0;
    const int s = cc % S;
    T* base = rho + (size_t)blockIdx.y * plane_stride + (size_t)(cc / S) * vstride + s;
    const int p0 = g * L;
    const int cnt = act ? max(0, min(L, n - p0)) : 0;
    const SegCtx c = make_ctx(tr, ta, tb, k0, n, S, s, p0, cnt);

    double v[LMAX];
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) v[t] = (double)base[(size_t)(p0 + t) * line_stride];

    double yend = 0.0, pa = 1.0;
    seg_forward<LMAX>(v, c, yend, pa);
    xa[tid] = yend;
    xb[tid] = pa;
    __syncthreads();
    double Y = 0.0;
#pragma unroll 4
    for (int u = 0; u < g; ++u) Y = fma(xb[u * COLS + lane], Y, xa[u * COLS + lane]);
    seg_fix_forward<LMAX>(v, c, Y);

    double zs = 0.0, pb = 1.0;
    seg_backward<LMAX>(v, c, zs, pb);
    __syncthreads();  // every forward carry has been read
    xa[tid] = zs;
    xb[tid] = pb;
    __syncthreads();
    double Z = 0.0;
#pragma unroll 4
    for (int u = nseg - 1; u > g; --u) Z = fma(xb[u * COLS + lane], Z, xa[u * COLS + lane]);
    seg_fix_backward<LMAX>(v, c, Z);

#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) base[(size_t)(p0 + t) * line_stride] = (T)v[t];
}

// One thread owns one whole (line, substrate) pair and runs the reference's
// forward elimination and back substitution along it in place
// (BioFVM_solvers.cpp:186-201).  Used for lines too long to segment.
template <typename T>
__global__ void k_line_sweep(T* __restrict__ rho, int n, size_t line_stride, size_t plane_stride,
                             int ncols, int nplanes, int S, const double* __restrict__ c1,
                             const double* __restrict__ denom, const double* __restrict__ cc)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncols * nplanes) return;
    int col = (int)(t % ncols);
    size_t plane = t / ncols;
    int s = col % S;
    T* p = rho + plane * plane_stride + col;
    // intermediate values stay FP64 in the store only for an FP64 store; an
    // FP32 store holds the rounded forward values between the two passes
    const double a = c1[s];
    double prev = (double)p[0] / denom[s];
    p[0] = (T)prev;
    for (int i = 1; i < n; ++i) {
        double v = ((double)p[(size_t)i * line_stride] + a * prev) / denom[i * S + s];
        p[(size_t)i * line_stride] = (T)v;
        prev = v;
    }
    for (int i = n - 2; i >= 0; --i) {
        double v = (double)p[(size_t)i * line_stride] - cc[i * S + s] * prev;
        p[(size_t)i * line_stride] = (T)v;
        prev = v;
    }
}

struct Geometry {
    int n, ncols, nplanes;
    size_t line_stride, plane_stride, vstride;
};

// Lines of one sweep as columns of planes (see k_sweep).
Geometry geometry(const Env& e, int axis)
{
    const size_t S = e.S, nx = e.nx, ny = e.ny;
    if (axis == 0) return {e.nx, (int)(ny * S), e.nz, S, nx * ny * S, nx * S};
    if (axis == 1) return {e.ny, (int)(nx * S), e.nz, nx * S, nx * ny * S, S};
    return {e.nz, (int)(nx * ny * S), 1, nx * ny * S, 0, S};
}

// Launch shape of one sweep: COLS lines per block and L nodes per segment
// (nseg = ceil(n/L) segments, COLS*nseg <= 1024, L <= 32).  cols == 0 means
// the line is too long for the segmented kernel.  A plan set by the
// autotuner (bfvm_autotune) wins over the default.
SweepPlan default_plan(int n)
{
    SweepPlan p;
    p.cols = 32;
    p.L = std::max(4, (n + 31) / 32);  // up to 32 segments of 32 lines
    if (p.L > 32) {
        p.cols = 16;  // up to 64 segments of 16 lines
        p.L = std::max(4, (n + 63) / 64);
    }
    if (p.L > 32) p.cols = 0;
    return p;
}

SweepPlan plan_for(const Env& e, int axis)
{
    const SweepPlan& t = e.plan[axis];
    if (t.cols > 0) return t;
    return default_plan(axis == 0 ? e.nx : axis == 1 ? e.ny : e.nz);
}

template <typename T>
hipError_t sweep_naive(Env& e, int axis)
{
    // one thread per whole line: lines as (plane, column) with the naive
    // kernel's plane/column convention
    const size_t S = e.S, nx = e.nx, ny = e.ny;
    int n, ncols, nplanes;
    size_t line_stride, plane_stride;
    if (axis == 0) {
        n = e.nx; ncols = e.S; nplanes = e.ny * e.nz; line_stride = S; plane_stride = nx * S;
    } else if (axis == 1) {
        n = e.ny; ncols = (int)(nx * S); nplanes = e.nz; line_stride = nx * S; plane_stride = nx * ny * S;
    } else {
        n = e.nz; ncols = (int)(nx * ny * S); nplanes = 1; line_stride = nx * ny * S; plane_stride = 0;
    }
    size_t threads = (size_t)ncols * nplanes;
    unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
    const AxisCoef& a = e.ax[axis];
    k_line_sweep<T><<<grid, kBlock, 0, e.stream>>>((T*)e.rho, n, line_stride, plane_stride, ncols,
                                                  nplanes, e.S, (const double*)e.c1,
                                                  (const double*)a.denom, (const double*)a.c);
    return hipGetLastError();
}

template <typename T, int LMAX, int COLS>
hipError_t launch_seg(Env& e, int axis, const Geometry& gm, int L, int nseg)
{
    const AxisCoef& a = e.ax[axis];
    dim3 grid((gm.ncols + COLS - 1) / COLS, gm.nplanes);
    int threads = (COLS * nseg + 63) / 64 * 64;
    k_sweep<T, LMAX, COLS><<<grid, threads, 0, e.stream>>>(
        (T*)e.rho, gm.n, L, nseg, gm.line_stride, gm.ncols, gm.vstride, gm.plane_stride, e.S,
        (const double*)a.r, (const double*)a.a, (const double*)a.b, a.k0);
    return hipGetLastError();
}

template <typename T, int COLS>
hipError_t launch_cols(Env& e, int axis, const Geometry& gm, int L, int nseg)
{
    if (L <= 8) return launch_seg<T, 8, COLS>(e, axis, gm, L, nseg);
    if (L <= 16) return launch_seg<T, 16, COLS>(e, axis, gm, L, nseg);
    return launch_seg<T, 32, COLS>(e, axis, gm, L, nseg);
}

template <typename T>
hipError_t sweep(Env& e, int axis)
{
    Geometry gm = geometry(e, axis);
    SweepPlan p = plan_for(e, axis);
    if (p.cols == 0) return sweep_naive<T>(e, axis);
    int nseg = (gm.n + p.L - 1) / p.L;
    switch (p.cols) {
    case 16: return launch_cols<T, 16>(e, axis, gm, p.L, nseg);
    case 32: return launch_cols<T, 32>(e, axis, gm, p.L, nseg);
    default: return launch_cols<T, 64>(e, axis, gm, p.L, nseg);
    }
}

}  // namespace

SweepPlan effective_plan(const Env& e, int axis)
{
    return plan_for(e, axis);
}

std::vector<SweepPlan> sweep_candidates(int n)
{
    std::vector<SweepPlan> out;
    for (int cols : {16, 32, 64}) {
        int lmin = std::max(4, (n + kMaxThreads / cols - 1) / (kMaxThreads / cols));
        for (int L : {lmin, 8, 12, 16, 24, 32}) {
            if (L < lmin || L > 32) continue;
            if (L > n && L != lmin) continue;
            bool dup = false;
            for (const SweepPlan& q : out) dup |= q.cols == cols && q.L == L;
            if (!dup) out.push_back(SweepPlan{cols, L});
        }
    }
    return out;
}

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis) : sweep<float>(e, axis);
}

}  // namespace bfvm
