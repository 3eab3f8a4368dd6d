// LOD Thomas sweeps (diffusion_decay_solver__constant_coefficients_LOD_3D,
// BioFVM_solvers.cpp:175-273) on the substrate-innermost store.
//
// Each sweep solves A z = x on every line, A the reference's tridiagonal
// matrix (BioFVM_solvers.cpp:104-142).  The solve is reorganised for the GPU:
//   1. A = A_inf - (rank 2 at the corners), A_inf = L U with constant
//      bidiagonal factors (see AxisCoef), so z = A_inf^-1 x plus a rank-2
//      Woodbury correction confined to the first kg0 / last kg1 nodes;
//   2. each line is cut into segments of L nodes held in registers, one
//      thread per (line, segment): the constant-coefficient recurrences are
//      run with zero carries, the carries are chained across segments
//      through LDS, and added back by linearity,
//        y_i = yhat_i + a^(i-p+1) Y,   z_i = zhat_i + a^(q-i) Z.
// Each sweep reads and writes the store once.  The arithmetic is FP64 for
// either store precision.  The result equals the reference's forward
// elimination / back substitution (:186-201) up to rounding.
//
// Kernels
//   k_sweep       all three sweeps: 16-64 lines per block, up to 1024/lines
//                 segments of 4-32 nodes per line; see its comment.
//   k_line_sweep  one thread per whole line with the reference's own
//                 recurrences, for lines of one node or longer than the
//                 segmented kernel holds (n > 64*32).
#include <algorithm>

#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxThreads = 1024;  // block = COLS columns x nseg segments
constexpr size_t kMaxCorrectionLds = 48 * 1024;  // g0/g1 rows staged per block
constexpr int kPad = 32;  // zero rows around the correction tables (>= the largest L)

// Raw buffer access: a per-lane byte offset plus a uniform (scalar) one, no
// per-element address arithmetic.  Reads past the resource's end return 0
// and writes there are dropped.  The descriptor's last word is the raw
// 32-bit-element format used for gfx9-class buffers.
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, size_t bytes)
{
    const int nrec = bytes >= 0xffffffffull ? -1 : (int)(unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nrec, kRsrcWord3);
}

__device__ __forceinline__ float buf_load(float*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ __forceinline__ double buf_load(double*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ void buf_store(float*, __amdgpu_buffer_rsrc_t r, double x, unsigned vo,
                                          unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)x), r, vo, so, 0);
}
__device__ __forceinline__ void buf_store(double*, __amdgpu_buffer_rsrc_t r, double x, unsigned vo,
                                          unsigned so)
{
    using u64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u64, x), r, vo, so, 0);
}

// Carry chain of one column across its nseg segments, run by one thread:
// slot u holds (alpha_u, beta_u) on entry and the carry into segment u on
// exit, for C' = alpha_u + beta_u * C walked from segment 0 up (forward) or
// from the top down (backward).  Returns the final C.
template <int COLS>
__device__ __forceinline__ double chain(double* xa, const double* xb, int lane, int nseg,
                                        bool forward)
{
    double C = 0.0;
    if (forward) {
#pragma unroll 4
        for (int u = 0; u < nseg; ++u) {
            const double a = xa[u * COLS + lane], b = xb[u * COLS + lane];
            xa[u * COLS + lane] = C;
            C = fma(b, C, a);
        }
    } else {
#pragma unroll 4
        for (int u = nseg - 1; u >= 0; --u) {
            const double a = xa[u * COLS + lane], b = xb[u * COLS + lane];
            xa[u * COLS + lane] = C;
            C = fma(b, C, a);
        }
    }
    return C;
}

// One kernel for the three sweeps.  The lines of a sweep are "columns": a
// column is one (voxel of a plane, substrate) pair, col = q*S + s, whose line
// starts at rho[plane*plane_stride + q*vstride + s] and advances line_stride
// per node.
//   x: plane = k, q = j, vstride = nx*S, line_stride = S
//   y: plane = k, q = i, vstride = S,    line_stride = nx*S
//   z: one plane, q = j*nx + i, vstride = S, line_stride = nx*ny*S
// Block = COLS columns x nseg segments of L nodes; thread tid -> column
// tid % COLS, segment tid / COLS, so a wave holds COLS columns of 64/COLS
// consecutive segments.  For y and z the COLS columns of a row are
// contiguous (coalesced); for x they are row pieces of S contiguous values,
// and consecutive nodes of a segment reuse the same cache lines.
//
// Each thread solves A_inf on its segment with constant coefficients (see
// AxisCoef): forward y_i = a*y_{i-1} + x_i*r and backward z_i = y_i + a*z_{i+1}
// (b = a except at the last node, where b = 0), first with zero carries;
// the segment-end values go through LDS, one thread per column walks the
// carry chain, and each segment adds its carry times the running product
// of a.  Segments within kg0 / kg1 of the line ends then add the rank-2
// correction.  Only the segment holding node n-1 can be cut short: its nodes
// past n-1 are clamped reads, are never stored, and cannot reach node n-1
// because b is 0 there.  No loop reads a table or needs a mask.
template <typename T, int L, int COLS>
__global__ void __launch_bounds__(kMaxThreads, L <= 16 ? 8 : 4)
k_sweep(T* __restrict__ rho, size_t nvs, int n, int nseg, unsigned line_stride, int cols_per_plane,
        unsigned vstride, size_t plane_stride, int S, const double* __restrict__ wb,
        const double* __restrict__ g0, const double* __restrict__ g1, int kg0, int kg1)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];  // w at node 0 of each column
    __shared__ double wlast[COLS];   // w at node n-1 of each column
    // correction tables with zero padding so that no index needs a clamp:
    //   G0 rows 0..kg0-1 = g0, then kPad zero rows
    //   G1 kPad zero rows, rows kPad+j = g1 at node n-1-j (j < kg1), kPad zero rows
    //   then wb (8 values per substrate)
    extern __shared__ double gtab[];
    const int tid = threadIdx.x;
    const int lane = tid % COLS;
    const int g = tid / COLS;
    const int col = blockIdx.x * COLS + lane;
    const bool act = col < cols_per_plane && g < nseg;
    const int cc = act ? col : 0;
    const int s = cc % S;
    const size_t plane0 = (size_t)blockIdx.y * plane_stride;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(rho + plane0, (nvs - plane0) * sizeof(T));
    const int p0 = act ? g * L : 0;
    const int cnt = act ? min(L, n - p0) : 0;
    const bool has_last = p0 + L >= n;  // holds node n-1 (or nodes past it)
    const double r = wb[s * 8 + 0], a = wb[s * 8 + 1];
    // byte offset of node p0 of this lane's line; node p0+t adds t*step
    const unsigned vo = ((unsigned)(cc / S) * vstride + (unsigned)s + (unsigned)p0 * line_stride) *
                        (unsigned)sizeof(T);
    const unsigned step = line_stride * (unsigned)sizeof(T);

    // The last segment clamps its reads to node n-1 (every address stays in
    // the store); its nodes past n-1 are never stored and cannot reach node
    // n-1 because b is 0 there.
    // raw store values, widened to FP64 one by one in the forward pass
    T raw[L];
    if (!has_last) {
#pragma unroll
        for (int t = 0; t < L; ++t) raw[t] = buf_load((T*)nullptr, rs, vo, t * step);
    } else {
        const unsigned vl = vo + (unsigned)(n - 1 - p0) * step;  // node n-1
#pragma unroll
        for (int t = 0; t < L; ++t)
            raw[t] = buf_load((T*)nullptr, rs, p0 + t < n ? vo + t * step : vl, 0);
    }

    // the correction tables, staged in LDS while the loads are in flight
    // (read after the first barrier below)
    double* G0 = gtab;
    double* G1 = gtab + (kg0 + kPad) * S;
    double* K = G1 + (kg1 + 2 * kPad) * S;
    const int ng = (kg0 + kg1 + 3 * kPad) * S;
    for (int i = tid; i < ng + 8 * S; i += blockDim.x) {
        const int row = i / S, col = i - row * S;
        double val = 0.0;
        if (row < kg0) val = g0[i];
        else if (row >= kg0 + 2 * kPad && row < kg0 + kg1 + 2 * kPad)
            val = g1[(row - kg0 - 2 * kPad) * S + col];
        else if (i >= ng) val = wb[i - ng];
        gtab[i] = val;
    }

    // forward, zero carry-in
    double v[L];
    double y = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) v[t] = y = fma(a, y, (double)raw[t] * r);
    double aL = a;
#pragma unroll
    for (int t = 1; t < L; ++t) aL *= a;
    xa[tid] = y;
    xb[tid] = aL;
    __syncthreads();
    if (tid < COLS) chain<COLS>(xa, xb, lane, nseg, true);
    __syncthreads();
    {
        double PY = xa[tid];
#pragma unroll
        for (int t = 0; t < L; ++t) {
            PY *= a;
            v[t] += PY;
        }
    }
    if (has_last && act) {  // z_{n-1} = y_{n-1}
        double wl = 0.0;
#pragma unroll
        for (int t = 0; t < L; ++t)
            if (p0 + t == n - 1) wl = v[t];
        wlast[lane] = wl;
    }

    // backward, zero carry-in from above
    double z = 0.0, bL;
    if (!has_last) {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(a, z, v[t]);
        bL = aL;
    } else {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(p0 + t < n - 1 ? a : 0.0, z, v[t]);
        bL = 0.0;  // nothing lies above the last segment
    }
    xa[tid] = z;  // each thread overwrites only the slot it alone has read
    xb[tid] = bL;
    __syncthreads();
    if (tid < COLS) wfirst[lane] = chain<COLS>(xa, xb, lane, nseg, false);
    __syncthreads();

    // Last pass, from the top of the segment down, each node finished and
    // stored as soon as it is final: the carry from above (a^(q-i) Z), the
    // rank-2 correction near the line ends (mu0 g0_i + mu1 g1_i), store.
    const double QZ0 = has_last ? 0.0 : xa[tid];  // the last segment has nothing above
    const bool head = p0 < kg0, tail = p0 + L > n - kg1;
    double mu0 = 0.0, mu1 = 0.0;
    if (head || tail) {
        const double w0 = wfirst[lane], wn = wlast[lane];
        const double* k = K + s * 8 + 2;  // K00, K01, K10, K11
        mu0 = head ? k[0] * w0 + k[1] * wn : 0.0;
        mu1 = tail ? k[2] * w0 + k[3] * wn : 0.0;
    }
    // g0 rows from node p0 up; g1 rows from node p0 down (zero padded)
    const double* g0r = G0 + (head ? p0 : 0) * S + s;
    const double* g1r = G1 + (tail ? kPad + n - 1 - p0 : 0) * S + s;
    const bool store_all = act && !has_last;
    double QZ = QZ0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
        QZ *= a;
        double z = v[t] + QZ;
        if (head) z = fma(mu0, g0r[t * S], z);
        if (tail) z = fma(mu1, g1r[-t * S], z);
        if (store_all || (act && t < cnt)) buf_store((T*)nullptr, rs, z, vo, t * step);
    }
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

// Launch shape of one sweep: COLS lines per block and L nodes per segment,
// L one of kNodes, nseg = ceil(n/L) segments with COLS*nseg <= 1024.
// cols == 0 means the line is too long for the segmented kernel.  A plan set
// by the autotuner (bfvm_autotune) wins over the default.
constexpr int kNodes[] = {4, 8, 12, 16, 24, 32};
constexpr int kColsSet[] = {16, 32, 64};

bool plan_valid(int n, int cols, int L)
{
    return cols * ((n + L - 1) / L) <= kMaxThreads;
}

SweepPlan default_plan(int n)
{
    for (int cols : {32, 16})
        for (int L : kNodes)
            if (plan_valid(n, cols, L)) return SweepPlan{cols, L};
    return SweepPlan{0, 0};
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

template <typename T, int L, int COLS>
hipError_t launch_seg(Env& e, int axis, const Geometry& gm, int nseg)
{
    const AxisCoef& a = e.ax[axis];
    dim3 grid((gm.ncols + COLS - 1) / COLS, gm.nplanes);
    int threads = (COLS * nseg + 63) / 64 * 64;
    const size_t lds = sizeof(double) * (size_t)(a.kg0 + a.kg1 + 3 * kPad + 8) * e.S;
    k_sweep<T, L, COLS><<<grid, threads, lds, e.stream>>>(
        (T*)e.rho, e.nvs, gm.n, nseg, (unsigned)gm.line_stride, gm.ncols, (unsigned)gm.vstride,
        gm.plane_stride, e.S, (const double*)a.wb, (const double*)a.g0, (const double*)a.g1, a.kg0,
        a.kg1);
    return hipGetLastError();
}

template <typename T, int COLS>
hipError_t launch_cols(Env& e, int axis, const Geometry& gm, int L, int nseg)
{
    switch (L) {
    case 4: return launch_seg<T, 4, COLS>(e, axis, gm, nseg);
    case 8: return launch_seg<T, 8, COLS>(e, axis, gm, nseg);
    case 12: return launch_seg<T, 12, COLS>(e, axis, gm, nseg);
    case 16: return launch_seg<T, 16, COLS>(e, axis, gm, nseg);
    case 24: return launch_seg<T, 24, COLS>(e, axis, gm, nseg);
    default: return launch_seg<T, 32, COLS>(e, axis, gm, nseg);
    }
}

template <typename T>
hipError_t sweep(Env& e, int axis)
{
    Geometry gm = geometry(e, axis);
    SweepPlan p = plan_for(e, axis);
    // buffer offsets are 32-bit: the segmented kernel addresses at most 4 GiB
    // from a plane's start
    const size_t span = (axis == 2 ? e.nvs : (size_t)e.nx * e.ny * e.S) * sizeof(T);
    // the end-correction tables live in LDS beside the 16 KiB carry arrays
    const size_t gbytes =
        sizeof(double) * (size_t)(e.ax[axis].kg0 + e.ax[axis].kg1 + 3 * kPad + 8) * e.S;
    if (gm.n < 2 || p.cols == 0 || !plan_valid(gm.n, p.cols, p.L) || span >= 0xffffffffull ||
        gbytes > kMaxCorrectionLds)
        return sweep_naive<T>(e, axis);
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
    for (int cols : kColsSet)
        for (int L : kNodes)
            if (plan_valid(n, cols, L) && (L <= 8 || L < 2 * n)) out.push_back(SweepPlan{cols, L});
    return out;
}

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis) : sweep<float>(e, axis);
}

}  // namespace bfvm
