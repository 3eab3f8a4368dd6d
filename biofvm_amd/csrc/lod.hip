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
// slot u*C + lane holds (alpha_u, beta_u) on entry and the carry into
// segment u on exit, for c' = alpha_u + beta_u * c walked from segment 0 up
// (forward) or from the top down (backward).  Returns the final carry.
__device__ __forceinline__ double chain(double* xa, const double* xb, int lane, int C, int nseg,
                                        bool forward)
{
    double C0 = 0.0;
    if (forward) {
#pragma unroll 4
        for (int u = 0; u < nseg; ++u) {
            const double a = xa[u * C + lane], b = xb[u * C + lane];
            xa[u * C + lane] = C0;
            C0 = fma(b, C0, a);
        }
    } else {
#pragma unroll 4
        for (int u = nseg - 1; u >= 0; --u) {
            const double a = xa[u * C + lane], b = xb[u * C + lane];
            xa[u * C + lane] = C0;
            C0 = fma(b, C0, a);
        }
    }
    return C0;
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
// Block-level body of the segmented sweep for one plane's column tile.
// C = lines per block (a power of two), tid -> column tid % C, segment tid / C.
template <typename T, int L>
__device__ __forceinline__ void sweep_tile(const SweepArgs& A, int C, int tile, int plane,
                                           double* xa, double* xb, double* wfirst, double* wlast,
                                           double* gtab)
{
    const int n = A.n, nseg = A.nseg, S = A.S, kg0 = A.kg0, kg1 = A.kg1;
    const unsigned line_stride = A.line_stride;
    const int tid = threadIdx.x;
    const int lane = tid & (C - 1);
    const int g = tid / C;
    const int col = tile * C + lane;
    const bool act = col < A.cols_per_plane && g < nseg;
    const int cc = act ? col : 0;
    const int s = cc % S;
    const size_t plane0 = (size_t)plane * A.plane_stride;
    T* rho = (T*)A.rho;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(rho + plane0, (A.nvs - plane0) * sizeof(T));
    const int p0 = act ? g * L : 0;
    const int cnt = act ? min(L, n - p0) : 0;
    const bool has_last = p0 + L >= n;  // holds node n-1 (or nodes past it)
    const double* wb = A.wb;
    const double r = wb[s * 8 + 0], a = wb[s * 8 + 1];
    // byte offset of node p0 of this lane's line; node p0+t adds t*step
    const unsigned vo = ((unsigned)(cc / S) * A.vstride + (unsigned)s + (unsigned)p0 * line_stride) *
                        (unsigned)sizeof(T);
    const unsigned step = line_stride * (unsigned)sizeof(T);

    // The last segment clamps its reads to node n-1 (every address stays in
    // the store); its nodes past n-1 are never stored and cannot reach node
    // n-1 because b is 0 there.  Raw store values are widened to FP64 one by
    // one in the forward pass.
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
    // (read after the first barrier below), zero padded so no index clamps:
    //   G0 rows 0..kg0-1 = g0, then kPad zero rows
    //   G1 kPad zero rows, rows kPad+j = g1 at node n-1-j (j < kg1), kPad zero rows
    //   K  = wb (8 values per substrate)
    double* G0 = gtab;
    double* G1 = gtab + (kg0 + kPad) * S;
    double* K = G1 + (kg1 + 2 * kPad) * S;
    const int ng = (kg0 + kg1 + 3 * kPad) * S;
    for (int i = tid; i < ng + 8 * S; i += blockDim.x) {
        const int row = i / S, cl = i - row * S;
        double val = 0.0;
        if (row < kg0) val = A.g0[i];
        else if (row >= kg0 + 2 * kPad && row < kg0 + kg1 + 2 * kPad)
            val = A.g1[(row - kg0 - 2 * kPad) * S + cl];
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
    if (tid < C) chain(xa, xb, lane, C, nseg, true);
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
    if (tid < C) wfirst[lane] = chain(xa, xb, lane, C, nseg, false);
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
        double zf = v[t] + QZ;
        if (head) zf = fma(mu0, g0r[t * S], zf);
        if (tail) zf = fma(mu1, g1r[-t * S], zf);
        if (store_all || (act && t < cnt)) buf_store((T*)nullptr, rs, zf, vo, t * step);
    }
}

// One grid: block (x, y) = (column tile, plane), COLS lines per block.
template <typename T, int L, int COLS>
__global__ void __launch_bounds__(kMaxThreads, L <= 16 ? 8 : 4) k_sweep(SweepArgs A)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];  // w at node 0 of each column
    __shared__ double wlast[COLS];   // w at node n-1 of each column
    extern __shared__ double gtab[];
    sweep_tile<T, L>(A, COLS, blockIdx.x, blockIdx.y, xa, xb, wfirst, wlast, gtab);
}

// An ensemble: member m owns blocks [first[m], first[m+1]) of a 1-D grid,
// laid out as planes x column tiles of A[m].cols lines; every block has
// kMaxThreads threads (A[m].cols * A[m].nseg of them active).
template <typename T, int L>
__global__ void __launch_bounds__(kMaxThreads, L <= 16 ? 8 : 4)
k_sweep_batch(const SweepArgs* __restrict__ args, const int* __restrict__ first, int nm)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[kMaxThreads];
    __shared__ double wlast[kMaxThreads];
    extern __shared__ double gtab[];
    const int b = blockIdx.x;
    int lo = 0, hi = nm;  // first[lo] <= b < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    const SweepArgs A = args[lo];
    const int local = b - first[lo];
    sweep_tile<T, L>(A, A.cols, local % A.tiles, local / A.tiles, xa, xb, wfirst, wlast, gtab);
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

// Kernel arguments of one axis of one grid, `cols` lines per block.
SweepArgs sweep_args(const Env& e, int axis, int cols, int nseg)
{
    Geometry gm = geometry(e, axis);
    const AxisCoef& a = e.ax[axis];
    SweepArgs A;
    A.rho = e.rho;
    A.nvs = e.nvs;
    A.plane_stride = gm.plane_stride;
    A.wb = (const double*)a.wb;
    A.g0 = (const double*)a.g0;
    A.g1 = (const double*)a.g1;
    A.n = gm.n;
    A.nseg = nseg;
    A.cols_per_plane = gm.ncols;
    A.S = e.S;
    A.kg0 = a.kg0;
    A.kg1 = a.kg1;
    A.line_stride = (unsigned)gm.line_stride;
    A.vstride = (unsigned)gm.vstride;
    A.cols = cols;
    A.tiles = (gm.ncols + cols - 1) / cols;
    A.planes = gm.nplanes;
    return A;
}

size_t sweep_lds(const SweepArgs& A)
{
    return sizeof(double) * (size_t)(A.kg0 + A.kg1 + 3 * kPad + 8) * A.S;
}

template <typename T, int L, int COLS>
hipError_t launch_seg(Env& e, int axis, const Geometry& gm, int nseg)
{
    const SweepArgs A = sweep_args(e, axis, COLS, nseg);
    dim3 grid(A.tiles, gm.nplanes);
    int threads = (COLS * nseg + 63) / 64 * 64;
    k_sweep<T, L, COLS><<<grid, threads, sweep_lds(A), e.stream>>>(A);
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

// Ensemble launch of one axis (see k_sweep_batch).  A member can join when
// its lines fit 16 segments of L nodes and its tables fit the kernel.
namespace {
template <typename T>
bool batch_member_ok_t(const Env& e, int axis, int L)
{
    Geometry gm = geometry(e, axis);
    const size_t span = (axis == 2 ? e.nvs : (size_t)e.nx * e.ny * e.S) * sizeof(T);
    const size_t gbytes =
        sizeof(double) * (size_t)(e.ax[axis].kg0 + e.ax[axis].kg1 + 3 * kPad + 8) * e.S;
    const int nseg = (gm.n + L - 1) / L;
    return gm.n >= 2 && nseg <= kMaxThreads / 64 && span < 0xffffffffull && gbytes <= kMaxCorrectionLds;
}

template <typename T, int L>
hipError_t launch_batch_t(const SweepArgs* d_args, const int* d_first, int nm, int blocks,
                          size_t lds, hipStream_t st)
{
    k_sweep_batch<T, L><<<blocks, kMaxThreads, lds, st>>>(d_args, d_first, nm);
    return hipGetLastError();
}
}  // namespace

bool batch_member_ok(const Env& e, int axis, int L)
{
    return e.precision == BFVM_FP64 ? batch_member_ok_t<double>(e, axis, L)
                                    : batch_member_ok_t<float>(e, axis, L);
}

SweepArgs batch_args(const Env& e, int axis, int L)
{
    Geometry gm = geometry(e, axis);
    const int nseg = (gm.n + L - 1) / L;
    int cols = 64;  // largest power of two with cols*nseg <= 1024, no wider than the plane needs
    while (cols * 2 * nseg <= kMaxThreads && cols < gm.ncols) cols *= 2;
    return sweep_args(e, axis, cols, nseg);
}

size_t batch_lds(const SweepArgs& A) { return sweep_lds(A); }

hipError_t launch_sweep_batch(int precision, int L, const SweepArgs* d_args, const int* d_first,
                              int nm, int blocks, size_t lds, hipStream_t st)
{
    if (precision == BFVM_FP64) {
        switch (L) {
        case 4: return launch_batch_t<double, 4>(d_args, d_first, nm, blocks, lds, st);
        case 8: return launch_batch_t<double, 8>(d_args, d_first, nm, blocks, lds, st);
        case 16: return launch_batch_t<double, 16>(d_args, d_first, nm, blocks, lds, st);
        default: return launch_batch_t<double, 32>(d_args, d_first, nm, blocks, lds, st);
        }
    }
    switch (L) {
    case 4: return launch_batch_t<float, 4>(d_args, d_first, nm, blocks, lds, st);
    case 8: return launch_batch_t<float, 8>(d_args, d_first, nm, blocks, lds, st);
    case 16: return launch_batch_t<float, 16>(d_args, d_first, nm, blocks, lds, st);
    default: return launch_batch_t<float, 32>(d_args, d_first, nm, blocks, lds, st);
    }
}

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis) : sweep<float>(e, axis);
}

}  // namespace bfvm
