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
// converges to its interior fixed point), so a thread keeps that row in
// registers and reads the table only at rows < k0 and at row n-1.
//
// Kernels
//   k_sweep_cols  y and z sweeps: lanes run across 64 adjacent columns (i,s)
//                 of a row (coalesced), the block's waves are the segments.
//   k_sweep_rows  x sweep: a tile of whole x-rows is staged through LDS with
//                 coalesced loads/stores; threads are (row, segment, substrate).
//   k_line_sweep  one thread per whole line, for lines longer than the
//                 segmented kernels hold (n > 16*64).
#include <algorithm>

#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxSeg = 16;  // segments per line in the segmented kernels

// Coefficient row p for substrate s: the converged interior from registers,
// else the table.  Tables and arithmetic are FP64 for either store
// precision; an FP32 store is rounded once per sweep, on the way out.
struct Coef {
    const double* __restrict__ r;
    const double* __restrict__ a;
    const double* __restrict__ b;
    double r_in, a_in, b_in;
    int k0, n, S, s;

    __device__ __forceinline__ bool special(int p) const { return p < k0 || p == n - 1; }
};

__device__ __forceinline__ Coef make_coef(const double* r, const double* a, const double* b, int k0,
                                          int n, int S, int s)
{
    Coef c;
    c.r = r; c.a = a; c.b = b; c.k0 = k0; c.n = n; c.S = S; c.s = s;
    int row = k0 < n - 1 ? k0 : n - 1;
    c.r_in = r[row * S + s];
    c.a_in = a[row * S + s];
    c.b_in = b[row * S + s];
    return c;
}

// Forward elimination over v[0..cnt) (positions p0..) with zero carry-in.
// Returns yhat at the segment end and the product of a over the segment.
template <int LMAX>
__device__ __forceinline__ void seg_forward(double (&v)[LMAX], int cnt, int p0, const Coef& c,
                                            double& yend, double& pa)
{
    double y = 0.0, P = 1.0;
#pragma unroll
    for (int t = 0; t < LMAX; ++t) {
        if (t < cnt) {
            int p = p0 + t;
            double r = c.r_in, a = c.a_in;
            if (c.special(p)) {
                r = c.r[p * c.S + c.s];
                a = c.a[p * c.S + c.s];
            }
            y = v[t] * r + a * y;
            v[t] = y;
            P *= a;
        }
    }
    yend = y;
    pa = P;
}

// y_i += (prod_{m=p0..i} a_m) * Y
template <int LMAX>
__device__ __forceinline__ void seg_fix_forward(double (&v)[LMAX], int cnt, int p0, const Coef& c,
                                                double Y)
{
    double P = 1.0;
#pragma unroll
    for (int t = 0; t < LMAX; ++t) {
        if (t < cnt) {
            int p = p0 + t;
            double a = c.special(p) ? c.a[p * c.S + c.s] : c.a_in;
            P *= a;
            v[t] += P * Y;
        }
    }
}

// Back substitution over v[0..cnt) with zero carry-in from above.  Returns
// zhat at the segment start and the product of b over the segment.
template <int LMAX>
__device__ __forceinline__ void seg_backward(double (&v)[LMAX], int cnt, int p0, const Coef& c,
                                             double& zstart, double& pb)
{
    double z = 0.0, Q = 1.0;
#pragma unroll
    for (int t = LMAX - 1; t >= 0; --t) {
        if (t < cnt) {
            int p = p0 + t;
            double b = c.special(p) ? c.b[p * c.S + c.s] : c.b_in;
            if (p == c.n - 1) b = 0.0;  // z_{n-1} = y_{n-1}
            z = v[t] + b * z;
            v[t] = z;
            Q *= b;
        }
    }
    zstart = z;
    pb = Q;
}

// z_i += (prod_{m=i..q-1} b_m) * Z
template <int LMAX>
__device__ __forceinline__ void seg_fix_backward(double (&v)[LMAX], int cnt, int p0, const Coef& c,
                                                 double Z)
{
    double Q = 1.0;
#pragma unroll
    for (int t = LMAX - 1; t >= 0; --t) {
        if (t < cnt) {
            int p = p0 + t;
            double b = c.special(p) ? c.b[p * c.S + c.s] : c.b_in;
            Q *= b;
            v[t] += Q * Z;
        }
    }
}

// y and z sweeps.  Block (64, nseg): lane = column within a 64-column tile of
// one plane, threadIdx.y = segment.  Column col of plane pl starts at
// rho[pl*plane_stride + col]; node p of its line is line_stride further per p.
template <typename T, int LMAX>
__global__ void __launch_bounds__(64 * kMaxSeg)
k_sweep_cols(T* __restrict__ rho, int n, int L, int nseg, size_t line_stride, int ncols,
             size_t plane_stride, int S, const double* __restrict__ tr, const double* __restrict__ ta,
             const double* __restrict__ tb, int k0)
{
    __shared__ double xa[kMaxSeg][64];
    __shared__ double xb[kMaxSeg][64];
    const int lane = threadIdx.x;
    const int g = threadIdx.y;
    const int col = blockIdx.x * 64 + lane;
    const bool act = col < ncols;
    const int s = (act ? col : 0) % S;
    T* base = rho + (size_t)blockIdx.y * plane_stride + (act ? col : 0);
    const int p0 = g * L;
    const int cnt = act ? max(0, min(L, n - p0)) : 0;
    const Coef c = make_coef(tr, ta, tb, k0, n, S, s);

    double v[LMAX];
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) v[t] = (double)base[(size_t)(p0 + t) * line_stride];

    double yend = 0.0, pa = 1.0;
    seg_forward<LMAX>(v, cnt, p0, c, yend, pa);
    xa[g][lane] = yend;
    xb[g][lane] = pa;
    __syncthreads();
    double Y = 0.0;
    for (int u = 0; u < g; ++u) Y = xa[u][lane] + xb[u][lane] * Y;
    seg_fix_forward<LMAX>(v, cnt, p0, c, Y);

    double zs = 0.0, pb = 1.0;
    seg_backward<LMAX>(v, cnt, p0, c, zs, pb);
    __syncthreads();  // all carries Y read
    xa[g][lane] = zs;
    xb[g][lane] = pb;
    __syncthreads();
    double Z = 0.0;
    for (int u = nseg - 1; u > g; --u) Z = xa[u][lane] + xb[u][lane] * Z;
    seg_fix_backward<LMAX>(v, cnt, p0, c, Z);

#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) base[(size_t)(p0 + t) * line_stride] = (T)v[t];
}

// x sweep.  A block owns `rows` consecutive x-rows (contiguous in memory:
// rows*nx*S values), staged in LDS.  Thread tid -> substrate s = tid % S,
// segment g = (tid / S) % nseg, row rr = tid / (S*nseg).
template <typename T, int LMAX>
__global__ void __launch_bounds__(1024)
k_sweep_rows(T* __restrict__ rho, int nx, int L, int nseg, int S, int nrows, int rows_per_block,
             const double* __restrict__ tr, const double* __restrict__ ta,
             const double* __restrict__ tb, int k0)
{
    extern __shared__ double smem_raw[];
    double* xa = smem_raw;                                   // rows_per_block*nseg*S
    double* xb = xa + (size_t)rows_per_block * nseg * S;     // rows_per_block*nseg*S
    T* tile = reinterpret_cast<T*>(xb + (size_t)rows_per_block * nseg * S);
    const int row0 = blockIdx.x * rows_per_block;
    const int rows = min(rows_per_block, nrows - row0);
    const size_t row_len = (size_t)nx * S;
    const size_t tile_len = (size_t)rows * row_len;
    T* src = rho + (size_t)row0 * row_len;

    for (size_t e = threadIdx.x; e < tile_len; e += blockDim.x) tile[e] = src[e];
    __syncthreads();

    const int tid = threadIdx.x;
    const int s = tid % S;
    const int g = (tid / S) % nseg;
    const int rr = tid / (S * nseg);
    const bool act = rr < rows;
    const int p0 = g * L;
    const int cnt = act ? max(0, min(L, nx - p0)) : 0;
    T* line = tile + (size_t)(act ? rr : 0) * row_len + s;
    const int xi = ((act ? rr : 0) * nseg + g) * S + s;
    const Coef c = make_coef(tr, ta, tb, k0, nx, S, s);

    double v[LMAX];
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) v[t] = (double)line[(size_t)(p0 + t) * S];

    double yend = 0.0, pa = 1.0;
    seg_forward<LMAX>(v, cnt, p0, c, yend, pa);
    if (act) {
        xa[xi] = yend;
        xb[xi] = pa;
    }
    __syncthreads();
    double Y = 0.0;
    if (act)
        for (int u = 0; u < g; ++u) Y = xa[xi - (g - u) * S] + xb[xi - (g - u) * S] * Y;
    seg_fix_forward<LMAX>(v, cnt, p0, c, Y);

    double zs = 0.0, pb = 1.0;
    seg_backward<LMAX>(v, cnt, p0, c, zs, pb);
    __syncthreads();
    if (act) {
        xa[xi] = zs;
        xb[xi] = pb;
    }
    __syncthreads();
    double Z = 0.0;
    if (act)
        for (int u = nseg - 1; u > g; --u) Z = xa[xi + (u - g) * S] + xb[xi + (u - g) * S] * Z;
    seg_fix_backward<LMAX>(v, cnt, p0, c, Z);

#pragma unroll
    for (int t = 0; t < LMAX; ++t)
        if (t < cnt) line[(size_t)(p0 + t) * S] = (T)v[t];
    __syncthreads();
    for (size_t e = threadIdx.x; e < tile_len; e += blockDim.x) src[e] = tile[e];
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

// Segment shape for a line of n nodes: L nodes per segment (<= LMAX), nseg
// segments (<= kMaxSeg).  Returns the LMAX instance to use, 0 if none fits.
int pick_lmax(int n, int& L, int& nseg)
{
    static const int lmax_set[] = {8, 16, 32, 64};
    L = std::max(4, (n + kMaxSeg - 1) / kMaxSeg);
    for (int lm : lmax_set) {
        if (L <= lm) {
            nseg = (n + L - 1) / L;
            return lm;
        }
    }
    return 0;
}

struct Geometry {
    int n, ncols, nplanes;
    size_t line_stride, plane_stride;
};

Geometry geometry(const Env& e, int axis)
{
    const size_t S = e.S, nx = e.nx, ny = e.ny;
    if (axis == 0) return {e.nx, e.S, e.ny * e.nz, S, nx * S};
    if (axis == 1) return {e.ny, (int)(nx * S), e.nz, nx * S, nx * ny * S};
    return {e.nz, (int)(nx * ny * S), 1, nx * ny * S, 0};
}

template <typename T>
hipError_t sweep_naive(Env& e, int axis)
{
    Geometry gm = geometry(e, axis);
    size_t threads = (size_t)gm.ncols * gm.nplanes;
    unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
    const AxisCoef& a = e.ax[axis];
    k_line_sweep<T><<<grid, kBlock, 0, e.stream>>>((T*)e.rho, gm.n, gm.line_stride, gm.plane_stride,
                                                  gm.ncols, gm.nplanes, e.S, (const double*)e.c1,
                                                  (const double*)a.denom, (const double*)a.c);
    return hipGetLastError();
}

template <typename T, int LMAX>
hipError_t launch_cols(Env& e, int axis, const Geometry& gm, int L, int nseg)
{
    const AxisCoef& a = e.ax[axis];
    dim3 block(64, nseg);
    dim3 grid((gm.ncols + 63) / 64, gm.nplanes);
    k_sweep_cols<T, LMAX><<<grid, block, 0, e.stream>>>(
        (T*)e.rho, gm.n, L, nseg, gm.line_stride, gm.ncols, gm.plane_stride, e.S,
        (const double*)a.r, (const double*)a.a, (const double*)a.b, a.k0);
    return hipGetLastError();
}

constexpr size_t kRowTileBytes = 48 * 1024;  // LDS budget of one x-row tile

template <typename T, int LMAX>
hipError_t launch_rows(Env& e, int L, int nseg, int rows_per_block)
{
    const AxisCoef& a = e.ax[0];
    const int nrows = e.ny * e.nz;
    int threads = rows_per_block * nseg * e.S;
    threads = (threads + 63) / 64 * 64;
    size_t lds = sizeof(T) * (size_t)rows_per_block * e.nx * e.S +
                 sizeof(double) * 2 * (size_t)rows_per_block * nseg * e.S;
    unsigned grid = (unsigned)((nrows + rows_per_block - 1) / rows_per_block);
    k_sweep_rows<T, LMAX><<<grid, threads, lds, e.stream>>>(
        (T*)e.rho, e.nx, L, nseg, e.S, nrows, rows_per_block, (const double*)a.r,
        (const double*)a.a, (const double*)a.b, a.k0);
    return hipGetLastError();
}

template <typename T>
hipError_t sweep(Env& e, int axis)
{
    Geometry gm = geometry(e, axis);
    int L = 0, nseg = 0;
    int lm = pick_lmax(gm.n, L, nseg);
    if (axis == 0) {
        // rows per block: keep the tile within the LDS budget and the block
        // within 1024 threads
        size_t row_bytes = sizeof(T) * (size_t)e.nx * e.S;
        int per_row_threads = nseg * e.S;
        int rpb = (int)std::min<size_t>(kRowTileBytes / std::max<size_t>(row_bytes, 1),
                                        (size_t)(1024 / std::max(per_row_threads, 1)));
        rpb = std::min(rpb, 8);
        if (lm == 0 || rpb < 1) return sweep_naive<T>(e, axis);
        switch (lm) {
        case 8: return launch_rows<T, 8>(e, L, nseg, rpb);
        case 16: return launch_rows<T, 16>(e, L, nseg, rpb);
        case 32: return launch_rows<T, 32>(e, L, nseg, rpb);
        default: return launch_rows<T, 64>(e, L, nseg, rpb);
        }
    }
    switch (lm) {
    case 8: return launch_cols<T, 8>(e, axis, gm, L, nseg);
    case 16: return launch_cols<T, 16>(e, axis, gm, L, nseg);
    case 32: return launch_cols<T, 32>(e, axis, gm, L, nseg);
    case 64: return launch_cols<T, 64>(e, axis, gm, L, nseg);
    default: return sweep_naive<T>(e, axis);
    }
}

}  // namespace

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis) : sweep<float>(e, axis);
}

}  // namespace bfvm
