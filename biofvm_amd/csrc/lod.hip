// LOD Thomas sweeps (diffusion_decay_solver__constant_coefficients_LOD_3D,
// BioFVM_solvers.cpp:175-273) on the substrate-innermost store.
#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;

// One thread owns one (line, substrate) pair and runs the reference's
// forward elimination and back substitution along it in place:
//   rho_0 /= denom_0;  rho_i = (rho_i + c1*rho_{i-1}) / denom_i
//   rho_i -= c_i * rho_{i+1}                     (BioFVM_solvers.cpp:186-201)
// A line is described by its first element (plane*plane_stride + col) and
// the stride between consecutive nodes; the substrate is col % S because
// every plane starts on a voxel boundary.
template <typename T>
__global__ void k_line_sweep(T* __restrict__ rho, int n, size_t line_stride, size_t plane_stride,
                             int ncols, int nplanes, int S, const T* __restrict__ c1,
                             const T* __restrict__ denom, const T* __restrict__ cc)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncols * nplanes) return;
    int col = (int)(t % ncols);
    size_t plane = t / ncols;
    int s = col % S;
    T* p = rho + plane * plane_stride + col;
    const T a = c1[s];
    T prev = p[0] / denom[s];
    p[0] = prev;
    for (int i = 1; i < n; ++i) {
        T v = (p[(size_t)i * line_stride] + a * prev) / denom[i * S + s];
        p[(size_t)i * line_stride] = v;
        prev = v;
    }
    for (int i = n - 2; i >= 0; --i) {
        T v = p[(size_t)i * line_stride] - cc[i * S + s] * prev;
        p[(size_t)i * line_stride] = v;
        prev = v;
    }
}

template <typename T>
hipError_t sweep_naive(Env& e, int axis)
{
    const size_t S = e.S, nx = e.nx, ny = e.ny;
    int n, ncols, nplanes;
    size_t line_stride, plane_stride;
    if (axis == 0) {  // x: lines along i, one plane per (j,k) row
        n = e.nx; ncols = e.S; nplanes = e.ny * e.nz;
        line_stride = S; plane_stride = nx * S;
    } else if (axis == 1) {  // y: lines along j, columns (i,s) of each k-plane
        n = e.ny; ncols = (int)(nx * S); nplanes = e.nz;
        line_stride = nx * S; plane_stride = nx * ny * S;
    } else {  // z: lines along k, columns (i,j,s)
        n = e.nz; ncols = (int)(nx * ny * S); nplanes = 1;
        line_stride = nx * ny * S; plane_stride = 0;
    }
    size_t threads = (size_t)ncols * nplanes;
    unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
    const AxisCoef& a = e.ax[axis];
    k_line_sweep<T><<<grid, kBlock, 0, e.stream>>>((T*)e.rho, n, line_stride, plane_stride, ncols,
                                                  nplanes, e.S, (const T*)e.c1,
                                                  (const T*)a.denom, (const T*)a.c);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep_naive<double>(e, axis) : sweep_naive<float>(e, axis);
}

}  // namespace bfvm
