// Bandwidth ceilings for the sweeps, measured by `bench.py --ceiling`.
//
// Built into its own library (libbfvm_ceiling.so), never linked into the
// product library: these kernels only move bytes.  Two kinds:
//  - streaming passes over one buffer (in-place read+write, copy, read only,
//    write only), one float4 per lane, plain or non-temporal; a region
//    smaller than the buffer, passed over again and again, stays in the
//    256 MB Infinity Cache;
//  - the sweeps' access pattern: the same column tiles, block shapes, line
//    and plane strides and XCD tile numbering as the segmented sweeps (C
//    lines per block, L nodes per thread, ceil(n/L) segments per line), each
//    thread loading its L values, one barrier, the same values stored back
//    (scaled by a runtime 1.0); no solve.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

constexpr unsigned kXcds = 8;

// Tile at launch position `lin` of N: 0 = as dispatched, k > 0 = XCD x runs
// chunks of k consecutive tiles, -1 = XCD x runs the x-th contiguous eighth
// (workgroups are dealt to the 8 XCDs round robin).
__device__ unsigned xcd_map(unsigned lin, unsigned N, int spread)
{
    if (spread == 0) return lin;
    if (spread < 0) {
        const unsigned q = N / kXcds, r = N % kXcds, x = lin % kXcds, i = lin / kXcds;
        return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    }
    const unsigned k = (unsigned)spread, R = k * kXcds;
    if (lin >= N / R * R) return lin;
    const unsigned j = lin % R;
    return lin - j + (j % kXcds) * k + j / kXcds;
}

typedef float f4v __attribute__((ext_vector_type(4)));

// One float4 per thread over an exact grid (every load of a wave issued at
// once, no loop), plain or non-temporal (NT) loads and stores: the fastest
// streaming form measured on this part (a grid-stride loop with one float4 in
// flight per iteration reached 4.7 TB/s in place against 5.9-6.7 here).
template <int NT>
__global__ void __launch_bounds__(256) k_pass(f4v* a, f4v* b, size_t n4, int kind, float s, float* sink)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    f4v v = {s, s, s, s};
    if (kind != 3) v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    if (kind == 2) {
        if (v.x + v.y + v.z + v.w == 12345.f) sink[0] = v.x;  // keeps the load
        return;
    }
    if (kind == 0) v *= s;
    f4v* dst = kind == 1 ? b + i : a + i;
    if (NT) __builtin_nontemporal_store(v, dst);
    else *dst = v;
}

// One column tile: threads (lane, g) = (t % C, t / C) load nodes g*L ..
// g*L + L-1 of line `col`, wait at one barrier, store them back.
//   axis 0: lines (k, j, s) along i, one plane of ny*nz*S columns
//   axis 1: plane k, columns (i, s) along j
//   axis 2: one plane of (j, i, s) columns along k
template <typename T, int L, int NT>
__global__ void __launch_bounds__(1024) k_pattern(T* rho, int n, int C, int nseg, unsigned tiles, unsigned total,
                                                  size_t line_stride, size_t plane_stride, int axis, int S,
                                                  int nx, int spread, T scale)
{
    const unsigned w = xcd_map(blockIdx.x, total, spread);
    const unsigned plane = w / tiles, tile = w - plane * tiles;
    const int lane = (int)threadIdx.x % C, g = (int)threadIdx.x / C;
    const size_t col = (size_t)tile * C + lane;
    size_t base;
    if (axis == 0) base = col / S * (size_t)nx * S + col % S;
    else if (axis == 1) base = plane * plane_stride + col;
    else base = col;
    const int p0 = g * L;
    const bool act = g < nseg;
    const int cnt = act ? min(L, n - p0) : 0;
    T v[L];
#pragma unroll
    for (int t = 0; t < L; ++t) {
        T* p = rho + base + (size_t)(p0 + t) * line_stride;
        v[t] = t < cnt ? (NT ? __builtin_nontemporal_load(p) : *p) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < L; ++t) {
        T* p = rho + base + (size_t)(p0 + t) * line_stride;
        if (t < cnt) {
            if (NT) __builtin_nontemporal_store(v[t] * scale, p);
            else *p = v[t] * scale;
        }
    }
}

template <typename T, int L>
hipError_t launch_pattern(T* rho, int n, int S, int axis, int C, int spread, int nt, hipStream_t st)
{
    const int nseg = (n + L - 1) / L;
    const size_t ncols = axis == 0 ? (size_t)n * n * S : axis == 1 ? (size_t)n * S : (size_t)n * n * S;
    const unsigned planes = axis == 1 ? (unsigned)n : 1u;
    const unsigned tiles = (unsigned)(ncols / C);
    const size_t line_stride = axis == 0 ? (size_t)S : axis == 1 ? (size_t)n * S : (size_t)n * n * S;
    const size_t plane_stride = (size_t)n * n * S;
    const int threads = (C * nseg + 63) / 64 * 64;
    if (nt)
        k_pattern<T, L, 1><<<tiles * planes, threads, 0, st>>>(rho, n, C, nseg, tiles, tiles * planes, line_stride,
                                                              plane_stride, axis, S, n, spread, T(1));
    else
        k_pattern<T, L, 0><<<tiles * planes, threads, 0, st>>>(rho, n, C, nseg, tiles, tiles * planes, line_stride,
                                                              plane_stride, axis, S, n, spread, T(1));
    return hipGetLastError();
}

template <typename T>
hipError_t pattern_dispatch(T* rho, int n, int S, int axis, int C, int L, int spread, int nt, hipStream_t st)
{
    switch (L) {
    case 8: return launch_pattern<T, 8>(rho, n, S, axis, C, spread, nt, st);
    case 16: return launch_pattern<T, 16>(rho, n, S, axis, C, spread, nt, st);
    case 20: return launch_pattern<T, 20>(rho, n, S, axis, C, spread, nt, st);
    case 32: return launch_pattern<T, 32>(rho, n, S, axis, C, spread, nt, st);
    default: return hipErrorInvalidValue;
    }
}

// The access pattern of the whole-cache-line x sweep (k_sweep_xt): blocks of
// R rows x nseg segments, 64/S segments per wave, each lane loading 16 bytes
// of its segment's next G-node chunk (G = 16 / value size), the chunk handed
// through a per-wave LDS buffer (pitch (G+1)*S values per segment) to the
// lanes of its lines, one barrier, and the same way back to 16-byte stores.
template <typename T, int L, int S>
__global__ void __launch_bounds__(1024) k_pattern_xt(T* rho, int n, int R, int nseg, int rows, int spread, T scale)
{
    constexpr int G = 16 / (int)sizeof(T), kSegW = 64 / S, kPitch = (G + 1) * S;
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_pattern[];
    const int tid = threadIdx.x, qq = (tid & 63) / S, s = tid % S;
    T* const mine = (T*)smem_pattern + (tid >> 6) * (kSegW * kPitch) + qq * kPitch;
    const int Q = R * nseg, q = (tid >> 6) * kSegW + qq, qc = min(q, Q - 1);
    const int rl = qc / nseg, g = qc - rl * nseg, p0 = g * L;
    const int row = (int)xcd_map(blockIdx.x, gridDim.x, spread) * R + rl;
    const bool act = q < Q && row < rows;
    const size_t row0 = (size_t)min(row, rows - 1) * n * S;
    const int uj = G * s / S, oj = G * s % S, wj = uj * S + oj;
    u32x4 piece[L / G];
#pragma unroll
    for (int k = 0; k < L / G; ++k)
        piece[k] = *(const u32x4*)(rho + row0 + (size_t)min(p0 + G * k + uj, n - 1) * S + oj);
    T v[L];
#pragma unroll
    for (int k = 0; k < L / G; ++k) {
        *(u32x4*)(mine + wj) = piece[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#pragma unroll
        for (int u = 0; u < G; ++u) v[G * k + u] = mine[u * S + s];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < L / G; ++k) {
#pragma unroll
        for (int u = 0; u < G; ++u) mine[u * S + s] = v[G * k + u] * scale;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        const u32x4 out = *(const u32x4*)(mine + wj);
        if (act && p0 + G * k + uj < n) *(u32x4*)(rho + row0 + (size_t)(p0 + G * k + uj) * S + oj) = out;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}

template <typename T, int L>
hipError_t launch_pattern_xt(T* rho, int n, int C, int spread, hipStream_t st)
{
    constexpr int S = 8, G = 16 / (int)sizeof(T);
    const int R = C / S, nseg = (n + L - 1) / L, rows = n * n;
    const int waves = (R * nseg + 64 / S - 1) / (64 / S);
    const size_t lds = (size_t)waves * 64 * (G + 1) * sizeof(T);
    const unsigned blocks = (unsigned)((rows + R - 1) / R);
    k_pattern_xt<T, L, S><<<blocks, waves * 64, lds, st>>>(rho, n, R, nseg, rows, spread, T(1));
    return hipGetLastError();
}

template <typename T>
hipError_t pattern_xt_dispatch(T* rho, int n, int C, int L, int spread, hipStream_t st)
{
    switch (L) {
    case 8: return launch_pattern_xt<T, 8>(rho, n, C, spread, st);
    case 16: return launch_pattern_xt<T, 16>(rho, n, C, spread, st);
    case 20: return launch_pattern_xt<T, 20>(rho, n, C, spread, st);
    default: return hipErrorInvalidValue;
    }
}

// Runs `launch` once untimed, then `reps` times between two events; average ms.
template <typename F>
int time_reps(F&& launch, int reps, double* ms)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return 1;
    if (hipEventCreate(&b) != hipSuccess) return 1;
    hipError_t st = launch();
    if (st == hipSuccess) st = hipEventRecord(a, nullptr);
    for (int r = 0; r < reps && st == hipSuccess; ++r) st = launch();
    if (st == hipSuccess) st = hipEventRecord(b, nullptr);
    if (st == hipSuccess) st = hipEventSynchronize(b);
    float t = 0.f;
    if (st == hipSuccess) st = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / reps;
    return st == hipSuccess ? 0 : 2;
}

}  // namespace

extern "C" {

// A streaming pass of `kind` (0 in-place read+write, 1 copy, 2 read, 3 write)
// over `region_bytes` of a `buffer_bytes` buffer (a multiple of 16 bytes),
// with plain (nt = 0) or non-temporal (nt = 1) loads and stores, `reps` times
// after one untimed pass.  *gbps = bytes moved per pass (2x the region for 0
// and 1) / average pass time.  Returns 0 on success.
int bfc_stream_pass(int kind, int nt, size_t buffer_bytes, size_t region_bytes, int reps, double* gbps, double* ms)
{
    if (kind < 0 || kind > 3 || nt < 0 || nt > 1 || region_bytes > buffer_bytes || region_bytes % 16 || reps < 1)
        return 1;
    float *a = nullptr, *b = nullptr, *sink = nullptr;
    if (hipMalloc((void**)&a, buffer_bytes) != hipSuccess) return 3;
    if (kind == 1 && hipMalloc((void**)&b, buffer_bytes) != hipSuccess) {
        (void)hipFree(a);
        return 3;
    }
    (void)hipMalloc((void**)&sink, 16);
    (void)hipMemset(a, 0, buffer_bytes);
    const size_t n4 = region_bytes / 16;
    const unsigned blocks = (unsigned)((n4 + 255) / 256);
    const int rc = time_reps(
        [&] {
            if (nt) k_pass<1><<<blocks, 256>>>((f4v*)a, (f4v*)b, n4, kind, 1.0f, sink);
            else k_pass<0><<<blocks, 256>>>((f4v*)a, (f4v*)b, n4, kind, 1.0f, sink);
            return hipGetLastError();
        },
        reps, ms);
    (void)hipDeviceSynchronize();
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    const double moved = (kind <= 1 ? 2.0 : 1.0) * (double)region_bytes;
    *gbps = rc == 0 ? moved / (*ms * 1e-3) / 1e9 : 0.0;
    return rc;
}

// The access pattern of one sweep on an n^3 grid of S substrates (elem = 4
// for an FP32 store, 8 for FP64): C lines per block, L in {8, 16, 20, 32} nodes
// per thread, XCD spread and cache policy (nt = 1: non-temporal loads and
// stores) as the sweeps take them.  *gbps = 2 x the store's bytes / average
// launch time.
int bfc_sweep_pattern(int elem, int n, int S, int axis, int C, int L, int spread, int nt, int reps, double* gbps,
                      double* ms)
{
    if ((elem != 4 && elem != 8) || axis < 0 || axis > 2 || C <= 0 || C * ((n + L - 1) / L) > 1024 || reps < 1 ||
        nt < 0 || nt > 1)
        return 1;
    const size_t ncols = axis == 1 ? (size_t)n * S : (size_t)n * n * S;
    if (ncols % (size_t)C) return 1;
    const size_t bytes = (size_t)n * n * n * S * elem;
    void* rho = nullptr;
    if (hipMalloc(&rho, bytes) != hipSuccess) return 3;
    (void)hipMemset(rho, 0, bytes);
    const int rc = time_reps(
        [&] {
            return elem == 4 ? pattern_dispatch<float>((float*)rho, n, S, axis, C, L, spread, nt, nullptr)
                             : pattern_dispatch<double>((double*)rho, n, S, axis, C, L, spread, nt, nullptr);
        },
        reps, ms);
    (void)hipDeviceSynchronize();
    (void)hipFree(rho);
    *gbps = rc == 0 ? 2.0 * (double)bytes / (*ms * 1e-3) / 1e9 : 0.0;
    return rc;
}

// The access pattern of the whole-cache-line x sweep on an n^3 grid of 8
// substrates: C = 8R lines per block, L in {8, 16, 20} nodes per segment.
int bfc_sweep_pattern_xt(int elem, int n, int C, int L, int spread, int reps, double* gbps, double* ms)
{
    constexpr int S = 8;
    if ((elem != 4 && elem != 8) || C <= 0 || C % S || reps < 1 || n < 2) return 1;
    const int nseg = (n + L - 1) / L;
    if ((C / S) * nseg * S > 1024) return 1;
    const size_t bytes = (size_t)n * n * n * S * elem;
    void* rho = nullptr;
    if (hipMalloc(&rho, bytes) != hipSuccess) return 3;
    (void)hipMemset(rho, 0, bytes);
    const int rc = time_reps(
        [&] {
            return elem == 4 ? pattern_xt_dispatch<float>((float*)rho, n, C, L, spread, nullptr)
                             : pattern_xt_dispatch<double>((double*)rho, n, C, L, spread, nullptr);
        },
        reps, ms);
    (void)hipDeviceSynchronize();
    (void)hipFree(rho);
    *gbps = rc == 0 ? 2.0 * (double)bytes / (*ms * 1e-3) / 1e9 : 0.0;
    return rc;
}

}  // extern "C"
