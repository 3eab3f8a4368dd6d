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
//   k_sweep        all three sweeps: 16-64 lines per block, up to 1024/lines
//                  segments of 4-32 nodes per line, one block per column tile.
//   k_sweep_stream the same solve in resident blocks that walk the tiles and
//                  load the next tile while solving the current one.
//   k_sweep_batch  one launch over the tiles of many grids (ensembles).
//   k_line_sweep   one thread per whole line with the reference's own
//                  recurrences, for lines of one node or longer than the
//                  segmented kernels hold (n > 64*32).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "bfvm_internal.h"

namespace bfvm {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxThreads = 1024;  // block = COLS columns x nseg segments
// g0/g1 rows and K staged per block (dynamic LDS beside the kernels' static
// 16-17 KiB, or 32 KiB for the ensemble kernel): up to 128 KiB, so grids
// with many substrates keep the segmented kernels (the tables grow with S
// times the longest correction, ~130 rows for the fastest diffusion here)
constexpr size_t kMaxCorrectionLds = 128 * 1024;
constexpr size_t kMaxBatchCorrectionLds = 112 * 1024;
constexpr int kPad = 32;  // zero rows around the correction tables (>= the largest L)

// Raw buffer access: a per-lane byte offset plus a uniform (scalar) one, no
// per-element address arithmetic.  Reads past the resource's end return 0
// and writes there are dropped.  The descriptor's last word is the raw
// 32-bit-element format used for gfx9-class buffers.
constexpr int kRsrcWord3 = 0x00020000;
// Cache policy bits of the loads and stores of the store (gfx950: bit 1 = NT,
// streaming).  The y and z sweeps read and write each value once, in whole
// cache lines, and can run faster with non-temporal loads and stores (at
// 400^3 x 8: FP32 y 4.5 %, z 3.7 %, and the agent kernel after them 15-25 %,
// the sweeps no longer sweeping the caches clean of its data; the FP64 step
// 7 %; at 200^3 x 4 FP64 the z sweep was 20 % slower), so the autotuner
// times both (SweepPlan::nt).  The x sweep, whose lanes read 4-64-byte
// pieces of lines that later loads of the same lines reuse, ran 8-65 %
// slower with them and keeps the default.
constexpr int kLoadPolicy = 0;
constexpr int kStorePolicy = 0;
constexpr int kStreamingPolicy = 2;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, size_t bytes)
{
    const int nrec = bytes >= 0xffffffffull ? -1 : (int)(unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nrec, kRsrcWord3);
}

template <int POL = kLoadPolicy>
__device__ __forceinline__ float buf_load(float*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, POL));
}
template <int POL = kLoadPolicy>
__device__ __forceinline__ double buf_load(double*, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, POL));
}
template <int POL = kStorePolicy>
__device__ __forceinline__ void buf_store(float*, __amdgpu_buffer_rsrc_t r, double x, unsigned vo,
                                          unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)x), r, vo, so, POL);
}
template <int POL = kStorePolicy>
__device__ __forceinline__ void buf_store(double*, __amdgpu_buffer_rsrc_t r, double x, unsigned vo,
                                          unsigned so)
{
    using u64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u64, x), r, vo, so, POL);
}

// Block barrier that orders LDS only: waits for this wave's LDS accesses,
// not for its global loads, so loads issued before it (the next tile's
// values) stay in flight across it.  Threads exchange no global data.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroups are dealt to the XCDs round robin (workgroup i to XCD i % 8),
// each XCD with its own L2.  Renumbering the tiles so that one XCD runs
// neighbouring column tiles, which share DRAM pages, at the same time is
// worth +7 % on the y-sweep pattern and +15 % on z with 32-column tiles.
// spread = 0: tile i as dispatched; k > 0: in each round of 8k tiles, XCD x
// runs tiles x*k .. x*k + k-1 (the last partial round as dispatched);
// -1: XCD x runs the x-th of 8 contiguous ranges.  A bijection of [0, N).
constexpr unsigned kXcds = 8;

__device__ __forceinline__ unsigned xcd_tile(unsigned lin, unsigned N, int spread)
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

// Carry chain of one column across its nseg segments, run by one thread:
// slot u*C + lane holds alpha_u on entry and the carry into segment u on
// exit, for c' = alpha_u + aL * c walked from segment 0 up (forward) or from
// the top down (backward); aL = a^L is the same for every segment (the
// backward step out of the last segment multiplies a zero carry).  Returns
// the final carry.
__device__ __forceinline__ double chain(double* x, double aL, int lane, int C, int nseg, bool forward)
{
    double C0 = 0.0;
    if (forward) {
#pragma unroll 4
        for (int u = 0; u < nseg; ++u) {
            const double al = x[u * C + lane];
            x[u * C + lane] = C0;
            C0 = fma(aL, C0, al);
        }
    } else {
#pragma unroll 4
        for (int u = nseg - 1; u >= 0; --u) {
            const double al = x[u * C + lane];
            x[u * C + lane] = C0;
            C0 = fma(aL, C0, al);
        }
    }
    return C0;
}

// One kernel for the three sweeps.  The lines of a sweep are "columns": a
// column is one (voxel of a plane, substrate) pair, col = q*S + s, whose line
// starts at rho[plane*plane_stride + q*vstride + s] and advances line_stride
// per node.
//   x: one plane (a run of z planes), q = k*ny + j, vstride = nx*S, line_stride = S
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
// past n-1 are clamped reads and cannot reach node n-1 because b is 0 there.
// No loop reads a table.  Barriers order LDS only (lds_barrier).

// Where one thread's segment of one (column tile, plane) lies.
// C = lines per block (a power of two), tid -> column tid % C, segment tid / C.
struct SegPos {
    size_t plane0;              // first value of the plane (the buffer runs from there to the end)
    unsigned vo;                // byte offset of node p0 of this thread's line
    int p0, s;                  // first node, substrate
    int slot;                   // LDS slot of the segment: g * C + lane
    bool act;                   // owns a real (line, segment), not a copy of one
    bool last_writer;           // records w at node n-1 for its lane
};

// Threads past the plane's last column work on that column, and threads past
// the last segment on that segment: every thread computes a real segment, so
// a copy stores exactly the values (same inputs, same instructions) that the
// real owner stores.
template <typename T, int L>
__device__ __forceinline__ SegPos seg_pos(const SweepArgs& A, int C, int tile, int plane)
{
    const int tid = threadIdx.x;
    const int lane = tid & (C - 1);
    const int graw = tid / C;
    const int craw = tile * C + lane;
    const int col = min(craw, A.cols_per_plane - 1);
    const int g = min(graw, A.nseg - 1);
    SegPos P;
    P.act = craw < A.cols_per_plane && graw < A.nseg;
    P.last_writer = graw == A.nseg - 1;
    P.s = col % A.S;
    P.p0 = g * L;
    P.slot = g * C + lane;
    // plane is the same in every lane; say so, so the buffer descriptor
    // built from it stays in scalar registers
    P.plane0 = (size_t)__builtin_amdgcn_readfirstlane(plane) * A.plane_stride;
    unsigned node0 = (unsigned)P.p0 * A.line_stride;  // values from the plane start to node p0
    if (A.seg_base) {
        // C = 64: the wave is one segment, p0 the same in every lane; the
        // descriptor starts at node p0 and offsets stay below L+1 nodes
        P.plane0 += (size_t)__builtin_amdgcn_readfirstlane(P.p0) * A.line_stride;
        node0 = 0;
    }
    P.vo = ((unsigned)(col / A.S) * A.vstride + (unsigned)P.s + node0) * (unsigned)sizeof(T);
    return P;
}

// The segment's L raw values.  The segment holding node n-1 clamps its reads
// to node n-1 (every address stays in the store); its nodes past n-1 are
// never stored and cannot reach node n-1 because b is 0 there.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const SweepArgs& A, const SegPos& P)
{
    return make_rsrc((T*)A.rho + P.plane0, (A.nvs - P.plane0) * sizeof(T));
}
// The same range of the output store (rho itself for an in-place sweep).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc_out(const SweepArgs& A, const SegPos& P)
{
    return make_rsrc((T*)A.out + P.plane0, (A.nvs - P.plane0) * sizeof(T));
}

// STREAM: one load per node for every lane with a per-lane offset and no
// branch, so a wait placed before the arithmetic can count every load issued
// after these (the prefetch of the streaming kernel).  Otherwise the
// segments below the last one use a scalar offset per node (fewer VGPRs) and
// only the last segment computes clamped offsets.
template <typename T, int L, bool STREAM, int POL = kLoadPolicy>
__device__ __forceinline__ void seg_load(const SweepArgs& A, const SegPos& P, T (&raw)[L])
{
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc<T>(A, P);
    const unsigned step = A.line_stride * (unsigned)sizeof(T);
    const int last = min(L, A.n - P.p0) - 1;  // nodes past n-1 read node n-1
    if (STREAM || P.p0 + L >= A.n) {
#pragma unroll
        for (int t = 0; t < L; ++t)
            raw[t] = buf_load<POL>((T*)nullptr, rs, P.vo + (unsigned)min(t, last) * step, 0);
    } else {
#pragma unroll
        for (int t = 0; t < L; ++t) raw[t] = buf_load<POL>((T*)nullptr, rs, P.vo, t * step);
    }
}

// The correction tables staged in LDS, zero padded so no index clamps:
//   G0 rows 0..kg0-1 = g0, then kPad zero rows
//   G1 kPad zero rows, rows kPad+j = g1 at node n-1-j (j < kg1), kPad zero rows
//   K  = wb (8 values per substrate)
// Read only after a barrier.  Each region is a contiguous run of values, so
// value i of the staged array comes from (its region's array)[i - start]
// (no per-value division by S).  The staging costs a 400^3 x 8 sweep 2-3 %
// (timed without it): the table's L2 reads, about 17 KB per block; copying
// it by direct-to-LDS loads instead, before or after the tile's loads, was
// no faster (x 2.5 % slower, z 1 % faster).
struct TableRuns {
    int e0, b1, e1, ng, total;  // end of G0, first G1 value, end of G1, first K value, end
};

__device__ __forceinline__ TableRuns table_runs(const SweepArgs& A)
{
    TableRuns t;
    t.e0 = A.kg0 * A.S;
    t.b1 = (A.kg0 + 2 * kPad) * A.S;
    t.e1 = t.b1 + A.kg1 * A.S;
    t.ng = t.e1 + kPad * A.S;
    t.total = t.ng + 8 * A.S;
    return t;
}

// Where staged value i comes from (nullptr: a zero row or past the end).
__device__ __forceinline__ const double* table_src(const SweepArgs& A, const TableRuns& t, int i)
{
    if (i < t.e0) return A.g0 + i;
    if (i >= t.b1 && i < t.e1) return A.g1 + (i - t.b1);
    if (i >= t.ng && i < t.total) return A.wb + (i - t.ng);
    return nullptr;
}

__device__ __forceinline__ void stage_tables(const SweepArgs& A, double* gtab)
{
    const TableRuns t = table_runs(A);
    for (int i = threadIdx.x; i < t.total; i += blockDim.x) {
        const double* src = table_src(A, t, i);
        gtab[i] = src ? *src : 0.0;
    }
}

// The solve of one tile from its raw values, and the stores; r = 1/d and
// a = c1/d of the thread's substrate.  Every thread of the block calls it (it
// holds barriers); raw store values are widened to FP64 one by one in the
// forward pass.  xa holds the forward partials and then the forward carries,
// xb the backward ones: a copy thread reads its segment's carries while the
// owner goes on writing, so the two phases do not share slots.
// With STREAM, every thread stores every node, nodes past n-1 to node n-1
// (copies store the owner's values; the true value of node n-1 is stored
// after the others by the same lane), so no store sits behind a lane mask;
// otherwise only owners store, and only their real nodes.
template <typename T, int L, bool STREAM, int POL = kStorePolicy>
__device__ __forceinline__ void seg_solve(const SweepArgs& A, int C, const SegPos& P, const T (&raw)[L],
                                          double r, double a, double* xa, double* xb, double* wfirst,
                                          double* wlast, const double* gtab)
{
    const int n = A.n, nseg = A.nseg, S = A.S, kg0 = A.kg0, kg1 = A.kg1;
    const int tid = threadIdx.x;
    const int lane = tid & (C - 1);
    const int s = P.s, p0 = P.p0;
    const int cnt = min(L, n - p0);     // real nodes of the segment (>= 1)
    const bool has_last = p0 + L >= n;  // holds node n-1 (and nodes past it)
    const double* K = gtab + (kg0 + kg1 + 3 * kPad) * S;  // read after the first barrier
    const unsigned step = A.line_stride * (unsigned)sizeof(T);

    // forward, zero carry-in
    double v[L];
    double y = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) v[t] = y = fma(a, y, (double)raw[t] * r);
    double aL = a;
#pragma unroll
    for (int t = 1; t < L; ++t) aL *= a;
    xa[tid] = y;
    lds_barrier();
    if (tid < C) chain(xa, aL, lane, C, nseg, true);
    lds_barrier();
    {
        double PY = xa[P.slot];
#pragma unroll
        for (int t = 0; t < L; ++t) {
            PY *= a;
            v[t] += PY;
        }
    }
    if (has_last && P.last_writer) {  // z_{n-1} = y_{n-1}
        double wl = 0.0;
#pragma unroll
        for (int t = 0; t < L; ++t)
            if (p0 + t == n - 1) wl = v[t];
        wlast[lane] = wl;
    }

    // backward, zero carry-in from above
    double z = 0.0;
    if (!has_last) {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(a, z, v[t]);
    } else {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(p0 + t < n - 1 ? a : 0.0, z, v[t]);
    }
    xb[tid] = z;
    lds_barrier();
    if (tid < C) wfirst[lane] = chain(xb, aL, lane, C, nseg, false);
    lds_barrier();

    // Last pass, from the top of the segment down, each node finished and
    // stored as soon as it is final: the carry from above (a^(q-i) Z), the
    // rank-2 correction near the line ends (mu0 g0_i + mu1 g1_i), store.
    const double QZ0 = has_last ? 0.0 : xb[P.slot];  // the last segment has nothing above
    const bool head = p0 < kg0, tail = p0 + L > n - kg1;
    double mu0 = 0.0, mu1 = 0.0;
    if (head || tail) {
        const double w0 = wfirst[lane], wn = wlast[lane];
        const double* k = K + s * 8 + 2;  // K00, K01, K10, K11
        mu0 = head ? k[0] * w0 + k[1] * wn : 0.0;
        mu1 = tail ? k[2] * w0 + k[3] * wn : 0.0;
    }
    // g0 rows from node p0 up; g1 rows from node p0 down (zero padded)
    const double* G0 = gtab;
    const double* G1 = gtab + (kg0 + kPad) * S;
    const double* g0r = G0 + (head ? p0 : 0) * S + s;
    const double* g1r = G1 + (tail ? kPad + n - 1 - p0 : 0) * S + s;
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc_out<T>(A, P);
    double QZ = QZ0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
        QZ *= a;
        double zf = v[t] + QZ;
        if (head) zf = fma(mu0, g0r[t * S], zf);
        if (tail) zf = fma(mu1, g1r[-t * S], zf);
        if (STREAM) buf_store<POL>((T*)nullptr, rs, zf, P.vo + (unsigned)min(t, cnt - 1) * step, 0);
        else if (P.act && t < cnt) buf_store<POL>((T*)nullptr, rs, zf, P.vo, t * step);
    }
}

// The same arithmetic as seg_solve (the same operations in the same order,
// so the same bits) with the values' source and destination left to the
// caller: lane = the segment's column in the block (0..C-1, C any count), s
// its substrate, p0 its first node, slot = g*C + lane its carry slot; gtab =
// the correction tables and K in stage_tables' layout; src(t, x) turns raw
// value t (widened to FP64) into the value the sweep solves for (sources
// applied before the sweep), in node order; sink(t, z) receives the final
// value of node p0 + t, from the top of the segment down.
struct NoSource {
    __device__ __forceinline__ double operator()(int, double x) const { return x; }
};

// xslot = the LDS slot this thread writes its partials to (its own slot, or
// the slot of the segment it copies: the same value); walker = this thread
// walks the carry chain of column `lane` (one thread per column).
template <typename T, int L, class Src, class Sink>
__device__ __forceinline__ void seg_core_at(const SweepArgs& A, int C, int lane, int s, int p0, int slot,
                                            bool last_writer, int xslot, bool walker, const T (&raw)[L],
                                            double r, double a, double* xa, double* xb, double* wfirst,
                                            double* wlast, const double* gtab, Src&& src, Sink&& sink)
{
    const int n = A.n, nseg = A.nseg, S = A.S, kg0 = A.kg0, kg1 = A.kg1;
    const bool has_last = p0 + L >= n;  // holds node n-1 (and nodes past it)
    const double* K = gtab + (kg0 + kg1 + 3 * kPad) * S;  // read after the first barrier

    // forward, zero carry-in
    double v[L];
    double y = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t) v[t] = y = fma(a, y, src(t, (double)raw[t]) * r);
    double aL = a;
#pragma unroll
    for (int t = 1; t < L; ++t) aL *= a;
    xa[xslot] = y;
    lds_barrier();
    if (walker) chain(xa, aL, lane, C, nseg, true);
    lds_barrier();
    {
        double PY = xa[slot];
#pragma unroll
        for (int t = 0; t < L; ++t) {
            PY *= a;
            v[t] += PY;
        }
    }
    if (has_last && last_writer) {  // z_{n-1} = y_{n-1}
        double wl = 0.0;
#pragma unroll
        for (int t = 0; t < L; ++t)
            if (p0 + t == n - 1) wl = v[t];
        wlast[lane] = wl;
    }

    // backward, zero carry-in from above
    double z = 0.0;
    if (!has_last) {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(a, z, v[t]);
    } else {
#pragma unroll
        for (int t = L - 1; t >= 0; --t) v[t] = z = fma(p0 + t < n - 1 ? a : 0.0, z, v[t]);
    }
    xb[xslot] = z;
    lds_barrier();
    if (walker) wfirst[lane] = chain(xb, aL, lane, C, nseg, false);
    lds_barrier();

    // Last pass, from the top of the segment down, each node finished and
    // stored as soon as it is final: the carry from above (a^(q-i) Z), the
    // rank-2 correction near the line ends (mu0 g0_i + mu1 g1_i), store.
    const double QZ0 = has_last ? 0.0 : xb[slot];  // the last segment has nothing above
    const bool head = p0 < kg0, tail = p0 + L > n - kg1;
    double mu0 = 0.0, mu1 = 0.0;
    if (head || tail) {
        const double w0 = wfirst[lane], wn = wlast[lane];
        const double* k = K + s * 8 + 2;  // K00, K01, K10, K11
        mu0 = head ? k[0] * w0 + k[1] * wn : 0.0;
        mu1 = tail ? k[2] * w0 + k[3] * wn : 0.0;
    }
    // g0 rows from node p0 up; g1 rows from node p0 down (zero padded)
    const double* G0 = gtab;
    const double* G1 = gtab + (kg0 + kPad) * S;
    const double* g0r = G0 + (head ? p0 : 0) * S + s;
    const double* g1r = G1 + (tail ? kPad + n - 1 - p0 : 0) * S + s;
    double QZ = QZ0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
        QZ *= a;
        double zf = v[t] + QZ;
        if (head) zf = fma(mu0, g0r[t * S], zf);
        if (tail) zf = fma(mu1, g1r[-t * S], zf);
        sink(t, zf);
    }
}

// Thread tid = segment * C + column (k_sweep's layout): partials in slot tid,
// threads 0..C-1 walk the chains.
template <typename T, int L, class Src, class Sink>
__device__ __forceinline__ void seg_core(const SweepArgs& A, int C, int lane, int s, int p0, int slot,
                                         bool last_writer, const T (&raw)[L], double r, double a, double* xa,
                                         double* xb, double* wfirst, double* wlast, const double* gtab,
                                         Src&& src, Sink&& sink)
{
    const int tid = threadIdx.x;
    seg_core_at<T, L>(A, C, lane, s, p0, slot, last_writer, tid, tid < C, raw, r, a, xa, xb, wfirst, wlast, gtab,
                      src, sink);
}

// Block-level body of the segmented sweep for one plane's column tile.
template <typename T, int L, int POL = 0>
__device__ __forceinline__ void sweep_tile(const SweepArgs& A, int C, int tile, int plane,
                                           double* xa, double* xb, double* wfirst, double* wlast,
                                           double* gtab)
{
    const SegPos P = seg_pos<T, L>(A, C, tile, plane);
    const double r = A.wb[P.s * 8 + 0], a = A.wb[P.s * 8 + 1];
    T raw[L];
    seg_load<T, L, false, POL>(A, P, raw);
    stage_tables(A, gtab);  // while the loads are in flight
    seg_solve<T, L, false, POL>(A, C, P, raw, r, a, xa, xb, wfirst, wlast, gtab);
}

// One grid: block (x, y) = (column tile, plane), COLS lines per block; POL =
// the cache policy of the store's loads and stores.
template <typename T, int L, int COLS, int POL>
__global__ void __launch_bounds__(kMaxThreads, L <= 20 ? 8 : 4) k_sweep(SweepArgs A)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];  // w at node 0 of each column
    __shared__ double wlast[COLS];   // w at node n-1 of each column
    extern __shared__ double gtab[];
    const unsigned w =
        xcd_tile(blockIdx.x + blockIdx.y * (unsigned)A.tiles, (unsigned)(A.tiles * A.planes), A.spread);
    const unsigned plane = w / (unsigned)A.tiles;
    sweep_tile<T, L, POL>(A, COLS, (int)(w - plane * (unsigned)A.tiles), (int)plane, xa, xb, wfirst, wlast,
                          gtab);
}

// One grid, resident blocks: block b solves tiles b, b + gridDim.x, ... (tile
// w = plane * tiles + column tile) and loads tile w + gridDim.x while it
// solves tile w, so the loads of one tile overlap the arithmetic, barriers
// and stores of the previous one.  The tables are staged once per block and
// r, a come from LDS, so no global load is issued behind the prefetch (the
// vector memory counter is in order: waiting for a later load would wait for
// the prefetch too).  nblocks = gridDim.x <= tiles * planes, passed as an
// argument so it sits in a scalar register.
template <typename T, int L, int COLS, int POL>
__global__ void __launch_bounds__(kMaxThreads, 4) k_sweep_stream(SweepArgs A, int nblocks)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];
    __shared__ double wlast[COLS];
    extern __shared__ double gtab[];
    const int total = A.tiles * A.planes;
    int w = blockIdx.x;  // the same in every lane, as is everything derived from it
    T raw[L];
    int wr = (int)xcd_tile((unsigned)w, (unsigned)total, A.spread);  // tile w, renumbered per XCD
    seg_load<T, L, true, POL>(A, seg_pos<T, L>(A, COLS, wr % A.tiles, wr / A.tiles), raw);
    stage_tables(A, gtab);
    lds_barrier();
    const double* K = gtab + (A.kg0 + A.kg1 + 3 * kPad) * A.S;
    for (;;) {
        const int wn = w + nblocks;
        const bool more = wn < total;
        // the last pass loads its own tile again rather than branch around
        // the loads: with no load behind a branch, the waits before the
        // arithmetic below count the prefetch as in flight
        const int wl = (int)xcd_tile((unsigned)(more ? wn : w), (unsigned)total, A.spread);
        T nxt[L];
        seg_load<T, L, true, POL>(A, seg_pos<T, L>(A, COLS, wl % A.tiles, wl / A.tiles), nxt);
        const SegPos P = seg_pos<T, L>(A, COLS, wr % A.tiles, wr / A.tiles);
        seg_solve<T, L, true, POL>(A, COLS, P, raw, K[P.s * 8 + 0], K[P.s * 8 + 1], xa, xb, wfirst, wlast,
                                   gtab);
        if (!more) break;
#pragma unroll
        for (int t = 0; t < L; ++t) raw[t] = nxt[t];
        w = wn;
        wr = wl;
    }
}

// An ensemble: member m owns a contiguous range of blocks of a 1-D grid, laid
// out as planes x column tiles of A[m].cols lines; map[2b], map[2b+1] = the
// member of block b and b's index within it.  Every block has kMaxThreads
// threads (A[m].cols * A[m].nseg of them active).
template <typename T, int L, int POL>
__global__ void __launch_bounds__(kMaxThreads, L <= 16 ? 8 : 4)
k_sweep_batch(const SweepArgs* __restrict__ args, const int* __restrict__ map, int nm, int spread)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[kMaxThreads];
    __shared__ double wlast[kMaxThreads];
    extern __shared__ double gtab[];
    const unsigned b = xcd_tile(blockIdx.x, gridDim.x, spread);
    const int member = __builtin_amdgcn_readfirstlane(map[2 * b]);
    const int local = __builtin_amdgcn_readfirstlane(map[2 * b + 1]);
    const SweepArgs A = args[member];
    sweep_tile<T, L, POL>(A, A.cols, local % A.tiles, local / A.tiles, xa, xb, wfirst, wlast, gtab);
}

// The x sweep with sources fused in (launch_sweep_x_fused).  The reference's
// order per step is bulk, agents, Dirichlet, x sweep, Dirichlet, y sweep ...
// (BioFVM_microenvironment.cpp:584-633, BioFVM_solvers.cpp:177-241).  An x
// line holds consecutive voxels, so the thread that loads a node for the x
// sweep can apply the updates that precede the sweep, and the store can write
// the Dirichlet value the y sweep must see:
//   bulk      x = (x + A)/B (FP64, then the store's precision, as k_bulk
//             leaves it), except on voxels with agents when the agent kernel
//             already applied it there (their bits in agent_bits);
//   Dirichlet the node's value where the voxel's bit is set and the substrate
//             is masked in: the bits of the segment's L voxels come from two
//             bitmap words, the value's index from the word's rank.
// Agents stay a kernel of their own (one thread per voxel group and
// substrate), run before this one: a voxel group's agents form a chain of
// dependent loads that one sweep thread, which owns up to L voxels, would run
// serially.  The same operations and roundings as the separate kernels: the
// step's bits do not depend on fusing.
struct FusedSources {
    const double* bulk_a;         // A, B per value, or null
    const double* bulk_b;
    size_t bulk_bytes;            // bytes of each (below 4 GiB)
    const unsigned* agent_bits;   // voxels whose bulk update is done already, or null
    const unsigned* dbits;        // Dirichlet bitmap, or null
    const int* drank;
    const void* dval;             // node values (store precision), node-major
    const unsigned char* dmask;   // per substrate
    int ny;
    int dir_after;                // also overwrite Dirichlet nodes at the stores
};

// Bits of voxels v0 .. v0+63 (bit t = voxel v0 + t) from a bitmap with a
// spare word past its last voxel.
__device__ __forceinline__ unsigned long long bit_window(const unsigned* bits, int v0)
{
    const int w0 = v0 >> 5, b0 = v0 & 31;
    return (((unsigned long long)bits[w0 + 1] << 32) | bits[w0]) >> b0;
}

constexpr int kBulkChunk = 8;  // 16 measured no faster (drop_in_200: 246 vs 242 us)

template <typename T, int L, int COLS>
__global__ void __launch_bounds__(kMaxThreads, L <= 8 ? 8 : 4) k_sweep_x_src(SweepArgs A, FusedSources F)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];
    __shared__ double wlast[COLS];
    extern __shared__ double gtab[];
    const unsigned w =
        xcd_tile(blockIdx.x + blockIdx.y * (unsigned)A.tiles, (unsigned)(A.tiles * A.planes), A.spread);
    const unsigned plane = w / (unsigned)A.tiles;
    const int tile = (int)(w - plane * (unsigned)A.tiles);
    const SegPos P = seg_pos<T, L>(A, COLS, tile, (int)plane);
    const double r = A.wb[P.s * 8 + 0], a = A.wb[P.s * 8 + 1];
    T raw[L];
    seg_load<T, L, false>(A, P, raw);
    stage_tables(A, gtab);  // while the loads are in flight

    const int S = A.S, s = P.s, p0 = P.p0;
    const int cnt = min(L, A.n - p0);  // nodes past n-1 use node n-1's sources
    const int col = min(tile * COLS + (int)(threadIdx.x & (COLS - 1)), A.cols_per_plane - 1);
    const int v0 = ((int)plane * F.ny + col / S) * A.n + p0;  // voxel of node p0

    if (F.bulk_a) {
        const unsigned long long done = F.agent_bits ? bit_window(F.agent_bits, v0) : 0ull;
        // buffer loads: node p0's byte offset per lane, then one uniform
        // offset per node (no per-node address arithmetic).  Nodes past n-1
        // keep their clamped raw value (finite, and they reach no stored
        // node), so their (A, B), which may lie past the buffer and read 0,
        // are not used.
        const __amdgpu_buffer_rsrc_t ra = make_rsrc(F.bulk_a, F.bulk_bytes);
        const __amdgpu_buffer_rsrc_t rb = make_rsrc(F.bulk_b, F.bulk_bytes);
        const unsigned vb = ((unsigned)v0 * (unsigned)S + (unsigned)s) * (unsigned)sizeof(double);
        const unsigned stepb = (unsigned)S * (unsigned)sizeof(double);
        // kBulkChunk nodes at a time, so that at most kBulkChunk pairs (A, B)
        // wait in registers
#pragma unroll
        for (int t0 = 0; t0 < L; t0 += kBulkChunk) {
            double a8[kBulkChunk], b8[kBulkChunk];
#pragma unroll
            for (int t = t0; t < t0 + kBulkChunk && t < L; ++t) {
                a8[t - t0] = buf_load((double*)nullptr, ra, vb, (unsigned)t * stepb);
                b8[t - t0] = buf_load((double*)nullptr, rb, vb, (unsigned)t * stepb);
            }
#pragma unroll
            for (int t = t0; t < t0 + kBulkChunk && t < L; ++t)
                if (t < cnt && !((done >> t) & 1ull)) raw[t] = (T)(((double)raw[t] + a8[t - t0]) / b8[t - t0]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // Dirichlet bits of voxels v0 .. v0+L-1 and the node index of the first
    unsigned long long dwin = 0;
    int dbase = 0;
    const bool dir_on = F.dbits && F.dmask[s];
    if (dir_on) {
        dwin = bit_window(F.dbits, v0);
        dbase = F.drank[v0 >> 5] + __builtin_popcount(F.dbits[v0 >> 5] & ((1u << (v0 & 31)) - 1u));
    }
    const T* dval = (const T*)F.dval;
    auto dir_value = [&](int t) {
        return (double)dval[(size_t)(dbase + __builtin_popcountll(dwin & ((1ull << t) - 1ull))) * S + s];
    };
    auto src = [&](int t, double x) { return dir_on && ((dwin >> t) & 1ull) ? dir_value(t) : x; };
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc_out<T>(A, P);
    const unsigned step = A.line_stride * (unsigned)sizeof(T);
    const bool dir_after = dir_on && F.dir_after;
    seg_core<T, L>(A, COLS, (int)(threadIdx.x & (COLS - 1)), s, p0, P.slot, P.last_writer, raw, r, a, xa, xb,
                   wfirst, wlast, gtab, src, [&](int t, double zf) {
                       if (P.act && t < cnt) {
                           if (dir_after && ((dwin >> t) & 1ull)) zf = dir_value(t);
                           buf_store((T*)nullptr, rs, zf, P.vo, t * step);
                       }
                   });
}

// x sweep on row tiles staged in LDS.  A row is the nx*S values of one (j, k)
// pair, contiguous in the store; a tile is R consecutive rows (one contiguous
// run of R*nx*S values), loaded and stored with one value per lane per
// instruction, consecutive lanes on consecutive values.  In LDS the tile
// keeps the store's order, row r at r*P (P = X.P values, at least nx*S plus
// the nodes the last segment holds past n-1, which are zeros, and P = S mod
// 64 so that the lanes of a wave, lines r*S + s, fall in different banks):
// the thread of line (r, s), segment g finds node t at r*P + (g*L + t)*S + s.
// The solve is seg_core's (bit-identical to k_sweep with the same L).
//
// LDS: xa, xb (one double per thread), wfirst, wlast (C each), the correction
// tables (stage_tables), then the tile.

__device__ __forceinline__ void val_store(float*, __amdgpu_buffer_rsrc_t r, float x, unsigned vo)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, 0, kStorePolicy);
}
__device__ __forceinline__ void val_store(double*, __amdgpu_buffer_rsrc_t r, double x, unsigned vo)
{
    using u64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u64, x), r, vo, 0, kStorePolicy);
}

// Row tile `tile` of X (blockDim >= C*nseg threads, L*blockDim >= R*nx*S).
template <typename T, int L>
__device__ __forceinline__ void rows_tile(const RowTile& X, unsigned tile_index)
{
    extern __shared__ double smem[];
    const SweepArgs& A = X.A;
    const int nthr = blockDim.x, tid = threadIdx.x;
    double* xa = smem;
    double* xb = xa + nthr;
    double* wfirst = xb + nthr;
    double* wlast = wfirst + X.C;
    double* gtab = wlast + X.C;
    T* tile = (T*)(gtab + (A.kg0 + A.kg1 + 3 * kPad + 8) * A.S);
    const unsigned nxS = (unsigned)X.nxS, P = (unsigned)X.P;
    // value e = tid + k*nthr of the tile is (row r, offset w): walked from
    // e = tid in steps of nthr = dr rows and dw values
    const unsigned r0 = __umulhi((unsigned)tid, X.mag_nxS), w0 = (unsigned)tid - r0 * nxS;
    const unsigned dr = (unsigned)nthr / nxS, dw = (unsigned)nthr - dr * nxS;

    // the buffer ends with the tile's last real value: reads past it give 0
    // (the rows past the last are zeros) and stores past it are dropped
    const size_t first = (size_t)tile_index * X.tile_vals;
    const unsigned range = (unsigned)min((size_t)X.tile_vals, X.vals - first) * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc((T*)A.rho + first, range);
    const __amdgpu_buffer_rsrc_t rs_out = make_rsrc((T*)A.out + first, range);
    {
        T v[L];
#pragma unroll
        for (int k = 0; k < L; ++k)
            v[k] = buf_load((T*)nullptr, rs, (unsigned)(tid + k * nthr) * (unsigned)sizeof(T), 0);
        stage_tables(A, gtab);  // while the loads are in flight
        // zeros past node n-1 of each line's last segment
        const int pad = (int)(P - nxS);
        for (int q = tid; q < X.R * pad; q += nthr) {
            const int r = q / pad;
            tile[(unsigned)r * P + nxS + (unsigned)(q - r * pad)] = (T)0;
        }
        unsigned r = r0, w = w0;
#pragma unroll
        for (int k = 0; k < L; ++k) {
            if (tid + k * nthr < X.tile_vals) tile[r * P + w] = v[k];
            w += dw;
            r += dr;
            if (w >= nxS) {
                w -= nxS;
                ++r;
            }
        }
    }
    lds_barrier();

    const int lane = tid % X.C, graw = tid / X.C;
    const int g = min(graw, A.nseg - 1);  // threads past the last segment are copies of it
    const int S = A.S;
    const int s = lane % S;
    unsigned base = (unsigned)(lane / S) * P + (unsigned)(g * L * S + s);
    T raw[L];
#pragma unroll
    for (int t = 0; t < L; ++t) raw[t] = tile[base + (unsigned)(t * S)];
    const double* K = gtab + (A.kg0 + A.kg1 + 3 * kPad) * A.S;
    const bool act = graw < A.nseg;
    // the addresses are recomputed after the solve rather than kept in
    // registers across it (the compiler would otherwise hold all L of them)
    asm volatile("" : "+v"(base));
    seg_core<T, L>(A, X.C, lane, s, g * L, g * X.C + lane, graw == A.nseg - 1, raw, K[s * 8 + 0], K[s * 8 + 1],
                   xa, xb, wfirst, wlast, gtab, NoSource{}, [&](int t, double zf) {
                       if (act) tile[base + (unsigned)(t * S)] = (T)zf;
                   });
    lds_barrier();
    unsigned r = r0, w = w0;
    asm volatile("" : "+v"(r), "+v"(w));
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const unsigned e = (unsigned)(tid + k * nthr);
        if (e < (unsigned)X.tile_vals) val_store((T*)nullptr, rs_out, tile[r * P + w], e * (unsigned)sizeof(T));
        w += dw;
        r += dr;
        if (w >= nxS) {
            w -= nxS;
            ++r;
        }
    }
}

// Grid: one block per row tile.
template <typename T, int L>
__global__ void __launch_bounds__(kMaxThreads, L <= 8 ? 8 : 4) k_sweep_rows(RowTile X)
{
    rows_tile<T, L>(X, xcd_tile(blockIdx.x, gridDim.x, X.A.spread));
}

// Row tiles of several grids' x sweeps: map[2b], map[2b+1] = the grid of block
// b and b's tile within it; blockDim and the LDS fit every grid's tiles.
template <typename T, int L>
__global__ void __launch_bounds__(kMaxThreads, L <= 8 ? 8 : 4)
k_sweep_rows_batch(const RowTile* __restrict__ tabs, const int* __restrict__ map, int spread)
{
    const unsigned b = xcd_tile(blockIdx.x, gridDim.x, spread);
    const int member = __builtin_amdgcn_readfirstlane(map[2 * b]);
    const int local = __builtin_amdgcn_readfirstlane(map[2 * b + 1]);
    const RowTile X = tabs[member];
    rows_tile<T, L>(X, (unsigned)local);
}

// x sweep with whole-cache-line loads and stores (k_sweep_xt).  In the
// column kernels a lane owns one x line, so a load instruction reads an
// S-value piece of each of several rows and the next nodes of those rows
// come from the same cache lines again (through L1).  Here a load
// instruction reads whole G-node chunks (G*S contiguous values, G = 16 bytes
// / value size): lane j of a segment loads the 16 bytes j of the segment's
// next chunk, and the chunk goes through a per-wave LDS buffer to the lanes
// that own its lines (lane s takes node u of line s); the solved values go
// back the same way and are stored 16 bytes per lane.  Two forms:
//   SC = S, a compile-time 4, 8 or 16 (FP32) or 2, 4, 8 or 16 (FP64): a
//     lane's 16 bytes lie in one node, any row length;
//   SC = 0, any S up to 64 (FP32): a lane's 16 bytes may span two nodes;
//     rows of a multiple of 4 nodes, so every chunk starts 16-byte aligned
//     and lies wholly inside or wholly past the row.
//
// Block: R rows (C = R*S lines, column rl*S + s) x nseg segments; its
// Q = R*nseg segments q = rl*nseg + g are dealt to the waves, segw = 64/S
// (rounded down) per wave (wave w holds q = w*segw + qq on lanes
// qq*S .. qq*S + S-1; lanes past segw*S copy the wave's last segment), so a
// wave's load instruction covers whole chunks of segw segments.  Threads
// past segment Q-1 copy it (same loads, same values, no stores); rows past
// the last are copies of it (no stores).  The solve is seg_core's:
// bit-identical to k_sweep with the same L.
//
// LDS: the correction tables (stage_tables), the carry slots xa, xb (Q*S
// each), wfirst, wlast (C each), then one chunk buffer per wave (at an even
// double offset): segment qq's chunk at qq*P values, node u of line s at
// u*S + s in it; P = (G+1)*S for SC (one node of padding: the 64 lanes
// reading one node fall in different banks and every lane's LDS offsets are
// one register plus constants), else SweepArgs::xt_pitch, a multiple of 4
// picked on the host for few bank conflicts.

// Orders the LDS accesses of one wave: a wave's LDS operations complete in
// order, so this only keeps the compiler from moving them across it.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename T, int S>
constexpr bool xt_shape_ok(int L)
{
    if (S == 0) return sizeof(T) == 4 && L % 4 == 0;  // runtime S: FP32
    return 16 % sizeof(T) == 0 && S % (16 / (int)sizeof(T)) == 0 && 64 % S == 0 && S <= 16 &&
           L % (16 / (int)sizeof(T)) == 0;
}

// Doubles before the chunk buffers (tables, carries, w at the line ends),
// rounded up to an even count so the buffers are 16-byte aligned.
__host__ __device__ __forceinline__ int xt_buffer_offset(int kg0, int kg1, int S, int Q, int C)
{
    const int d = (kg0 + kg1 + 3 * kPad + 8) * S + 2 * Q * S + 2 * C;
    return (d + 1) & ~1;
}

// What xt_tile tells its hook about the thread's segment: row0 = values
// before its row, p0 its first node, s its substrate, slot its carry slot
// (g*C + column), g its segment in the row, q its segment in the block (Q of
// them; threads with q >= Q copy segment Q-1), act = it stores.
struct XtPos {
    unsigned row0;
    int p0, s, S, slot, g, q, Q, C, nseg, n;
    bool act;
};

// The plain sweep's hook: nothing besides the sweep.
struct XtPlain {
    __device__ __forceinline__ void begin(const XtPos&) {}
    __device__ __forceinline__ void node(int, double) {}
    __device__ __forceinline__ void end(const XtPos&, double*, double*, double*, double*) {}
};

// One block's tile of the whole-cache-line x sweep (see above).  hook.begin
// runs once the tile's loads are issued, hook.node(t, z) receives each final
// value (from the top of the segment down, before it is stored), hook.end
// runs after the solve, with every thread of the block (the carry slots xa,
// xb and the line-end values wfirst, wlast are free once it has passed a
// block barrier).
template <typename T, int L, int SC, class Hook, int POL = kLoadPolicy>
__device__ __forceinline__ void xt_tile(const SweepArgs& A, unsigned tile, double* smem, Hook& hook)
{
    constexpr int G = 16 / (int)sizeof(T);  // nodes per chunk = values per lane piece
    static_assert(xt_shape_ok<T, SC>(L), "k_sweep_xt shape");
    const int S = SC ? SC : A.S;
    const int segw = SC ? 64 / SC : 64 / S;  // segments per wave
    const int P = SC ? (G + 1) * SC : A.xt_pitch;
    const int C = A.cols, nseg = A.nseg, n = A.n;
    const int R = C / S, Q = R * nseg;
    double* gtab = smem;
    double* xa = gtab + (A.kg0 + A.kg1 + 3 * kPad + 8) * S;
    double* xb = xa + Q * S;
    double* wfirst = xb + Q * S;
    double* wlast = wfirst + C;
    const int tid = threadIdx.x, l = tid & 63;
    int qq = l / S, s = l - qq * S;
    const bool dup = !SC && qq >= segw;
    if (dup) {  // lanes past the wave's whole segments copy its last one (and walk no chain)
        qq = segw - 1;
        s = l - segw * S;
    }
    // this lane's segment's part of the wave's chunk buffer
    T* const mine =
        (T*)(smem + xt_buffer_offset(A.kg0, A.kg1, S, Q, C)) + (tid >> 6) * (segw * P) + qq * P;
    const int q = (tid >> 6) * segw + qq;
    const int qc = min(q, Q - 1);
    const int rl = qc / nseg, g = qc - rl * nseg;
    const int col = rl * S + s, p0 = g * L;
    const int rows = A.cols_per_plane / S;
    const int row = (int)tile * R + rl;
    const bool act = q < Q && row < rows;
    const unsigned row0 = (unsigned)min(row, rows - 1) * A.vstride;  // values before the row
    const int uj = G * s / S;  // node of the chunk where this lane's piece starts
    const int wj = G * s;      // the piece's place in the chunk buffer
    // byte offset of this lane's piece of chunk 0; chunk k is k*G*S values on
    const unsigned chunk = (unsigned)(G * S * sizeof(T));
    const unsigned vo = (row0 + (unsigned)p0 * S + wj) * (unsigned)sizeof(T);

    const double r = A.wb[s * 8 + 0], a = A.wb[s * 8 + 1];
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(A.rho, A.nvs * sizeof(T));
    u32x4 piece[L / G];
    if (p0 + L <= n) {
#pragma unroll
        for (int k = 0; k < L / G; ++k)
            piece[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, vo, k * chunk, POL);
    } else if (SC) {  // the segment holding node n-1: nodes past it read node n-1
        const int oj = G * s - uj * S;
#pragma unroll
        for (int k = 0; k < L / G; ++k) {
            const unsigned node = (unsigned)min(p0 + G * k + uj, n - 1);
            piece[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, (row0 + node * S + oj) * (unsigned)sizeof(T), 0,
                                                             POL);
        }
    } else {  // rows of whole chunks: chunks past node n-1 read the row's last chunk
#pragma unroll
        for (int k = 0; k < L / G; ++k) {
            const int kk = min(k, (n - p0) / G - 1);
            piece[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, vo, kk * chunk, POL);
        }
    }
    const int slot = g * C + col;
    const XtPos pos{row0, p0, s, S, slot, g, q, Q, C, nseg, n, act && !dup};
    hook.begin(pos);
    stage_tables(A, gtab);  // while the loads are in flight
    T raw[L];
#pragma unroll
    for (int k = 0; k < L / G; ++k) {
        *(u32x4*)(mine + wj) = piece[k];
        wave_lds_sync();
#pragma unroll
        for (int u = 0; u < G; ++u) raw[G * k + u] = mine[u * S + s];
        wave_lds_sync();
    }

    const __amdgpu_buffer_rsrc_t ro = make_rsrc(A.out, A.nvs * sizeof(T));
    // a chunk is stored when its lane's piece lies in the row (SC: the
    // piece's node; SC = 0: the chunk's first node, chunks being whole)
    const int store_node = SC ? p0 + uj : p0;
    seg_core_at<T, L>(A, C, col, s, p0, slot, g == nseg - 1, slot, q < Q && g == 0 && !dup, raw, r, a, xa, xb,
                      wfirst, wlast, gtab, NoSource{}, [&](int t, double zf) {
                          hook.node(t, zf);
                          mine[(t % G) * S + s] = (T)zf;
                          if (t % G == 0) {  // chunk t / G is complete
                              wave_lds_sync();
                              const u32x4 out = *(const u32x4*)(mine + wj);
                              if (act && !dup && store_node + t < n)
                                  __builtin_amdgcn_raw_buffer_store_b128(out, ro, vo, (t / G) * chunk,
                                                                         POL);
                              wave_lds_sync();
                          }
                      });
    hook.end(pos, xa, xb, wfirst, wlast);
}

// POL: the cache policy of the store's loads and stores (SweepPlan::nt).
template <typename T, int L, int SC, int POL>
__global__ void __launch_bounds__(kMaxThreads, L <= 20 ? 8 : 4) k_sweep_xt(SweepArgs A)
{
    extern __shared__ __attribute__((aligned(16))) double smem_xt[];
    XtPlain plain;
    xt_tile<T, L, SC, XtPlain, POL>(A, xcd_tile(blockIdx.x, gridDim.x, A.spread), smem_xt, plain);
}

// Whole-cache-line x sweeps of several grids (runtime S): map[2b], map[2b+1] =
// the grid of block b and b's tile within it; blockDim and the LDS fit every
// grid's tiles (a grid needing fewer waves has its extra ones copy).
template <typename T, int L>
__global__ void __launch_bounds__(kMaxThreads, 8)
k_sweep_xt_batch(const SweepArgs* __restrict__ args, const int* __restrict__ map, int spread)
{
    extern __shared__ __attribute__((aligned(16))) double smem_xtb[];
    const unsigned b = xcd_tile(blockIdx.x, gridDim.x, spread);
    const int member = __builtin_amdgcn_readfirstlane(map[2 * b]);
    const int local = __builtin_amdgcn_readfirstlane(map[2 * b + 1]);
    const SweepArgs A = args[member];
    XtPlain plain;
    xt_tile<T, L, 0>(A, (unsigned)local, smem_xtb, plain);
}

// The adjoint's sweep of one axis with that axis's D/lambda sums fused in
// (k_sweep_adjoint_sums).  The sweep is the forward one (the sweep matrix A
// is symmetric) on the FP64 adjoint state, one block per column tile; as
// each node's final value l_i is stored, the thread adds, for its
// substrate, l_i y_i and, over the line's edges, (l_i - l_{i+1})(y_i - y_{i+1})
// (y = the forward sweep's output, store precision, loaded beside the
// adjoint values).  The edge between a segment's last node and the next
// segment's first goes through LDS.  The block's sums are added up in a
// fixed order (segments of a column, then the columns of a substrate) onto
// the tile's slot of `partial` (per chunk of kGradChunk substrates: tiles x
// 2*kGradChunk doubles, edge sums then l.y sums), the same every run.
constexpr int kGradChunk = kAdjointSumsChunk;

// ea += a[first + i*stride], eb += b[first + i*stride] for i = 0..n-1, in
// that order (the same bits as a plain loop), the LDS loads issued eight at
// a time so that their latencies overlap.
__device__ __forceinline__ void ordered_sums(const double* a, const double* b, int first, int n, int stride,
                                             double& ea, double& eb)
{
    int i = 0;
#pragma unroll 1
    for (; i + 8 <= n; i += 8) {
        double ta[8], tb[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ta[k] = a[first + (i + k) * stride];
            tb[k] = b[first + (i + k) * stride];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ea += ta[k];
            eb += tb[k];
        }
    }
#pragma unroll 1
    for (; i < n; ++i) {
        ea += a[first + i * stride];
        eb += b[first + i * stride];
    }
}

// The tile's slot of substrate s in `partial` (edge sums at [0], l.y sums at
// [kGradChunk]).
__device__ __forceinline__ double* grad_slot(double* partial, int s, unsigned ntiles, unsigned tile)
{
    return partial + ((size_t)(s / kGradChunk) * ntiles + tile) * 2 * kGradChunk + s % kGradChunk;
}

template <typename TY, int L, int COLS>
__global__ void __launch_bounds__(kMaxThreads, 4)
k_sweep_adjoint_sums(SweepArgs A, const TY* __restrict__ y, double* __restrict__ partial)
{
    __shared__ double xa[kMaxThreads];
    __shared__ double xb[kMaxThreads];
    __shared__ double wfirst[COLS];
    __shared__ double wlast[COLS];
    __shared__ double nl[kMaxThreads];  // a segment's first node: l, then its edge sum
    __shared__ double ny[kMaxThreads];  // a segment's first node: y, then its l.y sum
    extern __shared__ double gtab[];
    const unsigned ntiles = (unsigned)(A.tiles * A.planes);
    const unsigned w = xcd_tile(blockIdx.x + blockIdx.y * (unsigned)A.tiles, ntiles, A.spread);
    const unsigned plane = w / (unsigned)A.tiles;
    const int tile = (int)(w - plane * (unsigned)A.tiles);
    const SegPos P = seg_pos<double, L>(A, COLS, tile, (int)plane);
    const double r = A.wb[P.s * 8 + 0], a = A.wb[P.s * 8 + 1];
    double raw[L];
    seg_load<double, L, false>(A, P, raw);
    // the slot's running sums, read now so the add at the end does not wait
    // (FP32 y: the registers are there; FP64 y would lose a block per CU)
    constexpr bool pre = sizeof(TY) == 4;
    double old_e = 0.0, old_ly = 0.0;
    if (pre && (int)threadIdx.x < A.S) {
        const double* sl = grad_slot(partial, threadIdx.x, ntiles, w);
        old_e = sl[0];
        old_ly = sl[kGradChunk];
    }
    TY yv[L];
    {
        // the same nodes of the forward output (P.vo counts bytes of doubles)
        const __amdgpu_buffer_rsrc_t ry = make_rsrc(y + P.plane0, (A.nvs - P.plane0) * sizeof(TY));
        const unsigned vo = P.vo / (unsigned)sizeof(double) * (unsigned)sizeof(TY);
        const unsigned step = A.line_stride * (unsigned)sizeof(TY);
        const int last = min(L, A.n - P.p0) - 1;
#pragma unroll
        for (int t = 0; t < L; ++t) yv[t] = buf_load((TY*)nullptr, ry, vo + (unsigned)min(t, last) * step, 0);
    }
    stage_tables(A, gtab);  // while the loads are in flight
    const int lane = threadIdx.x & (COLS - 1);
    const int cnt = min(L, A.n - P.p0);
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc_out<double>(A, P);
    const unsigned stepd = A.line_stride * (unsigned)sizeof(double);
    double s_ly = 0.0, s_edge = 0.0, lnext = 0.0, lfirst = 0.0, llast = 0.0;
    seg_core<double, L>(A, COLS, lane, P.s, P.p0, P.slot, P.last_writer, raw, r, a, xa, xb, wfirst, wlast, gtab,
                        NoSource{}, [&](int t, double zf) {
                            if (P.act && t < cnt) {
                                buf_store((double*)nullptr, rs, zf, P.vo, t * stepd);
                                s_ly += zf * (double)yv[t];
                                if (t + 1 < cnt) s_edge += (zf - lnext) * ((double)yv[t] - (double)yv[t + 1]);
                                lnext = zf;
                            }
                            if (t == L - 1) llast = zf;
                            if (t == 0) lfirst = zf;
                        });
    // the edge from this segment's last node to the next segment's first
    nl[P.slot] = lfirst;
    ny[P.slot] = (double)yv[0];
    lds_barrier();
    const int g = P.p0 / L;
    if (P.act && g + 1 < A.nseg)  // a segment with a next one holds L real nodes
        s_edge += (llast - nl[P.slot + COLS]) * ((double)yv[L - 1] - ny[P.slot + COLS]);
    lds_barrier();
    const int tid = threadIdx.x;
    nl[tid] = P.act ? s_edge : 0.0;
    ny[tid] = P.act ? s_ly : 0.0;
    lds_barrier();
    // per column, its segments in order; then per substrate, its columns in order
    if (tid < COLS) {
        double e = 0.0, ly = 0.0;
        ordered_sums(nl, ny, tid, A.nseg, COLS, e, ly);
        xa[tid] = e;
        xb[tid] = ly;
    }
    lds_barrier();
    for (int s = tid; s < A.S; s += blockDim.x) {
        double e = 0.0, ly = 0.0;
        // the tile's columns of substrate s: c = c0, c0 + S, ... (c0 the first)
        const int c0 = ((s - tile * COLS) % A.S + A.S) % A.S;
        ordered_sums(xa, xb, c0, (COLS - c0 + A.S - 1) / A.S, A.S, e, ly);
        double* slot = grad_slot(partial, s, ntiles, w);
        slot[0] = (pre && s == tid ? old_e : slot[0]) + e;
        slot[kGradChunk] = (pre && s == tid ? old_ly : slot[kGradChunk]) + ly;
    }
}

// The fused adjoint sweep and sums for x in the whole-cache-line form
// (k_sweep_xt's tile on the FP64 adjoint state, compile-time S): the
// forward output y is loaded per thread beside the chunk loads (its nodes at
// stride S), the sums are added as k_sweep_adjoint_sums adds them, the edge
// to the next segment of the row goes through the carry slots, and the
// block's sums are added up in a fixed order (segments of a column, then the
// rows of a substrate) onto the tile's slot of `partial`, laid out as
// k_sweep_adjoint_sums lays it out.
template <typename TY, int L>
struct XtAdjointSums {
    const TY* y;
    size_t nvs;
    double* partial;
    unsigned ntiles, tile;
    static constexpr bool kPre = sizeof(TY) == 4;
    TY yv[L];
    int cnt;
    double s_ly, s_edge, lnext, lfirst, llast, old_e, old_ly;

    __device__ __forceinline__ void begin(const XtPos& p)
    {
        const __amdgpu_buffer_rsrc_t ry = make_rsrc(y, nvs * sizeof(TY));
        cnt = min(L, p.n - p.p0);
        const unsigned base = p.row0 + (unsigned)p.p0 * p.S + p.s;
#pragma unroll
        for (int t = 0; t < L; ++t)
            yv[t] = buf_load((TY*)nullptr, ry, (base + (unsigned)(min(t, cnt - 1) * p.S)) * (unsigned)sizeof(TY), 0);
        s_ly = s_edge = lnext = lfirst = llast = 0.0;
        // the slot's running sums, read now so the add at the end does not wait
        // (FP32 y: the registers are there; FP64 y would lose a block per CU)
        old_e = old_ly = 0.0;
        if (kPre && (int)threadIdx.x < p.S) {
            const double* sl = grad_slot(partial, threadIdx.x, ntiles, tile);
            old_e = sl[0];
            old_ly = sl[kGradChunk];
        }
    }

    __device__ __forceinline__ void node(int t, double zf)
    {
        if (t < cnt) {
            s_ly += zf * (double)yv[t];
            if (t + 1 < cnt) s_edge += (zf - lnext) * ((double)yv[t] - (double)yv[t + 1]);
            lnext = zf;
        }
        if (t == L - 1) llast = zf;
        if (t == 0) lfirst = zf;
    }

    __device__ __forceinline__ void end(const XtPos& p, double* xa, double* xb, double* wfirst, double* wlast)
    {
        const bool own = p.q < p.Q;  // copies of segment Q-1 write nothing
        lds_barrier();               // every carry read is done
        if (own) {
            xa[p.slot] = lfirst;
            xb[p.slot] = (double)yv[0];
        }
        lds_barrier();
        // a segment with a next one in its row holds L real nodes
        if (p.act && p.g + 1 < p.nseg)
            s_edge += (llast - xa[p.slot + p.C]) * ((double)yv[L - 1] - xb[p.slot + p.C]);
        lds_barrier();
        if (own) {
            xa[p.slot] = p.act ? s_edge : 0.0;
            xb[p.slot] = p.act ? s_ly : 0.0;
        }
        lds_barrier();
        const int tid = threadIdx.x;
        if (tid < p.C) {  // per column, its segments in order
            double e = 0.0, ly = 0.0;
            ordered_sums(xa, xb, tid, p.nseg, p.C, e, ly);
            wfirst[tid] = e;
            wlast[tid] = ly;
        }
        lds_barrier();
        for (int s = tid; s < p.S; s += blockDim.x) {  // per substrate, its rows in order
            double e = 0.0, ly = 0.0;
            ordered_sums(wfirst, wlast, s, p.C / p.S, p.S, e, ly);
            double* slot = grad_slot(partial, s, ntiles, tile);
            slot[0] = (kPre && s == tid ? old_e : slot[0]) + e;
            slot[kGradChunk] = (kPre && s == tid ? old_ly : slot[kGradChunk]) + ly;
        }
    }
};

template <typename TY, int L, int SC>
__global__ void __launch_bounds__(kMaxThreads, 4)
k_sweep_adjoint_sums_xt(SweepArgs A, const TY* __restrict__ y, double* __restrict__ partial)
{
    extern __shared__ __attribute__((aligned(16))) double smem_xa[];
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x, A.spread);
    XtAdjointSums<TY, L> hook;
    hook.y = y;
    hook.nvs = A.nvs;
    hook.partial = partial;
    hook.ntiles = gridDim.x;
    hook.tile = tile;
    xt_tile<double, L, SC>(A, tile, smem_xa, hook);
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
// x: every (j, k) row's start is nx*S after the previous row's, across planes
// too, so the x lines of the whole store (or of a run of z planes) form one
// plane of ny*nz*S columns: a tile never straddles planes, and small planes
// fill whole blocks.
Geometry geometry(const Env& e, int axis)
{
    const size_t S = e.S, nx = e.nx, ny = e.ny;
    // (a store whose rows overflow an int sweeps x in runs of planes: sweep())
    if (axis == 0) return {e.nx, (int)std::min<size_t>(ny * e.nz * S, 0x7fffffff), 1, S, 0, nx * S};
    if (axis == 1) return {e.ny, (int)(nx * S), e.nz, nx * S, nx * ny * S, S};
    return {e.nz, (int)(nx * ny * S), 1, nx * ny * S, 0, S};
}

// Launch shape of one sweep: COLS lines per block and L nodes per segment,
// nseg = ceil(n/L) segments with COLS*nseg <= 1024, either one block per
// tile (k_sweep, L one of kColNodes) or resident blocks streaming through
// the tiles (k_sweep_stream, L one of kStreamNodes; it holds two tiles'
// raw values, so an FP64 store is limited to 16 nodes to keep within 128
// VGPRs), or for x row tiles through LDS (k_sweep_rows, L one of kNodes).
// cols == 0 means the line is too long for the segmented kernel.  A plan set
// by the autotuner (bfvm_autotune) wins over the default.
constexpr int kNodes[] = {4, 8, 12, 16, 24, 32};
constexpr int kColNodes[] = {4, 8, 12, 16, 20, 24, 32};
constexpr int kStreamNodes[] = {8, 12, 16, 24};
constexpr int kColsSet[] = {16, 32, 64};

bool stream_nodes_ok(int L, int precision)
{
    for (int v : kStreamNodes)
        if (v == L) return precision == BFVM_FP32 || L <= 16;
    return false;
}

bool plan_valid(int n, int cols, int L)
{
    return cols * ((n + L - 1) / L) <= kMaxThreads;
}

// The shape of a handle the autotuner has not seen: 32 lines of 16-node
// segments where the line fits (the tuned shapes on the grids measured have
// 16 or 32 nodes; 4-node segments ran 30-45 % slower at 128^3 x 2), longer
// segments for longer lines, shorter ones for short lines.
SweepPlan default_plan(int n)
{
    for (int cols : {32, 16})
        for (int L : {16, 32, 8, 24, 12, 4})
            if (plan_valid(n, cols, L) && (L <= 8 || L < 2 * n)) return SweepPlan{cols, L};
    for (int cols : {32, 16})
        for (int L : kNodes)
            if (plan_valid(n, cols, L)) return SweepPlan{cols, L};
    return SweepPlan{0, 0};
}

}  // namespace

// The segment length of an axis when the autotuner may not choose it: a
// fixed function of the line length, axis and store precision, so results
// do not depend on timing (the segment length changes the rounding of the
// solve; every other launch parameter leaves the bits unchanged).  Measured
// on 64^3 .. 400^3 grids with 2-8 substrates, it is the fastest length or
// within 4 % of it: for y and z lines of 400 or more in an FP32 store 20
// nodes while 32 lines of them fit a block (n <= 640: the block's 64 VGPRs
// let three blocks share a compute unit, against two of the 32-node
// segments' 94), 32 beyond; 16 for lines of 200 or more, for FP32 y and z
// lines of 32 or more, and for FP32 x lines of 32 or more whose rows are
// whole 4-node chunks (the whole-cache-line x kernel takes any S up to 64
// there): twice the bytes in flight per thread of 8 nodes (ensembles of
// 32-128-node grids: diffusion step 0.62 ms at 8-node x and 16-node y/z
// segments, 0.58 ms at 16 on every axis); 8 below.  Ensembles batch each
// axis at its members' own length.
int default_segment(const Env& e, int axis)
{
    const int n = axis == 0 ? e.nx : axis == 1 ? e.ny : e.nz;
    int L = 8;
    if (n >= 400 && axis > 0 && e.precision == BFVM_FP32) L = plan_valid(n, 32, 20) ? 20 : 32;
    else if (n >= 200) L = 16;
    else if (n >= 32 && e.precision == BFVM_FP32 && (axis > 0 || (n % 4 == 0 && e.S <= 64))) L = 16;
    for (int cols : {32, 16, 64})
        if (plan_valid(n, cols, L)) return L;
    return default_plan(n).L;  // lines too long for this length
}

namespace {

SweepPlan plan_for(const Env& e, int axis)
{
    const SweepPlan& t = e.plan[axis];
    if (t.cols > 0) return t;
    const int n = axis == 0 ? e.nx : axis == 1 ? e.ny : e.nz;
    const int L = default_segment(e, axis);
    if (axis == 2 && e.nvs * e.elem >= 0xffffffffull) {
        // a store above 4 GiB: z lines one segment per wave (SweepArgs::seg_base)
        if (plan_valid(n, 64, L)) return SweepPlan{64, L};
        for (int Lz : {16, 32, 8, 24, 12, 4})
            if (plan_valid(n, 64, Lz) && (Lz <= 8 || Lz < 2 * n)) return SweepPlan{64, Lz};
    }
    for (int cols : {32, 16, 64})
        if (plan_valid(n, cols, L)) return SweepPlan{cols, L};
    return default_plan(n);
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

// Kernel arguments of one axis of one grid, `cols` lines per block, planes
// [k0, k1) (k1 < 0: all planes).
SweepArgs sweep_args(const Env& e, int axis, int cols, int nseg, int k0 = 0, int k1 = -1)
{
    Geometry gm = geometry(e, axis);
    const AxisCoef& a = e.ax[axis];
    size_t first = 0;  // first value swept
    if (axis == 0) {
        // z planes [k0, k1) as one plane of their rows
        if (k1 < 0) k1 = e.nz;
        first = (size_t)k0 * e.nx * e.ny * e.S;
        gm.ncols = (k1 - k0) * e.ny * e.S;
        k0 = 0;
        k1 = 1;
    } else {
        if (k1 < 0) k1 = gm.nplanes;
        first = (size_t)k0 * gm.plane_stride;
    }
    SweepArgs A;
    A.rho = (char*)e.rho + first * e.elem;
    A.out = e.out ? (char*)e.out + first * e.elem : A.rho;
    A.nvs = e.nvs - first;
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
    A.planes = k1 - k0;
    A.spread = e.spread[axis];
    A.seg_base = axis == 2 && cols == 64 && e.nvs * e.elem >= 0xffffffffull ? 1 : 0;
    A.xt_pitch = 0;
    return A;
}

size_t sweep_lds(const SweepArgs& A)
{
    return sizeof(double) * (size_t)(A.kg0 + A.kg1 + 3 * kPad + 8) * A.S;
}

// Lets `kernel` take more than 64 KiB of dynamic LDS (the correction tables
// of a grid with many substrates): set once per kernel.
hipError_t allow_table_lds(const void* kernel, size_t lds, size_t cap = kMaxCorrectionLds)
{
    if (lds <= 64 * 1024) return hipSuccess;
    static std::mutex mu;
    static std::map<const void*, bool> done;
    std::lock_guard<std::mutex> lock(mu);
    if (done[kernel]) return hipSuccess;
    const hipError_t st = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap);
    if (st == hipSuccess) done[kernel] = true;
    return st;
}

template <typename T, int L, int COLS, int POL>
hipError_t launch_seg_pol(Env& e, int axis, int nseg, int k0, int k1)
{
    const SweepArgs A = sweep_args(e, axis, COLS, nseg, k0, k1);
    dim3 grid(A.tiles, A.planes);
    int threads = (COLS * nseg + 63) / 64 * 64;
    const hipError_t st = allow_table_lds((const void*)k_sweep<T, L, COLS, POL>, sweep_lds(A));
    if (st != hipSuccess) return st;
    k_sweep<T, L, COLS, POL><<<grid, threads, sweep_lds(A), e.stream>>>(A);
    return hipGetLastError();
}

// Non-temporal loads and stores (plan.nt, see kStreamingPolicy) are offered
// for y and z sweeps whose tiles' row pieces are whole 128-byte lines (with
// 64-byte pieces they moved 20-25 % more bytes than the sweep's own), and
// for the whole-cache-line x kernel (each lane moves 16-byte pieces of
// chunks that no other load reads again): at 400^3 x 8 FP32 with 1e6
// agents, in full steps, x 0.711 -> 0.692 ms, and the kernels after it no
// longer meet its lines in the caches: y 0.776 -> 0.759 ms, the agent
// kernel of the next step 0.084 -> 0.065 ms.
template <typename T, int COLS>
constexpr bool streams_whole_lines()
{
    return COLS * sizeof(T) >= 128;
}

bool nt_shape(const Env& e, int axis, int cols, int stream)
{
    if (axis == 0) return stream == 3;
    return stream != 2 && (size_t)cols * e.elem >= 128;
}

template <typename T, int L, int COLS>
hipError_t launch_seg(Env& e, int axis, int nseg, int k0, int k1, bool nt)
{
    if constexpr (streams_whole_lines<T, COLS>())
        if (nt) return launch_seg_pol<T, L, COLS, kStreamingPolicy>(e, axis, nseg, k0, k1);
    return launch_seg_pol<T, L, COLS, kLoadPolicy>(e, axis, nseg, k0, k1);
}

template <typename T, int COLS>
hipError_t launch_cols(Env& e, int axis, int L, int nseg, int k0, int k1, bool nt)
{
    switch (L) {
    case 4: return launch_seg<T, 4, COLS>(e, axis, nseg, k0, k1, nt);
    case 8: return launch_seg<T, 8, COLS>(e, axis, nseg, k0, k1, nt);
    case 12: return launch_seg<T, 12, COLS>(e, axis, nseg, k0, k1, nt);
    case 16: return launch_seg<T, 16, COLS>(e, axis, nseg, k0, k1, nt);
    case 20: return launch_seg<T, 20, COLS>(e, axis, nseg, k0, k1, nt);
    case 24: return launch_seg<T, 24, COLS>(e, axis, nseg, k0, k1, nt);
    default: return launch_seg<T, 32, COLS>(e, axis, nseg, k0, k1, nt);
    }
}

// Blocks of `fn` resident on one compute unit (cached per kernel and shape).
int resident_blocks(const void* fn, int threads, size_t lds)
{
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(fn, threads, lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds) != hipSuccess || nb < 1)
        nb = 1;
    cache[key] = nb;
    return nb;
}

int compute_units(Env& e)
{
    if (e.cus == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e.device) != hipSuccess ||
            cus < 1)
            cus = 1;
        e.cus = cus;
    }
    return e.cus;
}

template <typename T, int L, int COLS, int POL>
hipError_t launch_stream_pol(Env& e, int axis, int nseg, int k0, int k1)
{
    const SweepArgs A = sweep_args(e, axis, COLS, nseg, k0, k1);
    const int threads = (COLS * nseg + 63) / 64 * 64;
    const size_t lds = sweep_lds(A);
    const long long total = (long long)A.tiles * A.planes;
    const hipError_t st = allow_table_lds((const void*)k_sweep_stream<T, L, COLS, POL>, lds);
    if (st != hipSuccess) return st;
    const int nb = resident_blocks((const void*)k_sweep_stream<T, L, COLS, POL>, threads, lds);
    const int grid = (int)std::min<long long>(total, (long long)nb * compute_units(e));
    k_sweep_stream<T, L, COLS, POL><<<grid, threads, lds, e.stream>>>(A, grid);
    return hipGetLastError();
}

template <typename T, int L, int COLS>
hipError_t launch_stream(Env& e, int axis, int nseg, int k0, int k1, bool nt)
{
    if constexpr (streams_whole_lines<T, COLS>())
        if (nt) return launch_stream_pol<T, L, COLS, kStreamingPolicy>(e, axis, nseg, k0, k1);
    return launch_stream_pol<T, L, COLS, kLoadPolicy>(e, axis, nseg, k0, k1);
}

template <typename T, int COLS>
hipError_t launch_cols_stream(Env& e, int axis, int L, int nseg, int k0, int k1, bool nt)
{
    switch (L) {
    case 8: return launch_stream<T, 8, COLS>(e, axis, nseg, k0, k1, nt);
    case 12: return launch_stream<T, 12, COLS>(e, axis, nseg, k0, k1, nt);
    case 16: return launch_stream<T, 16, COLS>(e, axis, nseg, k0, k1, nt);
    default:
        if constexpr (sizeof(T) == 4) return launch_stream<T, 24, COLS>(e, axis, nseg, k0, k1, nt);
        return hipErrorInvalidValue;
    }
}

// Row tiles of the x sweep (k_sweep_rows): plan.cols = C = R*S lines, R rows
// per tile.
constexpr size_t kMaxRowTileLds = 150 * 1024;

struct RowShape {
    int R = 0, C = 0, nseg = 0, threads = 0;
    int P = 0;  // LDS row pitch in values (rows_tile)
    size_t lds = 0;
};

RowShape row_shape(const Env& e, int C, int L, int elem)
{
    RowShape r;
    if (C <= 0 || L <= 0 || C % e.S != 0 || e.nx < 2) return r;
    r.C = C;
    r.R = C / e.S;
    r.nseg = (e.nx + L - 1) / L;
    r.threads = (C * r.nseg + 63) / 64 * 64;
    // a row's values, then the last segment's nodes past n-1, then up to the
    // next pitch = S (mod 64)
    const int nxS = e.nx * e.S;
    r.P = nxS + (r.nseg * L - e.nx) * e.S;
    r.P += ((e.S - r.P) % 64 + 64) % 64;
    const size_t tables = (size_t)(e.ax[0].kg0 + e.ax[0].kg1 + 3 * kPad + 8) * e.S;
    r.lds = sizeof(double) * (2 * r.threads + 2 * C + tables) + (size_t)elem * r.R * r.P;
    return r;
}

bool rows_plan_valid(const Env& e, int C, int L)
{
    const RowShape r = row_shape(e, C, L, (int)e.elem);
    const size_t nxS = (size_t)e.nx * e.S, tile_vals = (size_t)r.R * nxS;
    // the slot arithmetic divides by umulhi: exact for numerator * divisor < 2^32
    return r.C > 0 && r.C * r.nseg <= kMaxThreads && r.lds <= kMaxRowTileLds && tile_vals * nxS < (1ull << 32);
}

unsigned div_magic(unsigned d)
{
    return (unsigned)((((unsigned long long)1 << 32) + d - 1) / d);
}

// Row-tile arguments of planes [k0, k1) of e's x sweep, C lines per tile.
RowTile row_tile_args(const Env& e, const RowShape& sh, int k0, int k1)
{
    RowTile X;
    X.A = sweep_args(e, 0, sh.C, sh.nseg, k0, k1);
    const long long rows = (long long)e.ny * (k1 - k0);
    X.R = sh.R;
    X.C = sh.C;
    X.nxS = e.nx * e.S;
    X.mag_nxS = div_magic((unsigned)X.nxS);
    X.tile_vals = X.R * X.nxS;
    X.vals = (size_t)rows * X.nxS;
    X.P = sh.P;
    return X;
}

template <typename T, int L>
hipError_t launch_rows(Env& e, const SweepPlan& p, int k0, int k1)
{
    const RowShape sh = row_shape(e, p.cols, L, sizeof(T));
    const RowTile X = row_tile_args(e, sh, k0, k1);
    const long long rows = (long long)e.ny * (k1 - k0);
    const long long grid = (rows + X.R - 1) / X.R;
    if (grid >= (1ll << 31)) return hipErrorInvalidValue;
    static bool lds_set = false;  // per instantiation
    if (!lds_set) {
        hipError_t st = hipFuncSetAttribute((const void*)k_sweep_rows<T, L>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxRowTileLds);
        if (st != hipSuccess) return st;
        lds_set = true;
    }
    k_sweep_rows<T, L><<<(unsigned)grid, sh.threads, sh.lds, e.stream>>>(X);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rows_L(Env& e, const SweepPlan& p, int k0, int k1)
{
    switch (p.L) {
    case 4: return launch_rows<T, 4>(e, p, k0, k1);
    case 8: return launch_rows<T, 8>(e, p, k0, k1);
    case 12: return launch_rows<T, 12>(e, p, k0, k1);
    case 16: return launch_rows<T, 16>(e, p, k0, k1);
    case 24: return launch_rows<T, 24>(e, p, k0, k1);
    default: return launch_rows<T, 32>(e, p, k0, k1);
    }
}

// Whole-cache-line x sweeps (k_sweep_xt): plan.cols = C = R*S lines, R rows
// per block, L one of kXtNodes.  S of 4, 8 or 16 (FP32 or FP64) take the
// compile-time form; any other S up to 64 the runtime form (FP32, rows of a
// multiple of 4 nodes).
constexpr int kXtNodes[] = {8, 16, 20};
constexpr int kMaxXtRuntimeS = 64;

bool xt_compile_time_s(int S) { return S == 4 || S == 8 || S == 16; }

template <typename T>
bool xt_supported(const Env& e, int L)
{
    bool nodes = false;
    for (int v : kXtNodes) nodes = nodes || v == L;
    if (!nodes || e.nx < 2) return false;
    switch (e.S) {
    case 4: return xt_shape_ok<T, 4>(L);
    case 8: return xt_shape_ok<T, 8>(L);
    case 16: return xt_shape_ok<T, 16>(L);
    default: return e.S <= kMaxXtRuntimeS && e.nx % 4 == 0 && xt_shape_ok<T, 0>(L);
    }
}

// Chunk-buffer pitch (values per segment) of the runtime-S form: the
// multiple of 4 values (16-byte pieces) from 4*S up that spreads the 64
// lanes reading one node of their segments over the most LDS banks.
int xt_pitch(int S)
{
    const int segw = 64 / S;
    int best = 4 * S, best_worst = 1 << 30;
    for (int P = 4 * S; P <= 4 * S + 64; P += 4) {
        int worst = 0;
        for (int u = 0; u < 4; ++u) {
            int cnt[64] = {0};
            for (int qq = 0; qq < segw; ++qq)
                for (int s = 0; s < S; ++s) worst = std::max(worst, ++cnt[(qq * P + u * S + s) % 64]);
        }
        if (worst < best_worst) {
            best_worst = worst;
            best = P;
        }
    }
    return best;
}

struct XtShape {
    int Q = 0, threads = 0, pitch = 0;
    size_t lds = 0;
};

template <typename T>
XtShape xt_shape(const Env& e, int C, int L)
{
    XtShape x;
    const int S = e.S;
    if (C <= 0 || L <= 0 || C % S != 0 || S > kMaxXtRuntimeS) return x;
    constexpr int G = 16 / (int)sizeof(T);
    const int segw = 64 / S;
    x.Q = (C / S) * ((e.nx + L - 1) / L);
    const int waves = (x.Q + segw - 1) / segw;
    x.threads = waves * 64;
    // compile-time forms: FP32 S = 4, 8, 16; FP64 also S = 2 (the adjoint's)
    x.pitch = xt_compile_time_s(S) || (G == 2 && S == 2) ? (G + 1) * S : xt_pitch(S);
    x.lds = sizeof(double) * (size_t)xt_buffer_offset(e.ax[0].kg0, e.ax[0].kg1, S, x.Q, C) +
            (size_t)waves * segw * x.pitch * sizeof(T);
    return x;
}

template <typename T>
bool xt_plan_valid(const Env& e, int C, int L)
{
    if (!xt_supported<T>(e, L)) return false;
    const XtShape x = xt_shape<T>(e, C, L);
    return x.Q > 0 && x.threads <= kMaxThreads && x.lds <= kMaxCorrectionLds;
}

template <typename T, int L, int SC>
hipError_t launch_xt(Env& e, const SweepPlan& p, int k0, int k1)
{
    if constexpr (!xt_shape_ok<T, SC>(L)) {
        return hipErrorInvalidValue;
    } else {
        const XtShape x = xt_shape<T>(e, p.cols, L);
        SweepArgs A = sweep_args(e, 0, p.cols, (e.nx + L - 1) / L, k0, k1);
        A.xt_pitch = x.pitch;
        const bool nt = p.nt != 0 && nt_shape(e, 0, p.cols, p.stream);
        const void* k = nt ? (const void*)k_sweep_xt<T, L, SC, kStreamingPolicy>
                           : (const void*)k_sweep_xt<T, L, SC, kLoadPolicy>;
        const hipError_t st = allow_table_lds(k, x.lds);
        if (st != hipSuccess) return st;
        if (nt)
            k_sweep_xt<T, L, SC, kStreamingPolicy><<<(unsigned)A.tiles, x.threads, x.lds, e.stream>>>(A);
        else
            k_sweep_xt<T, L, SC, kLoadPolicy><<<(unsigned)A.tiles, x.threads, x.lds, e.stream>>>(A);
        return hipGetLastError();
    }
}

template <typename T, int L>
hipError_t launch_xt_S(Env& e, const SweepPlan& p, int k0, int k1)
{
    switch (e.S) {
    case 4: return launch_xt<T, L, 4>(e, p, k0, k1);
    case 8: return launch_xt<T, L, 8>(e, p, k0, k1);
    case 16: return launch_xt<T, L, 16>(e, p, k0, k1);
    default: return launch_xt<T, L, 0>(e, p, k0, k1);
    }
}

template <typename T>
hipError_t launch_xt_L(Env& e, const SweepPlan& p, int k0, int k1)
{
    switch (p.L) {
    case 8: return launch_xt_S<T, 8>(e, p, k0, k1);
    case 16: return launch_xt_S<T, 16>(e, p, k0, k1);
    case 20: return launch_xt_S<T, 20>(e, p, k0, k1);
    default: return hipErrorInvalidValue;
    }
}

// Whether the segmented kernels run this axis (else the per-line fallback).
template <typename T>
bool segmented(const Env& e, int axis, const SweepPlan& p)
{
    if (p.stream == 2) return axis == 0 && rows_plan_valid(e, p.cols, p.L);
    if (p.stream == 3 && (axis != 0 || !xt_plan_valid<T>(e, p.cols, p.L))) return false;
    Geometry gm = geometry(e, axis);
    // buffer offsets are 32-bit: the segmented kernel addresses at most 4 GiB
    // from a plane's start (x: from the first of a run of z planes, at least
    // one; z with 64 lines per block: from a segment's first node)
    size_t span = (axis == 2 ? e.nvs : (size_t)e.nx * e.ny * e.S) * sizeof(T);
    if (axis == 2 && span >= 0xffffffffull && p.cols == 64 && p.stream != 2)
        span = (size_t)e.nx * e.ny * e.S * (size_t)(p.L + 1) * sizeof(T);
    // the end-correction tables live in LDS beside the 16 KiB carry arrays
    const size_t gbytes =
        sizeof(double) * (size_t)(e.ax[axis].kg0 + e.ax[axis].kg1 + 3 * kPad + 8) * e.S;
    return gm.n >= 2 && p.cols != 0 && plan_valid(gm.n, p.cols, p.L) && span < 0xffffffffull &&
           gbytes <= kMaxCorrectionLds;
}

// Sweep of planes [k0, k1) of `axis` (k1 < 0: all; the per-line fallback
// always sweeps all planes).
template <typename T>
hipError_t sweep(Env& e, int axis, int k0 = 0, int k1 = -1)
{
    Geometry gm = geometry(e, axis);
    SweepPlan p = plan_for(e, axis);
    if (!segmented<T>(e, axis, p)) return sweep_naive<T>(e, axis);
    if (k1 < 0) k1 = axis == 0 ? e.nz : gm.nplanes;
    if (p.stream == 2) return launch_rows_L<T>(e, p, k0, k1);
    int planes = k1 - k0;
    if (axis == 0) {
        // x buffer offsets are 32-bit from the first value swept: runs of
        // planes that span less than 4 GiB (segmented() checked one fits)
        const int run = (int)std::min<size_t>(e.nz, 0xffffffffull / ((size_t)e.nx * e.ny * e.S * sizeof(T)));
        if (k1 - k0 > run) {
            for (int k = k0; k < k1; k += run) {
                const hipError_t st = sweep<T>(e, 0, k, std::min(k1, k + run));
                if (st != hipSuccess) return st;
            }
            return hipSuccess;
        }
        gm.ncols = (k1 - k0) * e.ny * e.S;
        planes = 1;
        if (p.stream == 3) return launch_xt_L<T>(e, p, k0, k1);
    }
    int nseg = (gm.n + p.L - 1) / p.L;
    const long long tiles = ((long long)gm.ncols + p.cols - 1) / p.cols;
    const bool nt = p.nt != 0 && nt_shape(e, axis, p.cols, p.stream);
    if (p.stream && stream_nodes_ok(p.L, e.precision) && tiles * planes < (1ll << 31)) {
        switch (p.cols) {
        case 16: return launch_cols_stream<T, 16>(e, axis, p.L, nseg, k0, k1, nt);
        case 32: return launch_cols_stream<T, 32>(e, axis, p.L, nseg, k0, k1, nt);
        default: return launch_cols_stream<T, 64>(e, axis, p.L, nseg, k0, k1, nt);
        }
    }
    switch (p.cols) {
    case 16: return launch_cols<T, 16>(e, axis, p.L, nseg, k0, k1, nt);
    case 32: return launch_cols<T, 32>(e, axis, p.L, nseg, k0, k1, nt);
    default: return launch_cols<T, 64>(e, axis, p.L, nseg, k0, k1, nt);
    }
}

template <typename T, int L, int COLS>
hipError_t launch_x_src(Env& e, int nseg, const FusedSources& F)
{
    const SweepArgs A = sweep_args(e, 0, COLS, nseg);
    dim3 grid(A.tiles, A.planes);
    const int threads = (COLS * nseg + 63) / 64 * 64;
    const hipError_t st = allow_table_lds((const void*)k_sweep_x_src<T, L, COLS>, sweep_lds(A));
    if (st != hipSuccess) return st;
    k_sweep_x_src<T, L, COLS><<<grid, threads, sweep_lds(A), e.stream>>>(A, F);
    return hipGetLastError();
}

template <typename T, int COLS>
hipError_t launch_x_src_L(Env& e, int L, int nseg, const FusedSources& F)
{
    switch (L) {
    case 4: return launch_x_src<T, 4, COLS>(e, nseg, F);
    case 8: return launch_x_src<T, 8, COLS>(e, nseg, F);
    case 12: return launch_x_src<T, 12, COLS>(e, nseg, F);
    case 16: return launch_x_src<T, 16, COLS>(e, nseg, F);
    case 20: return launch_x_src<T, 20, COLS>(e, nseg, F);
    case 24: return launch_x_src<T, 24, COLS>(e, nseg, F);
    default: return launch_x_src<T, 32, COLS>(e, nseg, F);
    }
}

template <typename T>
hipError_t sweep_x_fused(Env& e, const SweepPlan& p, const FusedSources& F)
{
    const int nseg = (e.nx + p.L - 1) / p.L;
    switch (p.cols) {
    case 16: return launch_x_src_L<T, 16>(e, p.L, nseg, F);
    case 32: return launch_x_src_L<T, 32>(e, p.L, nseg, F);
    default: return launch_x_src_L<T, 64>(e, p.L, nseg, F);
    }
}

}  // namespace

// The x shape of the fused sweep: the autotuner's fastest one-block-per-tile
// shape, else the tuned shape if it is one, else the default.
SweepPlan fused_x_plan(const Env& e)
{
    if (e.fused_plan.cols > 0) return e.fused_plan;
    const SweepPlan p = plan_for(e, 0);
    return p.stream == 0 ? p : default_plan(e.nx);
}

bool sweep_x_fusable(const Env& e)
{
    const SweepPlan p = fused_x_plan(e);
    const bool seg = e.precision == BFVM_FP64 ? segmented<double>(e, 0, p) : segmented<float>(e, 0, p);
    // one launch over all planes: 32-bit offsets from the store's start and
    // into the FP64 bulk fields
    return seg && p.stream == 0 && e.nvox < (1ull << 31) && e.nvs * e.elem < 0xffffffffull &&
           e.nvs * sizeof(double) < 0xffffffffull;
}

hipError_t launch_sweep_x_fused(Env& e, bool bulk, bool skip_agent_voxels, bool dir_after)
{
    FusedSources F{};
    if (bulk) {
        F.bulk_a = (const double*)e.bulk_A;
        F.bulk_b = (const double*)e.bulk_B;
        F.bulk_bytes = e.nvs * sizeof(double);
        if (skip_agent_voxels && e.n_agent_voxels > 0) {
            if (!e.agent_bits) return hipErrorInvalidValue;
            F.agent_bits = e.agent_bits;
        }
    }
    if (e.n_dir > 0) {
        F.dbits = e.dir_bits;
        F.drank = e.dir_rank;
        F.dval = e.dir_val;
        F.dmask = e.dir_mask;
    }
    F.ny = e.ny;
    F.dir_after = dir_after ? 1 : 0;
    const SweepPlan p = fused_x_plan(e);
    return e.precision == BFVM_FP64 ? sweep_x_fused<double>(e, p, F) : sweep_x_fused<float>(e, p, F);
}

bool sweep_nt_allowed(const Env& e, int axis, const SweepPlan& p)
{
    return nt_shape(e, axis, p.cols, p.stream);
}

// The shape that runs the axis; lines = 0 when the one-thread-per-line
// fallback runs it.
SweepPlan effective_plan(const Env& e, int axis)
{
    const SweepPlan p = plan_for(e, axis);
    const bool seg = e.precision == BFVM_FP64 ? segmented<double>(e, axis, p) : segmented<float>(e, axis, p);
    return seg ? p : SweepPlan{0, 0, 0};
}

std::vector<SweepPlan> sweep_candidates(const Env& e, int axis, bool any_length)
{
    const int n = axis == 0 ? e.nx : axis == 1 ? e.ny : e.nz;
    const int Lfix = any_length ? 0 : default_segment(e, axis);  // 0: any length
    std::vector<SweepPlan> out;
    for (int cols : kColsSet)
        for (int L : kColNodes)
            if (plan_valid(n, cols, L) && (L <= 8 || L < 2 * n) && (Lfix == 0 || L == Lfix)) {
                out.push_back(SweepPlan{cols, L, 0});
                if (stream_nodes_ok(L, e.precision)) out.push_back(SweepPlan{cols, L, 1});
                if (nt_shape(e, axis, cols, 0)) {
                    out.push_back(SweepPlan{cols, L, 0, 1});
                    if (stream_nodes_ok(L, e.precision)) out.push_back(SweepPlan{cols, L, 1, 1});
                }
            }
    if (axis == 0)
        for (int R = 1; R <= 64; R *= 2)
            for (int L : kNodes)
                if ((L <= 8 || L < 2 * n) && (Lfix == 0 || L == Lfix) && rows_plan_valid(e, R * e.S, L))
                    out.push_back(SweepPlan{R * e.S, L, 2});
    if (axis == 0)
        for (int R = 1; R <= 128; R = R < 8 ? R + 1 : 2 * R)  // rows per block
            for (int L : kXtNodes) {
                const bool ok = e.precision == BFVM_FP64 ? xt_plan_valid<double>(e, R * e.S, L)
                                                         : xt_plan_valid<float>(e, R * e.S, L);
                if ((L <= 8 || L < 2 * n) && (Lfix == 0 || L == Lfix) && ok) {
                    out.push_back(SweepPlan{R * e.S, L, 3});
                    out.push_back(SweepPlan{R * e.S, L, 3, 1});
                }
            }
    return out;
}

// Ensemble launch of one axis (see k_sweep_batch).  A member can join when
// its lines fit 16 segments of L nodes and its tables fit the kernel.
namespace {
template <typename T>
bool batch_member_ok_t(const Env& e, int axis, int L)
{
    Geometry gm = geometry(e, axis);
    // a member's x lines are one plane of all its rows
    const size_t span = (axis == 1 ? (size_t)e.nx * e.ny * e.S : e.nvs) * sizeof(T);
    const size_t gbytes =
        sizeof(double) * (size_t)(e.ax[axis].kg0 + e.ax[axis].kg1 + 3 * kPad + 8) * e.S;
    const int nseg = (gm.n + L - 1) / L;
    return gm.n >= 2 && nseg <= kMaxThreads / 64 && span < 0xffffffffull && gbytes <= kMaxBatchCorrectionLds;
}

template <typename T, int L, int POL>
hipError_t launch_batch_pol(const SweepArgs* d_args, const int* d_first, int nm, int blocks, size_t lds,
                            hipStream_t st, int spread)
{
    const hipError_t a = allow_table_lds((const void*)k_sweep_batch<T, L, POL>, lds, kMaxBatchCorrectionLds);
    if (a != hipSuccess) return a;
    k_sweep_batch<T, L, POL><<<blocks, kMaxThreads, lds, st>>>(d_args, d_first, nm, spread);
    return hipGetLastError();
}

// nt: non-temporal loads and stores (the batch's tiles are 64 lines or
// more, whole 128-byte row pieces in either precision)
template <typename T, int L>
hipError_t launch_batch_t(const SweepArgs* d_args, const int* d_first, int nm, int blocks,
                          size_t lds, hipStream_t st, int spread, bool nt)
{
    return nt ? launch_batch_pol<T, L, kStreamingPolicy>(d_args, d_first, nm, blocks, lds, st, spread)
              : launch_batch_pol<T, L, kLoadPolicy>(d_args, d_first, nm, blocks, lds, st, spread);
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

hipError_t launch_sweep_batch(int precision, int L, int spread, const SweepArgs* d_args, const int* d_first,
                              int nm, int blocks, size_t lds, hipStream_t st, bool nt)
{
    if (precision == BFVM_FP64) {
        switch (L) {
        case 4: return launch_batch_t<double, 4>(d_args, d_first, nm, blocks, lds, st, spread, nt);
        case 8: return launch_batch_t<double, 8>(d_args, d_first, nm, blocks, lds, st, spread, nt);
        case 16: return launch_batch_t<double, 16>(d_args, d_first, nm, blocks, lds, st, spread, nt);
        default: return launch_batch_t<double, 32>(d_args, d_first, nm, blocks, lds, st, spread, nt);
        }
    }
    switch (L) {
    case 4: return launch_batch_t<float, 4>(d_args, d_first, nm, blocks, lds, st, spread, nt);
    case 8: return launch_batch_t<float, 8>(d_args, d_first, nm, blocks, lds, st, spread, nt);
    case 16: return launch_batch_t<float, 16>(d_args, d_first, nm, blocks, lds, st, spread, nt);
    default: return launch_batch_t<float, 32>(d_args, d_first, nm, blocks, lds, st, spread, nt);
    }
}

// Row tiles of an ensemble's x sweeps (k_sweep_rows_batch): a member joins
// with C = the most lines (a multiple of S) whose tile fills at most 1024
// threads for segments of L nodes and fits the LDS limit, so that the blocks
// of short-lined members are as full as those of long-lined ones.
bool batch_rows_args(const Env& e, int L, RowTile* X, int* threads, size_t* lds, int* tiles)
{
    if (e.nx < 2 || e.nvox >= (1ull << 31)) return false;
    const int nseg = (e.nx + L - 1) / L;
    for (int C = kMaxThreads / nseg / e.S * e.S; C >= e.S; C -= e.S) {
        if (!rows_plan_valid(e, C, L)) continue;
        const RowShape sh = row_shape(e, C, L, (int)e.elem);
        // the batch's blocks may have up to kMaxThreads threads' carry arrays
        if (sh.lds - sizeof(double) * 2 * sh.threads + sizeof(double) * 2 * kMaxThreads > kMaxRowTileLds) continue;
        *X = row_tile_args(e, sh, 0, e.nz);
        *threads = sh.threads;
        *lds = sh.lds;
        *tiles = (int)(((long long)e.ny * e.nz + sh.R - 1) / sh.R);
        return true;
    }
    return false;
}

namespace {
template <typename T, int L>
hipError_t launch_rows_batch_t(const RowTile* d_tabs, const int* d_map, int blocks, int threads, size_t lds,
                               int spread, hipStream_t st)
{
    static bool lds_set = false;  // per instantiation
    if (!lds_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sweep_rows_batch<T, L>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxRowTileLds);
        if (e != hipSuccess) return e;
        lds_set = true;
    }
    k_sweep_rows_batch<T, L><<<blocks, threads, lds, st>>>(d_tabs, d_map, spread);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_rows_batch(int precision, int L, int spread, const RowTile* d_tabs, const int* d_map,
                             int blocks, int threads, size_t lds, hipStream_t st)
{
    if (blocks == 0) return hipSuccess;
    if (precision == BFVM_FP64) {
        switch (L) {
        case 4: return launch_rows_batch_t<double, 4>(d_tabs, d_map, blocks, threads, lds, spread, st);
        case 8: return launch_rows_batch_t<double, 8>(d_tabs, d_map, blocks, threads, lds, spread, st);
        case 16: return launch_rows_batch_t<double, 16>(d_tabs, d_map, blocks, threads, lds, spread, st);
        default: return launch_rows_batch_t<double, 32>(d_tabs, d_map, blocks, threads, lds, spread, st);
        }
    }
    switch (L) {
    case 4: return launch_rows_batch_t<float, 4>(d_tabs, d_map, blocks, threads, lds, spread, st);
    case 8: return launch_rows_batch_t<float, 8>(d_tabs, d_map, blocks, threads, lds, spread, st);
    case 16: return launch_rows_batch_t<float, 16>(d_tabs, d_map, blocks, threads, lds, spread, st);
    default: return launch_rows_batch_t<float, 32>(d_tabs, d_map, blocks, threads, lds, spread, st);
    }
}

// Whole-cache-line x sweep of an ensemble member (k_sweep_xt_batch, the
// runtime-S form: FP32, rows of a multiple of 4 nodes): the member's
// arguments for blocks of `threads` threads, R = the most rows whose
// segments fill them; false when the member cannot take part.
bool batch_xt_args(const Env& e, int L, int threads, SweepArgs* A, size_t* lds, int* tiles)
{
    if (e.precision != BFVM_FP32 || e.S > kMaxXtRuntimeS || e.nx % 4 != 0 || !xt_shape_ok<float, 0>(L) ||
        !xt_supported<float>(e, L) || !batch_member_ok(e, 0, L))
        return false;
    const int S = e.S, segw = 64 / S, nseg = (e.nx + L - 1) / L;
    const int rows = e.ny * e.nz;
    const int R = std::min(rows, (threads / 64) * segw / nseg);
    if (R < 1) return false;
    XtShape x = xt_shape<float>(e, R * S, L);
    x.pitch = xt_pitch(S);  // the batch kernel is the runtime form for every S
    x.lds = sizeof(double) * (size_t)xt_buffer_offset(e.ax[0].kg0, e.ax[0].kg1, S, x.Q, R * S) +
            (size_t)(threads / 64) * segw * x.pitch * sizeof(float);
    if (x.threads > threads || x.lds > kMaxBatchCorrectionLds) return false;
    *A = sweep_args(e, 0, R * S, nseg);
    A->xt_pitch = x.pitch;
    *lds = x.lds;
    *tiles = A->tiles * A->planes;
    return true;
}

hipError_t launch_xt_batch(int L, int spread, const SweepArgs* d_args, const int* d_map, int blocks, int threads,
                           size_t lds, hipStream_t st)
{
    if (blocks == 0) return hipSuccess;
    const void* fn = L == 8 ? (const void*)k_sweep_xt_batch<float, 8>
                   : L == 16 ? (const void*)k_sweep_xt_batch<float, 16>
                             : (const void*)k_sweep_xt_batch<float, 20>;
    if (L != 8 && L != 16 && L != 20) return hipErrorInvalidValue;
    const hipError_t a = allow_table_lds(fn, lds, kMaxBatchCorrectionLds);
    if (a != hipSuccess) return a;
    switch (L) {
    case 8: k_sweep_xt_batch<float, 8><<<blocks, threads, lds, st>>>(d_args, d_map, spread); break;
    case 16: k_sweep_xt_batch<float, 16><<<blocks, threads, lds, st>>>(d_args, d_map, spread); break;
    default: k_sweep_xt_batch<float, 20><<<blocks, threads, lds, st>>>(d_args, d_map, spread); break;
    }
    return hipGetLastError();
}

hipError_t launch_sweep(Env& e, int axis)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis) : sweep<float>(e, axis);
}

hipError_t launch_sweep_to(Env& e, int axis, const void* src, void* dst)
{
    const SweepPlan p = plan_for(e, axis);
    const bool seg = e.precision == BFVM_FP64 ? segmented<double>(e, axis, p) : segmented<float>(e, axis, p);
    void* keep = e.rho;
    hipError_t st;
    if (seg) {
        // the segmented kernels load a line before storing any of it
        e.rho = const_cast<void*>(src);
        e.out = dst;
        st = launch_sweep(e, axis);
    } else {
        // the one-thread-per-line fallback sweeps in place
        st = hipMemcpyAsync(dst, src, e.nvs * e.elem, hipMemcpyDeviceToDevice, e.stream);
        e.rho = dst;
        if (st == hipSuccess) st = launch_sweep(e, axis);
    }
    e.rho = keep;
    e.out = nullptr;
    return st;
}

// The fused adjoint sweep and D/lambda sums of `axis` (k_sweep_adjoint_sums),
// 8-node segments, the widest column tiles of up to 1024 threads: its shape
// (lines per block, 0 when it cannot run: a line of one node, or 32-bit
// buffer offsets short of the FP64 state) and its tiles.
AdjointSumsShape adjoint_sums_shape(const Env& e, int axis, int y_precision)
{
    AdjointSumsShape sh;
    const Geometry gm = geometry(e, axis);
    constexpr int L = kAdjointSumsNodes;
    const int nseg = (gm.n + L - 1) / L;
    if (gm.n < 2 || e.nvs * sizeof(double) >= 0xffffffffull || e.nvox >= (1ull << 31)) return sh;
    const size_t gbytes = sizeof(double) * (size_t)(e.ax[axis].kg0 + e.ax[axis].kg1 + 3 * kPad + 8) * e.S;
    if (gbytes > kMaxCorrectionLds) return sh;
    // x lines of 2-16 substrates (a node's FP64 values whole 16-byte pieces):
    // the whole-cache-line form, as many rows per block as 1024 threads hold
    // (512 with an FP64 y: its 66 VGPRs let 7 waves per SIMD run, one
    // 1024-thread block per compute unit)
    const int S = e.S;
    if (axis == 0 && (S == 2 || S == 4 || S == 8 || S == 16) && e.nx >= 2) {
        const int rows = gm.ncols / S;
        const int most = y_precision == BFVM_FP32 ? kMaxThreads : kMaxThreads / 2;
        const int R = std::min(rows, std::max(1, most / (nseg * S)));
        const XtShape x = xt_shape<double>(e, R * S, L);
        if (x.Q > 0 && x.threads <= kMaxThreads && x.lds <= kMaxCorrectionLds) {
            sh.cols = R * S;
            sh.xt_rows = R;
            sh.tiles = (size_t)((rows + R - 1) / R);
            return sh;
        }
    }
    for (int cols : {64, 32, 16})
        if (cols * nseg <= kMaxThreads) {
            sh.cols = cols;
            break;
        }
    if (sh.cols == 0) return sh;
    const SweepArgs A = sweep_args(e, axis, sh.cols, nseg);
    sh.tiles = (size_t)A.tiles * A.planes;
    return sh;
}

template <typename TY, int SC>
hipError_t launch_adjoint_sums_xt(Env& e, int cols, double* lam, const void* y, double* partial)
{
    constexpr int L = kAdjointSumsNodes;
    const XtShape x = xt_shape<double>(e, cols, L);
    void* keep = e.rho;
    e.rho = lam;
    const SweepArgs A = sweep_args(e, 0, cols, (e.nx + L - 1) / L);
    e.rho = keep;
    const hipError_t st = allow_table_lds((const void*)k_sweep_adjoint_sums_xt<TY, L, SC>, x.lds);
    if (st != hipSuccess) return st;
    k_sweep_adjoint_sums_xt<TY, L, SC><<<(unsigned)A.tiles, x.threads, x.lds, e.stream>>>(A, (const TY*)y, partial);
    return hipGetLastError();
}

template <typename TY>
hipError_t launch_adjoint_sums_xt_S(Env& e, int cols, double* lam, const void* y, double* partial)
{
    switch (e.S) {
    case 2: return launch_adjoint_sums_xt<TY, 2>(e, cols, lam, y, partial);
    case 4: return launch_adjoint_sums_xt<TY, 4>(e, cols, lam, y, partial);
    case 8: return launch_adjoint_sums_xt<TY, 8>(e, cols, lam, y, partial);
    case 16: return launch_adjoint_sums_xt<TY, 16>(e, cols, lam, y, partial);
    default: return hipErrorInvalidValue;
    }
}

template <typename TY, int COLS>
hipError_t launch_adjoint_sums_t(Env& e, int axis, double* lam, const void* y, double* partial)
{
    constexpr int L = kAdjointSumsNodes;
    const Geometry gm = geometry(e, axis);
    void* keep = e.rho;
    e.rho = lam;
    SweepArgs A = sweep_args(e, axis, COLS, (gm.n + L - 1) / L);
    e.rho = keep;
    const hipError_t st = allow_table_lds((const void*)k_sweep_adjoint_sums<TY, L, COLS>, sweep_lds(A));
    if (st != hipSuccess) return st;
    dim3 grid(A.tiles, A.planes);
    const int threads = (COLS * A.nseg + 63) / 64 * 64;
    k_sweep_adjoint_sums<TY, L, COLS><<<grid, threads, sweep_lds(A), e.stream>>>(A, (const TY*)y, partial);
    return hipGetLastError();
}

// Runs with the handle set up for the FP64 adjoint state (e.precision and
// e.elem FP64 while the arguments are built); y in the store's precision.
hipError_t launch_adjoint_sums(Env& e, int axis, double* lam, const void* y, int y_precision, double* partial)
{
    const AdjointSumsShape sh = adjoint_sums_shape(e, axis, y_precision);
    const bool f = y_precision == BFVM_FP32;
    if (sh.xt_rows)
        return f ? launch_adjoint_sums_xt_S<float>(e, sh.cols, lam, y, partial)
                 : launch_adjoint_sums_xt_S<double>(e, sh.cols, lam, y, partial);
    switch (sh.cols) {
    case 64: return f ? launch_adjoint_sums_t<float, 64>(e, axis, lam, y, partial)
                      : launch_adjoint_sums_t<double, 64>(e, axis, lam, y, partial);
    case 32: return f ? launch_adjoint_sums_t<float, 32>(e, axis, lam, y, partial)
                      : launch_adjoint_sums_t<double, 32>(e, axis, lam, y, partial);
    case 16: return f ? launch_adjoint_sums_t<float, 16>(e, axis, lam, y, partial)
                      : launch_adjoint_sums_t<double, 16>(e, axis, lam, y, partial);
    default: return hipErrorInvalidValue;
    }
}

bool sweep_chunkable(const Env& e, int axis)
{
    const SweepPlan p = plan_for(e, axis);
    return e.precision == BFVM_FP64 ? segmented<double>(e, axis, p) : segmented<float>(e, axis, p);
}

hipError_t launch_sweep_planes(Env& e, int axis, int k0, int k1)
{
    return e.precision == BFVM_FP64 ? sweep<double>(e, axis, k0, k1) : sweep<float>(e, axis, k0, k1);
}

}  // namespace bfvm
