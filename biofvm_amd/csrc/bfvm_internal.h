// Internal state of a biofvm_amd handle and the kernel launchers shared by
// the C-ABI (capi.hip) and the kernel files (sources.hip, lod.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "biofvm_amd.h"

namespace bfvm {

// Runs the enclosing scope with a private state of the C library's random()
// generator (which rand() shares), restoring the caller's state on exit.  The
// HIP runtime seeds that generator from the clock when it first loads a code
// object (at the first launch of a kernel of each source file), so without
// it a program's rand() sequence would depend on when this library first
// ran a kernel.  Every entry point of the C ABI holds one.  (Other threads
// that call rand() while an entry point runs draw from the private state.)
struct RandGuard {
    char state[8];
    char* saved;
    RandGuard() : saved(initstate(1u, state, sizeof state)) {}
    ~RandGuard() { setstate(saved); }
    RandGuard(const RandGuard&) = delete;
    RandGuard& operator=(const RandGuard&) = delete;
};

// Message of a failed HIP call.  A failed runtime call (an allocation the
// device refuses, say) also stays the thread's last error, which the next
// launch check (hipGetLastError) would report against an unrelated kernel;
// taking the message consumes it.
inline std::string hip_error_text(hipError_t st)
{
    (void)hipGetLastError();
    return hipGetErrorString(st);
}

// Tridiagonal solve of one axis, all substrates, FP64 on the device for
// either store precision.  Element (i, s) of an (n, S) table is at [i*S + s].
//
// The reference factorises A = tridiag(-c1, 1+2c1+c2, -c1) with end
// diagonals 1+c1+c2 (BioFVM_solvers.cpp:104-142); its Thomas coefficients
// denom_i converge to a fixed point d away from the ends.  The segmented
// sweep inverts instead the constant-coefficient product A_inf = L U with
// L = bidiag(-c1, d), U = bidiag(1, -c1/d), which differs from A only in
// the two corner diagonals, and restores A^-1 exactly with a rank-2
// (Woodbury) correction:
//   w = A_inf^-1 x,   z = A^-1 x = w + mu0*g0 + mu1*g1,
//   (mu0, mu1) = K (w_0, w_{n-1}),  g0 = A_inf^-1 (delta0 e_0),
//   g1 = A_inf^-1 (delta1 e_{n-1}),  K = (I - [[g0_0, g1_0], [g0_{n-1}, g1_{n-1}]])^-1
// with delta0 = d - (1+c1+c2), delta1 = d + c1^2/d - (1+c1+c2).  g0 and g1
// decay geometrically from their end, so only their first kg0 / last kg1
// rows (above 1e-18 of their peak) are stored.
struct AxisCoef {
    int n = 0;
    void* denom = nullptr;  // denom_i (BioFVM_solvers.cpp:130-141), for the per-line kernel
    void* c = nullptr;      // c_i = -c1/denom_i (:137-141)
    void* wb = nullptr;     // per substrate, 8 values: 1/d, c1/d, K00, K01, K10, K11, 0, 0
    void* g0 = nullptr;     // (kg0, S): g0 rows 0..kg0-1
    void* g1 = nullptr;     // (kg1, S): g1 rows n-1, n-2, ..., n-kg1
    int kg0 = 0, kg1 = 0;
};

// Arguments of the segmented sweep kernel for one axis of one grid (see
// lod.hip); also the per-member table of the ensemble launch.
struct SweepArgs {
    void* rho;               // density store (float or double)
    void* out;               // where the swept values go: rho (in place) or a store of the same shape
    size_t nvs;              // values in the store
    size_t plane_stride;     // values between planes
    const double* wb;        // per substrate: 1/d, c1/d, K00, K01, K10, K11, 0, 0
    const double* g0;        // (kg0, S)
    const double* g1;        // (kg1, S)
    int n, nseg;             // nodes per line, segments per line
    int cols_per_plane, S, kg0, kg1;
    unsigned line_stride;    // values between nodes of a line
    unsigned vstride;        // values between voxels of a plane's column set
    int cols;                // lines per block (a power of two)
    int tiles, planes;       // column tiles per plane, planes
    int spread;              // tiles per XCD chunk (see xcd_tile in lod.hip)
    // 1: every wave's buffer starts at its own segment's first node (cols =
    // 64, one segment per wave), so a z line may span more than 4 GiB of
    // addresses: the store's z sweep above 4 GiB
    int seg_base;
    // values per segment in the chunk buffers of the whole-cache-line x
    // sweep's runtime-S form (k_sweep_xt in lod.hip); 0 for other kernels
    int xt_pitch;
};

// Arguments of the x sweep on LDS-staged row tiles (k_sweep_rows in lod.hip):
// a tile is R consecutive x rows, C = R*S lines.
struct RowTile {
    SweepArgs A;             // axis-0 arguments of the planes swept (rho at the first row)
    size_t vals;             // values swept: rows * nx * S
    int R, C;                // rows per tile, lines per tile (R*S)
    int nxS;                 // values per row
    unsigned mag_nxS;        // ceil(2^32/nxS), nxS >= 2: e/nxS = umulhi(e, mag) for e*nxS < 2^32
    int tile_vals;           // R*nx*S
    int P;                   // LDS row pitch in values: nx*S, the last segment's nodes past n-1, = S mod 64
};

// Launch shape of the segmented sweep of one axis (see lod.hip).
struct SweepPlan {
    int cols = 0;    // lines per block (16, 32 or 64; row tiles: a multiple of S); 0 = default
    int L = 0;       // nodes per segment
    int stream = 0;  // kernel: 0 one block per tile, 1 resident blocks that load the
                     // next tile while solving one, 2 x sweep on LDS-staged row tiles,
                     // 3 x sweep with whole-cache-line loads and stores (k_sweep_xt)
    int nt = 0;      // 1: non-temporal loads and stores (y and z sweeps whose tiles'
                     // row pieces are whole 128-byte cache lines, kernel 3 x sweeps)
};

constexpr int kDefaultSpread = -1;

struct Env {
    int nx = 0, ny = 0, nz = 0, S = 0;
    double dx = 0, dy = 0, dz = 0;
    bool uniform = true;
    int precision = BFVM_FP64;
    size_t nvox = 0, nvs = 0;  // voxels, voxel-substrate values
    size_t elem = 8;           // bytes per device value
    hipStream_t stream = nullptr;
    hipStream_t ensemble_stream = nullptr;  // stream of the ensemble holding this handle, if any
    int device = 0;
    int cus = 0;  // compute units of the device

    void* rho = nullptr;  // density store, nvs values
    void* out = nullptr;  // set only around an out-of-place sweep (launch_sweep_to)

    // Coefficients: D, lambda as set by the caller; the factorisation is
    // rebuilt when they or dt change.
    std::vector<double> D, lambda;
    bool coef_valid = false;
    double coef_dt = 0;
    unsigned coef_epoch = 0;  // bumped whenever the device tables are rebuilt
    void* c1 = nullptr;  // dt*D/dx^2 per substrate (FP64)
    AxisCoef ax[3];
    SweepPlan plan[3];  // per-axis launch shape chosen by the autotuner
    SweepPlan fused_plan;  // the fastest one-block-per-tile x shape (for the fused x sweep)
    // per axis: how tiles are spread over the XCDs (xcd_tile in lod.hip):
    // 0 as dispatched, k > 0 chunks of k consecutive tiles per XCD, -1 one
    // contiguous range per XCD.  Chosen by the autotuner.
    int spread[3] = {kDefaultSpread, kDefaultSpread, kDefaultSpread};
    // x and y sweeps run in chunks of this many z planes (x then y per chunk),
    // so the y sweep reads the x sweep's output from the last-level cache;
    // 0 = whole-store sweeps.  Chosen by the autotuner.
    int xy_chunk = 0;
    // 3: diffusion_decay_solver__constant_coefficients_LOD_3D (x, y, z sweeps,
    // decay split in thirds); 2: ..._LOD_2D (BioFVM_solvers.cpp:278-427: x and
    // y sweeps, decay split in halves; nz must be 1).
    int lod_dims = 3;
    // bulk, agents and the Dirichlet overwrites before the x and y sweeps run
    // inside the x sweep (load prologue and store epilogue) when it can
    bool fused_sources = true;
    // whether bfvm_autotune may choose the segment length (which changes the
    // solve's rounding); otherwise it keeps default_segment's
    bool tune_segments = false;
    // the adjoint's backward pass keeps each step's states when they fit
    // (bfvm_set_tape_keep_states)
    bool tape_keep_states = true;
    // bumped whenever a launch shape, cache policy or XCD spread is set, so
    // that an ensemble holding this handle rebuilds its tables and graph
    unsigned plan_epoch = 0;
    // number of times the Thomas/Woodbury tables were built
    unsigned coef_builds = 0;

    // Dirichlet nodes, one per voxel (the last of repeated entries, as the
    // reference's in-order overwrite leaves it), sorted by voxel index.
    int n_dir = 0;
    int* dir_idx = nullptr;
    void* dir_val = nullptr;          // n_dir*S, device precision
    unsigned char* dir_mask = nullptr;  // S bytes
    // The same nodes as a bitmap over voxels (bit v%32 of word v/32, two
    // spare zero words at the end) and, per word, the number of nodes in the
    // words before it: node of voxel v = dir_rank[v/32] + popcount of the
    // lower bits of its word.  Read by the fused x sweep.
    unsigned* dir_bits = nullptr;
    int* dir_rank = nullptr;
    unsigned source_epoch = 0;  // bumped whenever Dirichlet nodes or bulk fields are set

    // Bulk fields and the implicit-step terms built from them for one dt:
    // A = dt*(T*S), B = (U+S)*dt + 1 (all FP64 on the device).
    bool have_bulk = false;
    double* bulk_supply = nullptr;
    double* bulk_target = nullptr;
    double* bulk_uptake = nullptr;
    void* bulk_A = nullptr;
    void* bulk_B = nullptr;
    double bulk_dt = 0;
    bool bulk_valid = false;

    // Agents grouped by voxel (list order kept within a voxel).
    int n_agents = 0;
    int n_agent_voxels = 0;
    unsigned agent_epoch = 0;  // bumped whenever the agent arrays change
    int* agent_seg_voxel = nullptr;  // n_agent_voxels
    int* agent_seg_start = nullptr;  // n_agent_voxels + 1
    unsigned* agent_bits = nullptr;  // voxels with agents: bit v%32 of word v/32 (+2 spare words)
    void* agent_t1 = nullptr;        // n_agents*S, grouped order, FP64
    void* agent_t2 = nullptr;
    // Agents set from rates, for the adjoint: c = dt*V/V_voxel per agent and
    // the saturation densities (n_agents*S), grouped order, the dt they were
    // built for, and the list index of each grouped agent.
    bool agent_rates = false;
    double agent_dt = 0;
    double* agent_c = nullptr;
    double* agent_sat = nullptr;
    std::vector<int> agent_order;

    // Recorded steps for the adjoint (adjoint.hip).
    struct Tape* tape = nullptr;

    // Gradient vectors (nvs*3, store precision), allocated on first use.
    void* grad = nullptr;

    // Workspace of bfvm_mass_balance (reduce.hip), allocated on first use.
    void* red = nullptr;
    size_t red_bytes = 0;

    // MAT v4 snapshot buffers and writer thread (io.hip).
    struct MatWriter* mat = nullptr;

    // Steps replayed from one HIP graph by bfvm_run and bfvm_timed_run
    // (without per-launch events); 0 = every kernel launched on its own.
    // The graph (capi.hip, StepGraph) is captured from the same launches an
    // eager step makes and rebuilt when anything baked into them changes.
    int graph_steps = 8;
    struct StepGraph* graph = nullptr;

    // Per-launch timing (bfvm_timed_run with per-kernel events).
    bool kernel_timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_marks;
    size_t ev_next = 0;

    std::string err;
};

// Launchers (device precision chosen from env.precision).  Each returns the
// hipError_t of the launch.
hipError_t launch_dirichlet(Env& e);
hipError_t launch_bulk_prepare(Env& e, double dt);
hipError_t launch_bulk(Env& e);
// with_bulk: the bulk update of each agent voxel first (the fused x sweep
// then applies it to the voxels without agents only)
hipError_t launch_agents(Env& e, bool with_bulk = false);
// launch_agents (no bulk update) with each agent voxel's values as they were
// before it copied to save (n_agent_voxels * S values of the store's type),
// and the copy written back to the store.
hipError_t launch_agents_saving(Env& e, void* save);
hipError_t launch_agents_restore(Env& e, const void* save);
// launch_agents_restore on prev_rho (prev_save) and launch_agents_saving on
// e.rho (save) in one launch; prev_rho is another store than e.rho.
hipError_t launch_agents_restore_then_save(Env& e, void* save, void* prev_rho, const void* prev_save);
// Gradient vectors of the store into out (nvs*3 values, store precision).
hipError_t launch_gradient(Env& e, void* out);
hipError_t launch_sweep(Env& e, int axis);
// The sweep of `axis` reading store `src` and writing store `dst` (same shape
// and precision as the handle's; src is left unchanged).
hipError_t launch_sweep_to(Env& e, int axis, const void* src, void* dst);
// The adjoint's sweep of `axis` on the FP64 adjoint state lam (in place) with
// that axis's D/lambda sums over lam and the forward sweep's output y fused
// in, added to per-tile partial slots: per chunk of 8 substrates, tiles x 16
// doubles (the edge sums of the chunk's substrates, then their l.y sums).
// Call with e.precision and e.elem set to FP64 (the adjoint state's).
constexpr int kAdjointSumsNodes = 8;
constexpr int kAdjointSumsChunk = 8;  // substrates per chunk of the partial slots
struct AdjointSumsShape {
    int cols = 0;      // lines per block; 0: the fused kernel cannot run this axis
    size_t tiles = 0;  // blocks of one launch (slots per chunk of `partial`)
    int xt_rows = 0;   // x: rows per block of the whole-cache-line form (0: column tiles)
};
AdjointSumsShape adjoint_sums_shape(const Env& e, int axis, int y_precision);
hipError_t launch_adjoint_sums(Env& e, int axis, double* lam, const void* y, int y_precision, double* partial);
// Whether this grid's x sweep can run with sources fused in (a segmented
// one-block-per-tile shape: the autotuner's fastest, or the default); if so,
// launch_sweep_x_fused runs it with, in its load prologue, the bulk update
// (if `bulk`; skipping voxels with agents if `skip_agent_voxels`, whose bulk
// update the agent kernel applied) and the Dirichlet overwrite before the
// sweep, and in its store epilogue the Dirichlet overwrite of the next sweep
// (if `dir_after`).
bool sweep_x_fusable(const Env& e);
hipError_t launch_sweep_x_fused(Env& e, bool bulk, bool skip_agent_voxels, bool dir_after);
// agent_bits from the grouped agents (after each agent set-up).
hipError_t build_agent_bits(Env& e);
// x or y sweep of planes (z index) [k0, k1) only, for plane-chunked steps;
// valid when sweep_chunkable(e, axis) (the segmented kernels run the axis).
bool sweep_chunkable(const Env& e, int axis);
hipError_t launch_sweep_planes(Env& e, int axis, int k0, int k1);
// Ensemble sweeps: whether a member's axis can join a launch with L-node
// segments, its table entry, its dynamic LDS bytes, and the launch itself
// (kMaxThreads threads per block, `blocks` blocks in total).
bool batch_member_ok(const Env& e, int axis, int L);
SweepArgs batch_args(const Env& e, int axis, int L);
size_t batch_lds(const SweepArgs& A);
// Row-tile arguments of a grid's whole x sweep for an ensemble's batched
// row-tile launch (false: the grid cannot use row tiles of L-node segments).
bool batch_rows_args(const Env& e, int L, RowTile* X, int* threads, size_t* lds, int* tiles);
hipError_t launch_rows_batch(int precision, int L, int spread, const RowTile* d_tabs, const int* d_map,
                             int blocks, int threads, size_t lds, hipStream_t st);
hipError_t launch_sweep_batch(int precision, int L, int spread, const SweepArgs* d_args, const int* d_map,
                              int nm, int blocks, size_t lds, hipStream_t st, bool nt = false);
// Whole-cache-line x sweep arguments of a grid for an ensemble's batched
// launch of `threads`-thread blocks (false: the grid cannot take part: FP32,
// rows of a multiple of 4 nodes, L of 8, 16 or 20), and that launch.
bool batch_xt_args(const Env& e, int L, int threads, SweepArgs* A, size_t* lds, int* tiles);
hipError_t launch_xt_batch(int L, int spread, const SweepArgs* d_args, const int* d_map, int blocks, int threads,
                           size_t lds, hipStream_t st);

// Agents of several grids in one launch: one entry per member.
struct AgentBatch {
    void* rho;
    const int* gvox;
    const int* gstart;
    const double* t1;
    const double* t2;
    int S, n_groups;
    long long first;  // first global thread of this member (a multiple of kAgentBatchBlock)
};
// Block size of k_agents_batch: each member's threads start on a block
// boundary, so block b works for member block_member[b] alone.
constexpr int kAgentBatchBlock = 256;
hipError_t launch_agents_batch(int precision, const AgentBatch* d_tab, const int* d_block_member, int blocks,
                               hipStream_t st);

// Thomas/Woodbury tables for dt (rebuilt when dt, D or lambda changed);
// bumps Env::coef_epoch when it rebuilds.
int ensure_coefficients(Env& e, double dt);
// Bulk implicit terms for dt (rebuilt when dt changed).
hipError_t ensure_bulk(Env& e, double dt);

// The decay part of one sweep: dt*lambda/3 for LOD_3D (BioFVM_solvers.cpp:118-119),
// dt*lambda*0.5 for LOD_2D (:313-314), in the reference's operation order.
inline double decay_split(double lambda, double dt, int lod_dims)
{
    double c2 = lambda;
    c2 *= dt;
    if (lod_dims == 2) c2 *= 0.5;
    else c2 /= 3.0;
    return c2;
}

// Mass balance (reduce.hip): the workspace layout for this grid (pointers
// into `workspace`, which may be null to learn the size) and the two
// launches that write the S sums and the non-finite count to device memory.
struct MassLayout {
    int nblk = 0;
    double* part = nullptr;
    unsigned long long* bad = nullptr;
    double* sums = nullptr;
    long long* count = nullptr;
    size_t bytes = 0;
};
MassLayout mass_layout(const Env& e, void* workspace);
hipError_t launch_mass_balance(Env& e, const MassLayout& m, double* sums, long long* nonfinite);

// Frees the adjoint's tape, if any.
// Stable ascending sort of n (key, value) pairs by the low `bits` bits of
// the keys (sort.hip), and an exclusive prefix sum of n ints; tmp holds at
// least the bytes the *_tmp_bytes functions give.  Enqueued on `st`.
size_t radix_sort_tmp_bytes(int n);
hipError_t radix_sort_pairs(const unsigned* key_in, const int* val_in, unsigned* key_out, int* val_out, int n,
                            int bits, void* tmp, size_t tmp_bytes, hipStream_t st);
size_t exclusive_scan_tmp_bytes(int n);
hipError_t exclusive_scan(const int* in, int* out, int n, void* tmp, size_t tmp_bytes, hipStream_t st);

void free_tape(Env& e);
// Waits for and frees the snapshot writer, if any.
void free_writer(Env& e);

// Segment length of an axis unless the autotuner may choose it (lod.hip).
int default_segment(const Env& e, int axis);
// Valid segmented shapes of one axis for this grid and store precision: of
// every segment length, or only of default_segment's (what the autotuner
// times unless the handle's tune_segments is set).
std::vector<SweepPlan> sweep_candidates(const Env& e, int axis, bool any_length);
// The launch shape a sweep of `axis` uses now (autotuned or default).
SweepPlan effective_plan(const Env& e, int axis);
// Whether plan p of `axis` may use non-temporal loads and stores (SweepPlan::nt).
bool sweep_nt_allowed(const Env& e, int axis, const SweepPlan& p);

}  // namespace bfvm

// The opaque handle type of the C ABI is the internal state.
struct bfvm_env : public bfvm::Env {};
