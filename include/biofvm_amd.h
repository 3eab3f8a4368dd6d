/*
 * biofvm_amd — MI355X (gfx950) diffusion-decay engine, C ABI.
 *
 * Drop-in for BioFVM 1.0.4's hot path: the LOD diffusion-decay solver plugged
 * into Microenvironment::diffusion_decay_solver, plus the bulk and per-agent
 * source/sink steps that run beside it each time step.  Densities live in HBM
 * as one array rho[((k*ny + j)*nx + i)*S + s] (voxel-major, substrate
 * innermost: the reference's vector<vector<double>> store flattened).
 *
 * Conventions
 *   - Every function returns 0 on success or a negative BFVM_E* code; the
 *     message is available from bfvm_last_error(handle) (or
 *     bfvm_last_error(NULL) for failures that have no handle yet).
 *   - Work is enqueued on the handle's own HIP stream and is asynchronous;
 *     functions that read results back to host memory synchronise first.
 *   - One host thread per handle; handles are independent.
 *   - Host arrays are float64, voxel-major / substrate-innermost, whatever
 *     the device precision.
 */
#ifndef BIOFVM_AMD_H
#define BIOFVM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfvm_env* bfvm_handle;

/* Device arithmetic of the density store and the sweeps. */
enum bfvm_precision {
    BFVM_FP64 = 0, /* bit-for-bit reference semantics up to FMA rounding */
    BFVM_FP32 = 1  /* FP32 store (half the HBM bytes); sweep and source arithmetic and
                      all coefficient tables stay FP64, one rounding per store write */
};

/* Phases of one time step, in the order they run
 * (examples/convergence_test5.cpp:236-238). */
enum bfvm_phase {
    BFVM_BULK = 1,      /* Microenvironment::simulate_bulk_sources_and_sinks */
    BFVM_CELLS = 2,     /* Microenvironment::simulate_cell_sources_and_sinks */
    BFVM_DIFFUSION = 4, /* Microenvironment::simulate_diffusion_decay (LOD_3D) */
    BFVM_ALL = 7
};

enum bfvm_status {
    BFVM_OK = 0,
    BFVM_E_ARG = -1,     /* invalid argument (sizes, null pointers, ranges) */
    BFVM_E_DEVICE = -2,  /* HIP runtime error (allocation, launch, copy) */
    BFVM_E_STATE = -3,   /* call not valid in the handle's current state */
    BFVM_E_MESH = -4     /* non-uniform mesh: the LOD solver refuses it */
};

/*
 * Creates a uniform Cartesian grid of nx*ny*nz voxels (edge dx) carrying
 * n_substrates densities, all zero.  Replaces Microenvironment's
 * set_density/add_density + resize_space_uniform
 * (BioFVM_microenvironment.cpp:287-290, 325-458; mesh BioFVM_mesh.cpp:796-922).
 */
int bfvm_create(int nx, int ny, int nz, int n_substrates, double dx, int precision,
                bfvm_handle* out);
int bfvm_destroy(bfvm_handle h);

/*
 * Same, for a mesh with separate dx, dy, dz.  The LOD solver only accepts a
 * uniform mesh (|dx-dy|, |dy-dz|, |dx-dz| <= 1e-16, BioFVM_mesh.cpp:811-815);
 * otherwise the handle is created but every BFVM_DIFFUSION step returns
 * BFVM_E_MESH and leaves the densities untouched, as the reference prints an
 * error and returns (BioFVM_solvers.cpp:79-83).
 */
int bfvm_create_mesh(int nx, int ny, int nz, int n_substrates, double dx, double dy, double dz,
                     int precision, bfvm_handle* out);

/* Diffusion coefficients D[s] and decay rates lambda[s]
 * (Microenvironment::diffusion_coefficients / decay_rates, microenvironment.h:118-119).
 * The Thomas factorisation (BioFVM_solvers.cpp:87-173) is recomputed in FP64
 * whenever D, lambda or dt change. */
int bfvm_set_coefficients(bfvm_handle h, const double* D, const double* lambda);

/* Whole density store, host <-> device (density_vector(n), microenvironment.h:186-191). */
int bfvm_upload_density(bfvm_handle h, const double* host);
int bfvm_download_density(bfvm_handle h, double* host);

/*
 * Dirichlet nodes: n voxel indices and n*S values.  Replaces the whole node
 * list built by add_dirichlet_node (BioFVM_microenvironment.cpp:146-151) and
 * applied by apply_dirichlet_conditions (:192-197) before each of the x, y
 * and z sweeps.  mask[S] (optional) limits the overwrite to substrates with
 * mask[s] != 0; NULL = whole vector, the reference's behaviour.  n = 0 clears.
 */
int bfvm_set_dirichlet(bfvm_handle h, int n, const int* voxel_index, const double* values,
                       const unsigned char* mask);

/*
 * Bulk source/sink fields, nvox*S each: supply rate S, supply target T and
 * uptake rate U — the values the three bulk callbacks write
 * (microenvironment.h:130-132, used at BioFVM_microenvironment.cpp:595-610).
 * All NULL removes bulk sources.
 */
int bfvm_set_bulk(bfvm_handle h, const double* supply_rate, const double* supply_target,
                  const double* uptake_rate);

/*
 * Agents in list order (the order of all_basic_agents, BioFVM_basic_agent.cpp:35):
 * voxel index and the implicit constants temp1/temp2 (n*S each) that
 * Basic_Agent::set_internal_uptake_constants computes
 * (BioFVM_basic_agent.cpp:114-148).  Each step applies, per agent in list
 * order, rho[v] = (rho[v] + temp1) / temp2 (:229-238) — on the device the
 * agents are grouped by voxel, keeping list order inside a voxel, so the
 * result is deterministic.  n = 0 clears.
 */
int bfvm_set_agents(bfvm_handle h, int n, const int* voxel_index, const double* temp1,
                    const double* temp2);

/*
 * Agents from rates: positions (3n), volumes (n), secretion, saturation and
 * uptake rates (n*S each).  Voxels come from Cartesian_Mesh::nearest_voxel_index
 * (BioFVM_mesh.cpp:930-948) on a mesh spanning [0, n*dx] per axis, and the
 * constants from set_internal_uptake_constants(dt) (BioFVM_basic_agent.cpp:114-148).
 * A position outside the box is rejected with BFVM_E_ARG (the reference's
 * assign_position refuses it, BioFVM_basic_agent.cpp:66-71).
 * out_voxel (optional, n ints) receives the voxel of each agent.
 */
int bfvm_set_agent_rates(bfvm_handle h, int n, const double* position, const double* volume,
                         const double* secretion, const double* saturation,
                         const double* uptake, double dt, int* out_voxel);

/*
 * bfvm_set_agent_rates with every array in device memory (FP64, the caller's):
 * voxels, constants and the grouping by voxel are computed on the GPU (a
 * stable radix sort keeps list order inside a voxel), with the same results
 * as bfvm_set_agent_rates on the same values.  Agents that move are re-binned
 * by calling it again with the new positions.  out_voxel (optional) is a
 * device array of n ints.  Returns BFVM_E_ARG if an agent lies outside the
 * domain (after the kernels ran; the handle then has no agents).
 */
int bfvm_set_agent_rates_device(bfvm_handle h, int n, const double* position, const double* volume,
                                const double* secretion, const double* saturation,
                                const double* uptake, double dt, int* out_voxel);

/*
 * One time step: the phases in `flags`, in the order bulk -> cells ->
 * diffusion.  BFVM_DIFFUSION replaces
 * diffusion_decay_solver__constant_coefficients_LOD_3D
 * (BioFVM_solvers.cpp:77-276) called through simulate_diffusion_decay
 * (BioFVM_microenvironment.cpp:502-513).
 */
int bfvm_step(bfvm_handle h, double dt, int flags);

/* `steps` consecutive calls of bfvm_step, enqueued back to back (whole
 * groups of steps replayed from a HIP graph, bfvm_set_graph_steps). */
int bfvm_run(bfvm_handle h, double dt, int steps, int flags);

/* Steps per HIP graph for bfvm_run and bfvm_timed_run (without per-launch
 * events).  A run of K steps makes its first step launch by launch, captures
 * the launches of the next steps_per_graph steps into a graph (once: it is
 * kept for later runs), replays it while whole graphs fit and launches the
 * rest one by one.  The graph holds the kernels of the eager steps with the
 * same arguments in the same order, so the results are the same bits as with
 * steps_per_graph = 0 (every kernel launched on its own).  It is captured
 * again after the tables, sources, agents, launch shapes, dt or phases
 * change.  Default 8.  No counterpart in the reference (its steps are host
 * loops). */
int bfvm_set_graph_steps(bfvm_handle h, int steps_per_graph);
int bfvm_get_graph_steps(bfvm_handle h, int* steps_per_graph);

/* Waits for all work enqueued on the handle. */
int bfvm_sync(bfvm_handle h);

/*
 * Runs `steps` steps and times them on the handle's stream with HIP events.
 * ms_total receives the elapsed time of the whole run.  If per_kernel_ms is
 * non-NULL, each kernel launch is bracketed by its own events and
 * per_kernel_ms[kind] (kind in enum bfvm_kernel_kind, array of
 * BFVM_KERNEL_KINDS doubles) receives the summed time of that kind, and
 * launches[kind] (optional) the number of launches.
 */
enum bfvm_kernel_kind {
    BFVM_K_DIRICHLET = 0,
    BFVM_K_BULK = 1,
    BFVM_K_AGENTS = 2,
    BFVM_K_SWEEP_X = 3,
    BFVM_K_SWEEP_Y = 4,
    BFVM_K_SWEEP_Z = 5,
    BFVM_KERNEL_KINDS = 6
};
int bfvm_timed_run(bfvm_handle h, double dt, int steps, int flags, double* ms_total,
                   double* per_kernel_ms, int* launches);

/*
 * Autotuner: times every valid launch shape of the three segmented sweeps
 * (lines per block x kernel) for this grid, substrate count and precision on
 * a scratch copy of the store, keeps the fastest per axis (y and z: with
 * either cache policy, bfvm_set_sweep_cache_policy), then its XCD spread
 * (bfvm_set_sweep_spread), then (with bulk fields or Dirichlet nodes set and
 * fused sources on) the one-block-per-tile shapes of the x sweep with the
 * sources fused in, then times the x/y plane chunk sizes
 * (bfvm_set_sweep_chunk).  The densities are left unchanged.  No counterpart
 * in the reference (its OpenMP loops have no launch shape).
 *
 * Reproducibility: none of the knobs above changes a result bit, so results
 * do not depend on which shape won the timing.  The one parameter that does
 * change the rounding of the solve, the number of nodes per segment, stays at
 * a fixed function of the line length, axis and store precision (for y and
 * z lines of 400-640 nodes in an FP32 store 20, longer ones 32; 16 for lines
 * of >= 200 nodes, 8 below) unless bfvm_set_autotune_segments(h, 1) lets the autotuner time the
 * segment lengths too; then results may differ in their last bits between
 * runs whose timings pick different lengths.
 */
int bfvm_autotune(bfvm_handle h, double dt);
int bfvm_set_autotune_segments(bfvm_handle h, int on);
int bfvm_get_autotune_segments(bfvm_handle h, int* on);

/* Number of times the handle built its Thomas/Woodbury tables (at the first
 * diffusion step and after D, lambda, dt or the LOD dimension changed;
 * bfvm_set_coefficients with the values already set changes nothing).  The
 * reference builds them once per Microenvironment (BioFVM_solvers.cpp:87,172). */
int bfvm_get_coefficient_builds(bfvm_handle h, unsigned* builds);

/* The launch shape in use for one axis (0 = x, 1 = y, 2 = z):
 * lines per block, nodes per segment, and the kernel ("streaming"): 0 one
 * block per column tile, 1 resident blocks that load the next tile while
 * solving one, 2 (x only) tiles of whole rows loaded contiguously and
 * transposed through LDS, lines_per_block = rows per tile x substrates,
 * 3 (x only, a node's values whole 16-byte pieces: FP32 with 4, 8 or 16
 * substrates, FP64 with 2, 4, 8 or 16) 16-byte loads and stores of whole
 * 4-node (FP32) or 2-node (FP64) chunks handed to the lines' lanes through
 * LDS, lines_per_block = rows per block x substrates, nodes per segment 8,
 * 16 or 20; *lines_per_block = 0 means the one-thread-per-line fallback.
 * Any output pointer may be NULL. */
int bfvm_get_sweep_plan(bfvm_handle h, int axis, int* lines_per_block, int* nodes_per_segment,
                        int* streaming);

/* Sets the launch shape of one axis by hand (lines_per_block = 0 restores
 * the default).  Returns BFVM_E_ARG for a shape the kernel cannot run
 * (for kernels 2 and 3, the shapes bfvm_autotune would try).  Every shape
 * with the same nodes per segment gives the same result bits. */
int bfvm_set_sweep_plan(bfvm_handle h, int axis, int lines_per_block, int nodes_per_segment,
                        int streaming);

/* Cache policy of one axis's sweep: non_temporal = 1 when its loads and
 * stores of the store carry the non-temporal (streaming) hint, 0 for the
 * default.  Offered for the y and z sweeps whose tiles' row pieces are whole
 * 128-byte lines (lines per block x value size >= 128, column kernels) and
 * for x sweeps on the whole-cache-line kernel (streaming = 3): each value is
 * read and written once per sweep, and the hint keeps the sweep from
 * evicting the other kernels' data and from leaving its own lines for the
 * next kernel to meet; bfvm_autotune times both (for the x sweep, together
 * with the y sweep after it).
 * Setting 1 elsewhere returns BFVM_E_ARG; setting a launch shape by hand
 * (bfvm_set_sweep_plan) resets the axis to the default policy.  Results do
 * not depend on it.  No counterpart in the reference. */
int bfvm_get_sweep_cache_policy(bfvm_handle h, int axis, int* non_temporal);
int bfvm_set_sweep_cache_policy(bfvm_handle h, int axis, int non_temporal);

/* Whether a step runs its bulk update and the Dirichlet overwrites before the
 * x and y sweeps inside the x sweep (on = 1, the default) or as kernels of
 * their own (0).  Fused, the thread that loads a node for the x sweep applies
 * the bulk update and the Dirichlet value, in the reference's order (bulk,
 * agents, Dirichlet: BioFVM_microenvironment.cpp:584-633, BioFVM_solvers.cpp:
 * 177-207), and the store writes the Dirichlet value the y sweep must see;
 * the agent kernel, which runs before the x sweep, applies the bulk update of
 * the voxels holding agents.  The store pass of the bulk kernel and two
 * Dirichlet launches are saved.  Same operations and roundings either way:
 * the same bits for either store precision.  Used when the x sweep can run
 * a one-block-per-tile shape (the autotuner's fastest for the fused sweep,
 * timed with the handle's sources) and no x/y plane chunking is set. */
int bfvm_set_fused_sources(bfvm_handle h, int on);
int bfvm_get_fused_sources(bfvm_handle h, int* on);

/* How the column tiles of one axis's sweep are spread over the GPU's 8 XCDs
 * (each with its own L2; workgroups are dealt to them round robin):
 * 0 = as dispatched, k > 0 = each XCD runs chunks of k consecutive tiles,
 * -1 = each XCD runs one contiguous eighth of the tiles (the default).  It
 * changes which workgroup computes which tile, never the result.  Set by
 * bfvm_autotune; no counterpart in the reference. */
int bfvm_get_sweep_spread(bfvm_handle h, int axis, int* tiles_per_chunk);
int bfvm_set_sweep_spread(bfvm_handle h, int axis, int tiles_per_chunk);

/*
 * Mass balance: sums[s] = sum over voxels of rho[v*S + s], in FP64 whatever
 * the store precision, with a fixed summation order (the same store gives the
 * same bits on every run), and *nonfinite = the number of values that are NaN
 * or infinite.  Either output may be NULL (not both).  Every column of each
 * LOD sweep matrix sums to 1 + c2 (BioFVM_solvers.cpp:121-135), so without
 * sources, Dirichlet nodes or agents one LOD_3D step multiplies each sum by
 * (1 + dt*lambda/3)^-3 exactly; the reference has no such check (its solver
 * and examples never sum the store).  Waits for the handle's work; the
 * workspace (about 8*(S+1) KiB) is allocated on the first call.
 */
int bfvm_mass_balance(bfvm_handle h, double* sums, long long* nonfinite);

/*
 * bfvm_mass_balance for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same kernels and the
 * same bits as bfvm_mass_balance, but nothing is allocated and nothing waits:
 * the reduction is enqueued on the handle's stream and writes sums[S] and
 * *nonfinite to DEVICE memory owned by the caller (either may be NULL).  Its
 * workspace is a static device array of 520 KiB per GPU, held until the
 * process ends, which limits it to n_substrates <= 64 (BFVM_E_ARG beyond).
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams at once) race; calls in order on one stream do not.
 */
int bfvm_mass_balance_noalloc(bfvm_handle h, double* sums, long long* nonfinite);

/*
 * Gradient vectors of the current densities (Microenvironment::
 * compute_all_gradient_vectors, BioFVM_microenvironment.cpp:699-779): per
 * voxel n and substrate s, for each axis with n strictly inside,
 * (rho[n+jump] - rho[n-jump]) / (2*d_axis) in FP64, else 0.  Enqueued on the
 * handle's stream into a device array of nvox*S*3 values in the store's
 * precision, index (n*S + s)*3 + axis.  Unlike the reference (whose
 * neighbour strides are set only once its solver has run, and whose 2*dx is
 * fixed by the first mesh of the process), the strides and spacings are the
 * handle's own.
 */
int bfvm_compute_gradients(bfvm_handle h);
int bfvm_download_gradients(bfvm_handle h, double* host);
int bfvm_gradient_device_ptr(bfvm_handle h, void** ptr);

/*
 * Snapshot of the densities as a MAT v4 file in the layout of
 * Microenvironment::write_to_matlab (BioFVM_microenvironment.cpp:557-580,
 * header BioFVM_matlab.cpp:515-559): a (4+S) x voxels FP64 matrix named
 * "multiscale_microenvironment", one column per voxel: centre x, y, z
 * (x0 + (i+0.5)*dx, ...: the mesh's origin is the caller's), volume, the S
 * densities.  bfvm_write_matlab returns when the file is written;
 * bfvm_write_matlab_async returns at once (the store is captured at that
 * point of the handle's stream, later steps may run) and bfvm_io_wait waits
 * for the file and reports its errors.  One asynchronous snapshot is in flight
 * at a time (a second call waits for the first).
 */
int bfvm_write_matlab(bfvm_handle h, const char* filename, double x0, double y0, double z0);
int bfvm_write_matlab_async(bfvm_handle h, const char* filename, double x0, double y0, double z0);
int bfvm_io_wait(bfvm_handle h);

/*
 * Solver choice: dims = 3 is diffusion_decay_solver__constant_coefficients_LOD_3D
 * (the default), dims = 2 is ..._LOD_2D (BioFVM_solvers.cpp:278-427): Dirichlet,
 * x sweep, Dirichlet, y sweep, with the decay split in halves
 * (dt*lambda*0.5 per sweep).  dims = 2 needs nz == 1 (BFVM_E_MESH otherwise;
 * the reference's 2-D solver sweeps only the k = 0 plane).
 */
int bfvm_set_lod_dimensions(bfvm_handle h, int dims);

/* Plane chunking of the x and y sweeps: with planes > 0 (and fewer than nz)
 * the x and y sweeps run chunk by chunk of `planes` z planes, x then y per
 * chunk, so the y sweep reads what the x sweep just wrote from the GPU's
 * last-level cache.  Results are bit-identical to whole-store sweeps;
 * 0 = whole-store sweeps.  bfvm_autotune picks it. */
int bfvm_get_sweep_chunk(bfvm_handle h, int* planes);
int bfvm_set_sweep_chunk(bfvm_handle h, int planes);

/* Grid and store description. */
int bfvm_get_info(bfvm_handle h, int* nx, int* ny, int* nz, int* n_substrates, int* precision,
                  size_t* density_bytes);

/* Device address of the density store (for zero-copy interop). */
int bfvm_density_device_ptr(bfvm_handle h, void** ptr);

/* Message for the last failure on `h` (or of the calling thread when h is NULL). */
const char* bfvm_last_error(bfvm_handle h);

/*
 * Ensembles.  The reference advances one Microenvironment per call of
 * simulate_bulk_sources_and_sinks / simulate_cell_sources_and_sinks /
 * simulate_diffusion_decay (BioFVM_microenvironment.cpp:502-633); an
 * ensemble advances `count` independent handles (uneven grids and substrate
 * counts, one store precision) with one batched launch per phase and sweep
 * axis.  The members keep their own data and entry points; while an
 * ensemble runs, do not enqueue work on a member's own handle without
 * bfvm_ensemble_sync first.  Each member's results equal those of
 * bfvm_step on that member alone.
 */
typedef struct bfvm_ensemble_t* bfvm_ensemble;

/* Builds an ensemble over `count` existing handles (all of one precision). */
int bfvm_ensemble_create(int count, const bfvm_handle* members, bfvm_ensemble* out);
int bfvm_ensemble_destroy(bfvm_ensemble e);

/* `steps` steps of every member (phases in `flags`, as bfvm_step).  With
 * use_graph != 0 one step is captured into a HIP graph (once per dt and
 * flags) and replayed `steps` times.  Asynchronous. */
int bfvm_ensemble_run(bfvm_ensemble e, double dt, int steps, int flags, int use_graph);

/* Steps per graph replay of bfvm_ensemble_run (default 8): runs of that many
 * steps replay one graph holding them (one launch gap instead of one per
 * step), the rest one-step graphs; the same launches in the same order, the
 * same bits.  No counterpart in the reference. */
int bfvm_ensemble_set_graph_steps(bfvm_ensemble e, int steps_per_graph);
int bfvm_ensemble_get_graph_steps(bfvm_ensemble e, int* steps_per_graph);

/* Waits for all work enqueued on the ensemble. */
int bfvm_ensemble_sync(bfvm_ensemble e);

/* bfvm_ensemble_run timed with HIP events on the ensemble's stream
 * (preparation of coefficients and launch tables is outside the timing). */
int bfvm_ensemble_timed_run(bfvm_ensemble e, double dt, int steps, int flags, int use_graph,
                            double* ms_total);

/* Times the batched sweeps for each segment length (4, 8, 16, 32 nodes; only
 * the ensemble's own unless bfvm_ensemble_set_autotune_segments), spread of
 * their blocks over the XCDs (as dispatched, chunks of 64 tiles,
 * contiguous eighths), x kernel (column tiles, row tiles, whole-cache-line
 * chunks) and y/z cache policy (default or non-temporal) on scratch copies
 * of the stores and keeps the fastest.  Densities unchanged. */
int bfvm_ensemble_autotune(bfvm_ensemble e, double dt);
/* Nodes per segment of the batched sweeps: 0 (the default) = each axis at
 * the length the members' own sweeps take (the most common one among them,
 * by values; a member whose own length differs sweeps that axis on its own),
 * so every member's result equals its solo run bit for bit; 4, 8, 16 or 32 =
 * that length for every axis and member. */
int bfvm_ensemble_get_segment(bfvm_ensemble e, int* nodes_per_segment);
int bfvm_ensemble_set_segment(bfvm_ensemble e, int nodes_per_segment);
/* The length the batched sweep of `axis` runs at (0 = x, 1 = y, 2 = z). */
int bfvm_ensemble_get_axis_segment(bfvm_ensemble e, int axis, int* nodes_per_segment);
/* Whether bfvm_ensemble_autotune also times the segment lengths (1: the
 * members' own, 4, 8, 16 and 32), which changes the rounding of the batched
 * solve, or keeps the ensemble's segment length (0, the default) and tunes
 * only the knobs that leave the bits unchanged. */
int bfvm_ensemble_set_autotune_segments(bfvm_ensemble e, int on);
/* Whether the members' x sweeps run on LDS-staged row tiles (on = 1, the
 * default; members that cannot take part run with column tiles) or all on
 * column tiles (0).  Row tiles load whole x rows contiguously, which matters
 * for few substrates per voxel; the result is the same bit for bit.  Set by
 * bfvm_ensemble_autotune. */
int bfvm_ensemble_set_row_tiles(bfvm_ensemble e, int on);
int bfvm_ensemble_get_row_tiles(bfvm_ensemble e, int* on);
/* Whether the members' x sweeps load and store whole cache lines (on = 1,
 * the default: 16-byte pieces of 4-node chunks handed to the lines' lanes
 * through LDS, for FP32 members whose x rows are a multiple of 4 nodes and
 * at segment lengths of 8, 16 or 20; the others run on row or column tiles).
 * The result is the same bit for bit.  Set by bfvm_ensemble_autotune.  No
 * counterpart in the reference. */
int bfvm_ensemble_set_x_chunks(bfvm_ensemble e, int on);
int bfvm_ensemble_get_x_chunks(bfvm_ensemble e, int* on);

/* Whether the batched y and z sweeps load and store with the non-temporal
 * (streaming) hint (1) or the default policy (0; see
 * bfvm_set_sweep_cache_policy); bfvm_ensemble_autotune times both.  Results
 * do not depend on it.  No counterpart in the reference. */
int bfvm_ensemble_set_cache_policy(bfvm_ensemble e, int non_temporal);
int bfvm_ensemble_get_cache_policy(bfvm_ensemble e, int* non_temporal);

const char* bfvm_ensemble_last_error(bfvm_ensemble e);

/*
 * Differentiation: gradients of a scalar loss J(rho after `steps` steps)
 * with respect to the initial densities, D, lambda and the agents'
 * secretion and uptake rates, by a hand-written adjoint of the step.  No
 * counterpart in the reference (it has no derivatives); the forward steps
 * are exactly bfvm_step's.
 *
 *   bfvm_tape_record    runs `steps` steps from the current store (as
 *                       bfvm_run) keeping the state every checkpoint_every
 *                       steps; the store ends at rho_N.
 *   bfvm_tape_seed      dJ/drho_N from the caller (host, nvox*S), or
 *   bfvm_tape_seed_l2   J = sum (rho_N - target)^2 / 2, seed rho_N - target;
 *                       *loss receives J (FP64 sum).
 *   bfvm_tape_backward  runs the adjoint (recomputing each checkpoint
 *                       segment) and writes the gradients: grad_rho0 (nvox*S),
 *                       grad_D and grad_lambda (S), grad_secretion and
 *                       grad_uptake (n_agents*S, list order; agents must have
 *                       been set with bfvm_set_agent_rates).  Any output may be
 *                       NULL.  The coefficients, sources and agents must be
 *                       those the tape was recorded with.
 *   bfvm_tape_free      releases the tape's device memory
 *                       ((steps/K + K + 4) stores and an FP64 adjoint state,
 *                       plus 2K + 1 stores, 3K + 1 with bulk fields, while
 *                       the backward pass keeps every step's states).
 * The recomputed forward states are in the store's precision; the adjoint
 * state, all arithmetic and the gradient sums are FP64.
 *   bfvm_set_tape_keep_states  with on = 1 (the default) the backward pass
 *                       recomputes a segment's steps once, keeping the states
 *                       each step's transposes read (2 or 3 stores per step
 *                       of the segment), where they fit in half the free
 *                       device memory; otherwise (on = 0, or too little
 *                       memory) it recomputes each step twice (the step
 *                       starts, then each step again).  The same gradients
 *                       bit for bit either way.  No counterpart in the
 *                       reference.
 */
int bfvm_tape_record(bfvm_handle h, double dt, int steps, int flags, int checkpoint_every);
int bfvm_tape_seed(bfvm_handle h, const double* seed);
int bfvm_tape_seed_l2(bfvm_handle h, const double* target, double* loss);
int bfvm_tape_backward(bfvm_handle h, double* grad_rho0, double* grad_D, double* grad_lambda,
                       double* grad_secretion, double* grad_uptake);
int bfvm_tape_free(bfvm_handle h);
int bfvm_set_tape_keep_states(bfvm_handle h, int on);
int bfvm_get_tape_keep_states(bfvm_handle h, int* on);
/* *kept = 1 when the tape's last backward pass kept every step's states. */
int bfvm_tape_states_kept(bfvm_handle h, int* kept);

#ifdef __cplusplus
}
#endif

#endif /* BIOFVM_AMD_H */
