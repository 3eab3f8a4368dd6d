// Ensembles: many independent grids advanced together, one launch per phase
// for all of them, optionally replayed from a HIP graph of one step.
//
// Members are ordinary handles (each with its own store, coefficients,
// Dirichlet nodes, bulk fields and agents, all of one precision).  A step of
// the ensemble runs, in the reference's order (bulk -> cells -> diffusion):
//   - bulk: one k_bulk launch per member that has bulk fields;
//   - cells: one k_agents_batch launch over every member's agent groups;
//   - diffusion, per axis: Dirichlet launches for members that have nodes,
//     then one k_sweep_batch launch over every member (members whose lines
//     do not fit the batched kernel take their own sweep launch).
// The batched sweep of an axis runs at the segment length the members take
// on their own (default_segment in lod.hip: the most common one among the
// members; members with another length sweep on their own), so a member's
// result equals its solo run bit for bit; bfvm_ensemble_set_segment fixes
// one length for every axis and member instead.
#include <algorithm>
#include <string>
#include <vector>

#include "bfvm_internal.h"

using bfvm::Env;

struct bfvm_ensemble_t {
    std::vector<Env*> m;
    int precision = BFVM_FP64;
    hipStream_t stream = nullptr;
    int L = 0;  // nodes per segment of every batched sweep; 0 = the members' own, per axis
    int spread = 0;  // XCD spread of the batched sweeps' blocks (xcd_tile in lod.hip)
    int nt = 0;      // y and z batched sweeps with non-temporal loads and stores (autotuned)
    int xt_x = 1;    // x sweeps with whole-cache-line chunk loads (k_sweep_xt_batch) where a member can
    int rows_x = 1;  // x sweeps on LDS-staged row tiles (k_sweep_rows_batch) where a member can
    // whether bfvm_ensemble_autotune may change L (which changes the rounding
    // of the batched solve); otherwise it tunes the other knobs at this L
    bool tune_segments = false;

    // tables, valid for (dt, L, member coefficient/agent epochs)
    bool tables_valid = false;
    double tab_dt = 0;
    int tab_L[3] = {0, 0, 0};
    std::vector<unsigned> coef_epoch, agent_epoch, source_epoch, plan_epoch;
    bfvm::SweepArgs* d_args[3] = {nullptr, nullptr, nullptr};
    int* d_first[3] = {nullptr, nullptr, nullptr};  // per block: member, block within the member
    int nbatch[3] = {0, 0, 0};       // members in the batched launch
    int blocks[3] = {0, 0, 0};
    size_t lds[3] = {0, 0, 0};
    std::vector<int> solo[3];        // members sweeping on their own
    // x sweeps on row tiles: per-member arguments, per-block (member, tile)
    bfvm::RowTile* d_rows = nullptr;
    int* d_rmap = nullptr;
    int rblocks = 0, rthreads = 0;
    size_t rlds = 0;
    int tab_rows_x = -1;
    // x sweeps with whole-cache-line chunk loads: per-member arguments, per-block (member, tile)
    bfvm::SweepArgs* d_xt = nullptr;
    int* d_xtmap = nullptr;
    int xtblocks = 0;
    size_t xtlds = 0;
    int tab_xt_x = -1;
    bfvm::AgentBatch* d_agents = nullptr;
    int* d_agent_blocks = nullptr;  // the member of each block of k_agents_batch
    int n_agent_members = 0;
    int agent_blocks = 0;

    // graphs of one step and of graph_steps steps (replays of the latter
    // carry runs of that many steps: one launch gap instead of one per step)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec_multi = nullptr;
    hipGraph_t graph_multi = nullptr;
    int graph_steps = 8;
    double graph_dt = 0;
    int graph_flags = -1;

    std::string err;
};

namespace {

// Threads per block of the batched whole-cache-line x sweep: each member's
// blocks hold as many rows as fill 8 waves with segments.
constexpr int kXtBatchThreads = 512;

int efail(bfvm_ensemble_t* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}

#define EHIP(e, call)                                                                          \
    do {                                                                                       \
        hipError_t st_ = (call);                                                               \
        if (st_ != hipSuccess) return efail(e, BFVM_E_DEVICE, std::string(#call) + ": " +      \
                                                                  bfvm::hip_error_text(st_));     \
    } while (0)

// Runs `f` with member `m` launching on the ensemble's stream.
template <typename F>
hipError_t on_stream(Env& m, hipStream_t st, F&& f)
{
    hipStream_t keep = m.stream;
    m.stream = st;
    hipError_t r = f();
    m.stream = keep;
    return r;
}

void free_tables(bfvm_ensemble_t& E)
{
    for (int a = 0; a < 3; ++a) {
        if (E.d_args[a]) (void)hipFree(E.d_args[a]);
        if (E.d_first[a]) (void)hipFree(E.d_first[a]);
        E.d_args[a] = nullptr;
        E.d_first[a] = nullptr;
        E.solo[a].clear();
    }
    if (E.d_agents) (void)hipFree(E.d_agents);
    E.d_agents = nullptr;
    if (E.d_agent_blocks) (void)hipFree(E.d_agent_blocks);
    E.d_agent_blocks = nullptr;
    if (E.d_rows) (void)hipFree(E.d_rows);
    if (E.d_rmap) (void)hipFree(E.d_rmap);
    E.d_rows = nullptr;
    E.d_rmap = nullptr;
    E.rblocks = 0;
    if (E.d_xt) (void)hipFree(E.d_xt);
    if (E.d_xtmap) (void)hipFree(E.d_xtmap);
    E.d_xt = nullptr;
    E.d_xtmap = nullptr;
    E.xtblocks = 0;
    E.xtlds = 0;
    E.tables_valid = false;
}

void drop_graph(bfvm_ensemble_t& E)
{
    if (E.exec) (void)hipGraphExecDestroy(E.exec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    if (E.exec_multi) (void)hipGraphExecDestroy(E.exec_multi);
    if (E.graph_multi) (void)hipGraphDestroy(E.graph_multi);
    E.exec = nullptr;
    E.graph = nullptr;
    E.exec_multi = nullptr;
    E.graph_multi = nullptr;
    E.graph_flags = -1;
}

// Host-side preparation of one step: coefficients and bulk terms of every
// member for dt, then the batched launch tables.  No kernel is enqueued
// here except bulk preparation, so a step captured after this is launches only.
// The segment length of the batched sweep of `axis`: the fixed one, or the
// one the members' own sweeps of that axis take most (by values swept; the
// shorter on a tie).
int axis_segment(const bfvm_ensemble_t& E, int axis)
{
    if (E.L > 0) return E.L;
    size_t best_n = 0;
    int best = 8;
    for (int L : {4, 8, 12, 16, 20, 24, 32}) {
        size_t n = 0;
        for (const Env* pm : E.m)
            if (axis < pm->lod_dims && bfvm::default_segment(*pm, axis) == L) n += pm->nvs;
        if (n > best_n) {
            best_n = n;
            best = L;
        }
    }
    return best;
}

// Whether member m takes part in the batched sweep of `axis` at length L:
// it fits the batched kernel and, with the members' own lengths, L is its own.
bool joins_batch(const bfvm_ensemble_t& E, const Env& m, int axis, int L)
{
    return m.coef_valid && bfvm::batch_member_ok(m, axis, L) && (E.L > 0 || bfvm::default_segment(m, axis) == L);
}

int prepare(bfvm_ensemble_t& E, double dt, int flags)
{
    int Lax[3];
    for (int a = 0; a < 3; ++a) Lax[a] = axis_segment(E, a);
    bool changed = !E.tables_valid || E.tab_dt != dt || E.tab_rows_x != E.rows_x || E.tab_xt_x != E.xt_x;
    for (int a = 0; a < 3; ++a) changed = changed || E.tab_L[a] != Lax[a];
    for (size_t i = 0; i < E.m.size(); ++i) {
        Env& m = *E.m[i];
        if (flags & BFVM_DIFFUSION) {
            if (!m.uniform) return efail(&E, BFVM_E_MESH, "member " + std::to_string(i) + " has a non-uniform mesh");
            if (m.D.empty()) return efail(&E, BFVM_E_STATE, "member " + std::to_string(i) + " has no coefficients");
            int rc = bfvm::ensure_coefficients(m, dt);
            if (rc) return efail(&E, rc, "member " + std::to_string(i) + ": " + m.err);
        }
        if (flags & BFVM_BULK) EHIP(&E, on_stream(m, E.stream, [&] { return bfvm::ensure_bulk(m, dt); }));
        if (E.coef_epoch[i] != m.coef_epoch || E.agent_epoch[i] != m.agent_epoch ||
            E.source_epoch[i] != m.source_epoch || E.plan_epoch[i] != m.plan_epoch)
            changed = true;
    }
    if (!changed) return BFVM_OK;
    drop_graph(E);
    free_tables(E);
    std::vector<bool> on_rows(E.m.size(), false), on_xt(E.m.size(), false);
    if (E.xt_x) {
        // x sweeps with whole-cache-line chunk loads, one launch for every member that can
        std::vector<bfvm::SweepArgs> args;
        std::vector<int> map;
        for (size_t i = 0; i < E.m.size(); ++i) {
            Env& m = *E.m[i];
            bfvm::SweepArgs A;
            size_t lds = 0;
            int tiles = 0;
            if (!joins_batch(E, m, 0, Lax[0]) || !bfvm::batch_xt_args(m, Lax[0], kXtBatchThreads, &A, &lds, &tiles))
                continue;
            on_xt[i] = true;
            for (int t = 0; t < tiles; ++t) {
                map.push_back((int)args.size());
                map.push_back(t);
            }
            args.push_back(A);
            E.xtlds = std::max(E.xtlds, lds);
        }
        if (!args.empty()) {
            E.xtblocks = (int)(map.size() / 2);
            EHIP(&E, hipMalloc((void**)&E.d_xt, sizeof(bfvm::SweepArgs) * args.size()));
            EHIP(&E, hipMalloc((void**)&E.d_xtmap, sizeof(int) * map.size()));
            EHIP(&E, hipMemcpy(E.d_xt, args.data(), sizeof(bfvm::SweepArgs) * args.size(), hipMemcpyHostToDevice));
            EHIP(&E, hipMemcpy(E.d_xtmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        }
    }
    if (E.rows_x) {
        // x sweeps on row tiles, one launch for every member that can
        std::vector<bfvm::RowTile> tabs;
        std::vector<int> map;
        size_t rest = 0;  // LDS beyond the two per-thread carry arrays
        E.rthreads = 0;
        for (size_t i = 0; i < E.m.size(); ++i) {
            Env& m = *E.m[i];
            bfvm::RowTile X;
            int threads = 0, tiles = 0;
            size_t lds = 0;
            if (on_xt[i] || !joins_batch(E, m, 0, Lax[0]) ||
                !bfvm::batch_rows_args(m, Lax[0], &X, &threads, &lds, &tiles))
                continue;
            on_rows[i] = true;
            for (int t = 0; t < tiles; ++t) {
                map.push_back((int)tabs.size());
                map.push_back(t);
            }
            tabs.push_back(X);
            E.rthreads = std::max(E.rthreads, threads);
            rest = std::max(rest, lds - sizeof(double) * 2 * (size_t)threads);
        }
        // every block has the largest member's thread count, and its carry
        // arrays sized for it; beyond the LDS a block may use, no row tiles
        E.rlds = rest + sizeof(double) * 2 * (size_t)E.rthreads;
        if (E.rlds > 150 * 1024) {
            tabs.clear();
            on_rows.assign(E.m.size(), false);
        }
        if (!tabs.empty()) {
            E.rblocks = (int)(map.size() / 2);
            EHIP(&E, hipMalloc((void**)&E.d_rows, sizeof(bfvm::RowTile) * tabs.size()));
            EHIP(&E, hipMalloc((void**)&E.d_rmap, sizeof(int) * map.size()));
            EHIP(&E, hipMemcpy(E.d_rows, tabs.data(), sizeof(bfvm::RowTile) * tabs.size(), hipMemcpyHostToDevice));
            EHIP(&E, hipMemcpy(E.d_rmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        }
    }
    for (int axis = 0; axis < 3; ++axis) {
        std::vector<bfvm::SweepArgs> args;
        std::vector<int> first{0};
        size_t lds = 0;
        for (size_t i = 0; i < E.m.size(); ++i) {
            Env& m = *E.m[i];
            if (axis >= m.lod_dims) continue;  // LOD_2D members have no z sweep
            if (axis == 0 && (on_rows[i] || on_xt[i])) continue;
            if (!joins_batch(E, m, axis, Lax[axis])) {
                E.solo[axis].push_back((int)i);
                continue;
            }
            bfvm::SweepArgs A = bfvm::batch_args(m, axis, Lax[axis]);
            args.push_back(A);
            first.push_back(first.back() + A.tiles * A.planes);
            lds = std::max(lds, bfvm::batch_lds(A));
        }
        E.nbatch[axis] = (int)args.size();
        E.blocks[axis] = first.back();
        E.lds[axis] = lds;
        if (!args.empty()) {
            EHIP(&E, hipMalloc((void**)&E.d_args[axis], sizeof(bfvm::SweepArgs) * args.size()));
            // per block: its member and its block index within the member (one
            // load at the block's start instead of a search over the members)
            std::vector<int> map(2 * (size_t)first.back());
            for (size_t i = 0; i + 1 < first.size(); ++i)
                for (int b = first[i]; b < first[i + 1]; ++b) {
                    map[2 * (size_t)b] = (int)i;
                    map[2 * (size_t)b + 1] = b - first[i];
                }
            EHIP(&E, hipMalloc((void**)&E.d_first[axis], sizeof(int) * map.size()));
            EHIP(&E, hipMemcpy(E.d_args[axis], args.data(), sizeof(bfvm::SweepArgs) * args.size(),
                               hipMemcpyHostToDevice));
            EHIP(&E, hipMemcpy(E.d_first[axis], map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        }
    }
    // each member's (voxel group, substrate) threads start on a block boundary
    std::vector<bfvm::AgentBatch> ag;
    std::vector<int> agent_blocks;
    for (Env* pm : E.m) {
        if (pm->n_agent_voxels == 0) continue;
        bfvm::AgentBatch b;
        b.rho = pm->rho;
        b.gvox = pm->agent_seg_voxel;
        b.gstart = pm->agent_seg_start;
        b.t1 = (const double*)pm->agent_t1;
        b.t2 = (const double*)pm->agent_t2;
        b.S = pm->S;
        b.n_groups = pm->n_agent_voxels;
        b.first = (long long)agent_blocks.size() * bfvm::kAgentBatchBlock;
        const long long items = (long long)pm->n_agent_voxels * pm->S;
        const long long nb = (items + bfvm::kAgentBatchBlock - 1) / bfvm::kAgentBatchBlock;
        if ((long long)agent_blocks.size() + nb > 0x7fffffffll)
            return efail(&E, BFVM_E_ARG, "too many agent voxels in the ensemble");
        agent_blocks.insert(agent_blocks.end(), (size_t)nb, (int)ag.size());
        ag.push_back(b);
    }
    E.n_agent_members = (int)ag.size();
    E.agent_blocks = (int)agent_blocks.size();
    if (!ag.empty()) {
        EHIP(&E, hipMalloc((void**)&E.d_agents, sizeof(bfvm::AgentBatch) * ag.size()));
        EHIP(&E, hipMemcpy(E.d_agents, ag.data(), sizeof(bfvm::AgentBatch) * ag.size(),
                           hipMemcpyHostToDevice));
        EHIP(&E, hipMalloc((void**)&E.d_agent_blocks, sizeof(int) * agent_blocks.size()));
        EHIP(&E, hipMemcpy(E.d_agent_blocks, agent_blocks.data(), sizeof(int) * agent_blocks.size(),
                           hipMemcpyHostToDevice));
    }
    for (size_t i = 0; i < E.m.size(); ++i) {
        E.coef_epoch[i] = E.m[i]->coef_epoch;
        E.agent_epoch[i] = E.m[i]->agent_epoch;
        E.source_epoch[i] = E.m[i]->source_epoch;
        E.plan_epoch[i] = E.m[i]->plan_epoch;
    }
    E.tab_dt = dt;
    for (int a = 0; a < 3; ++a) E.tab_L[a] = Lax[a];
    E.tab_rows_x = E.rows_x;
    E.tab_xt_x = E.xt_x;
    E.tables_valid = true;
    return BFVM_OK;
}

// The launches of one step (after prepare), all on the ensemble's stream.
hipError_t enqueue(bfvm_ensemble_t& E, int flags)
{
    hipError_t st = hipSuccess;
    if (flags & BFVM_BULK)
        for (Env* pm : E.m)
            if (pm->have_bulk && (st = on_stream(*pm, E.stream, [&] { return bfvm::launch_bulk(*pm); })))
                return st;
    if ((flags & BFVM_CELLS) && E.agent_blocks > 0)
        if ((st = bfvm::launch_agents_batch(E.precision, E.d_agents, E.d_agent_blocks, E.agent_blocks,
                                            E.stream)))
            return st;
    if (flags & BFVM_DIFFUSION) {
        for (int axis = 0; axis < 3; ++axis) {
            for (Env* pm : E.m)
                if (pm->n_dir > 0 && axis < pm->lod_dims &&
                    (st = on_stream(*pm, E.stream, [&] { return bfvm::launch_dirichlet(*pm); })))
                    return st;
            if (axis == 0 && E.xtblocks > 0 &&
                (st = bfvm::launch_xt_batch(E.tab_L[0], E.spread, E.d_xt, E.d_xtmap, E.xtblocks, kXtBatchThreads,
                                            E.xtlds, E.stream)))
                return st;
            if (axis == 0 && E.rblocks > 0 &&
                (st = bfvm::launch_rows_batch(E.precision, E.tab_L[0], E.spread, E.d_rows, E.d_rmap, E.rblocks,
                                              E.rthreads, E.rlds, E.stream)))
                return st;
            if (E.nbatch[axis] > 0 &&
                (st = bfvm::launch_sweep_batch(E.precision, E.tab_L[axis], E.spread, E.d_args[axis], E.d_first[axis],
                                               E.nbatch[axis], E.blocks[axis], E.lds[axis], E.stream,
                                               axis > 0 && E.nt)))
                return st;
            for (int i : E.solo[axis])
                if ((st = on_stream(*E.m[i], E.stream, [&] { return bfvm::launch_sweep(*E.m[i], axis); })))
                    return st;
        }
    }
    return hipSuccess;
}

// Captures `steps` steps' launches into a graph and instantiates it.
int capture_steps(bfvm_ensemble_t& E, int flags, int steps, hipGraph_t* graph, hipGraphExec_t* exec)
{
    EHIP(&E, hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
    hipError_t st = hipSuccess;
    for (int n = 0; n < steps && st == hipSuccess; ++n) st = enqueue(E, flags);
    hipGraph_t g = nullptr;
    hipError_t st2 = hipStreamEndCapture(E.stream, &g);
    if (st != hipSuccess || st2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return efail(&E, BFVM_E_DEVICE, std::string("graph capture: ") +
                                            bfvm::hip_error_text(st != hipSuccess ? st : st2));
    }
    *graph = g;
    EHIP(&E, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    // stage the graph's launch resources now rather than at its first replay
    EHIP(&E, hipGraphUpload(*exec, E.stream));
    return BFVM_OK;
}

// The one-step graph, and the graph_steps-step one when a run holds that many.
int ensure_graph(bfvm_ensemble_t& E, double dt, int flags, int steps)
{
    if (!(E.exec && E.graph_dt == dt && E.graph_flags == flags)) {
        drop_graph(E);
        int rc = capture_steps(E, flags, 1, &E.graph, &E.exec);
        if (rc) return rc;
        E.graph_dt = dt;
        E.graph_flags = flags;
    }
    if (E.graph_steps > 1 && steps >= E.graph_steps && !E.exec_multi)
        return capture_steps(E, flags, E.graph_steps, &E.graph_multi, &E.exec_multi);
    return BFVM_OK;
}

// Everything before the launches of `steps` steps: members' own work
// drained, coefficients and tables for dt, the graph if asked for.
int ready(bfvm_ensemble_t& E, double dt, int flags, int use_graph, int steps)
{
    for (Env* pm : E.m) EHIP(&E, hipStreamSynchronize(pm->stream));
    int rc = prepare(E, dt, flags);
    if (rc == BFVM_OK && use_graph) rc = ensure_graph(E, dt, flags, steps);
    return rc;
}

// Graph runs: whole graph_steps-step graphs while they fit, then one-step ones
// (the same launches in the same order as the eager steps).
int launch_steps(bfvm_ensemble_t& E, int steps, int flags, int use_graph)
{
    int n = 0;
    if (use_graph && E.exec_multi)
        for (; n + E.graph_steps <= steps; n += E.graph_steps) EHIP(&E, hipGraphLaunch(E.exec_multi, E.stream));
    for (; n < steps; ++n) {
        if (use_graph) EHIP(&E, hipGraphLaunch(E.exec, E.stream));
        else EHIP(&E, enqueue(E, flags));
    }
    return BFVM_OK;
}

}  // namespace

extern "C" {

int bfvm_ensemble_create(int count, const bfvm_handle* members, bfvm_ensemble* out)
{
    bfvm::RandGuard rand_guard;
    if (!out) return BFVM_E_ARG;
    *out = nullptr;
    if (count < 1 || !members) return BFVM_E_ARG;
    for (int i = 0; i < count; ++i) {
        if (!members[i] || members[i]->ensemble_stream) return BFVM_E_ARG;  // one ensemble per handle
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) return BFVM_E_ARG;
    }
    const int prec = members[0]->precision;
    for (int i = 1; i < count; ++i)
        if (members[i]->precision != prec) return BFVM_E_ARG;
    auto* E = new bfvm_ensemble_t;
    E->precision = prec;
    for (int i = 0; i < count; ++i) E->m.push_back(members[i]);
    E->coef_epoch.assign(count, ~0u);
    E->agent_epoch.assign(count, ~0u);
    E->source_epoch.assign(count, ~0u);
    E->plan_epoch.assign(count, ~0u);
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        delete E;
        return BFVM_E_DEVICE;
    }
    for (Env* pm : E->m) pm->ensemble_stream = E->stream;
    *out = E;
    return BFVM_OK;
}

int bfvm_ensemble_destroy(bfvm_ensemble e)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_OK;
    (void)hipStreamSynchronize(e->stream);
    for (Env* pm : e->m) pm->ensemble_stream = nullptr;
    drop_graph(*e);
    free_tables(*e);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return BFVM_OK;
}

int bfvm_ensemble_run(bfvm_ensemble e, double dt, int steps, int flags, int use_graph)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    if (steps < 0) return efail(e, BFVM_E_ARG, "steps < 0");
    int rc = ready(*e, dt, flags, use_graph, steps);
    return rc ? rc : launch_steps(*e, steps, flags, use_graph);
}

int bfvm_ensemble_sync(bfvm_ensemble e)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    EHIP(e, hipStreamSynchronize(e->stream));
    return BFVM_OK;
}

int bfvm_ensemble_timed_run(bfvm_ensemble e, double dt, int steps, int flags, int use_graph,
                            double* ms)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    if (steps < 0) return efail(e, BFVM_E_ARG, "steps < 0");
    int rc = ready(*e, dt, flags, use_graph, steps);
    if (rc) return rc;
    hipEvent_t t0, t1;
    EHIP(e, hipEventCreate(&t0));
    EHIP(e, hipEventCreate(&t1));
    EHIP(e, hipStreamSynchronize(e->stream));
    EHIP(e, hipEventRecord(t0, e->stream));
    rc = launch_steps(*e, steps, flags, use_graph);
    EHIP(e, hipEventRecord(t1, e->stream));
    EHIP(e, hipEventSynchronize(t1));
    float f = 0;
    EHIP(e, hipEventElapsedTime(&f, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (ms) *ms = f;
    return rc;
}

int bfvm_ensemble_autotune(bfvm_ensemble e, double dt)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    // time the three batched sweeps for each segment length on scratch copies
    for (Env* pm : e->m) EHIP(e, hipStreamSynchronize(pm->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    std::vector<void*> keep, scratch(e->m.size(), nullptr);
    hipError_t st = hipSuccess;
    for (size_t i = 0; i < e->m.size() && st == hipSuccess; ++i) {
        Env& m = *e->m[i];
        st = hipMalloc(&scratch[i], m.nvs * m.elem);
        if (st == hipSuccess)
            st = hipMemcpyAsync(scratch[i], m.rho, m.nvs * m.elem, hipMemcpyDeviceToDevice, e->stream);
    }
    if (st != hipSuccess) {
        (void)hipStreamSynchronize(e->stream);
        for (void* p : scratch)
            if (p) (void)hipFree(p);
        return efail(e, BFVM_E_DEVICE, std::string("autotune scratch: ") + bfvm::hip_error_text(st));
    }
    for (size_t i = 0; i < e->m.size(); ++i) {
        keep.push_back(e->m[i]->rho);
        e->m[i]->rho = scratch[i];
        ++e->m[i]->coef_epoch;  // launch tables point at the scratch stores
    }
    int best_L = e->L, best_spread = e->spread, best_rows = e->rows_x, best_nt = e->nt, best_xt = e->xt_x;
    double best = 1e300;
    int rc = BFVM_OK;
    std::vector<int> lengths{e->L};
    if (e->tune_segments) lengths = {0, 4, 8, 16, 32};  // 0: the members' own lengths
    for (int L : lengths) {
        for (int spread : {0, 64, -1}) {
            for (int xrow : {0, 1, 2}) {  // x: column tiles, row tiles, whole-cache-line chunks (+ row tiles)
                for (int nt : {0, 1}) {
                    e->L = L;
                    e->spread = spread;
                    e->rows_x = xrow > 0;
                    e->xt_x = xrow == 2;
                    e->nt = nt;
                    double ms = 0;
                    if ((rc = bfvm_ensemble_timed_run(e, dt, 1, BFVM_DIFFUSION, 0, &ms))) break;
                    if ((rc = bfvm_ensemble_timed_run(e, dt, 3, BFVM_DIFFUSION, 0, &ms))) break;
                    if (ms < best) {
                        best = ms;
                        best_L = L;
                        best_spread = spread;
                        best_rows = e->rows_x;
                        best_xt = e->xt_x;
                        best_nt = nt;
                    }
                }
                if (rc) break;
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipStreamSynchronize(e->stream);
    for (size_t i = 0; i < e->m.size(); ++i) {
        e->m[i]->rho = keep[i];
        ++e->m[i]->coef_epoch;
        (void)hipFree(scratch[i]);
    }
    e->L = best_L;
    e->spread = best_spread;
    e->rows_x = best_rows;
    e->xt_x = best_xt;
    e->nt = best_nt;
    e->tables_valid = false;
    drop_graph(*e);
    return rc;
}

int bfvm_ensemble_set_graph_steps(bfvm_ensemble e, int steps_per_graph)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    if (steps_per_graph < 1) return efail(e, BFVM_E_ARG, "steps per graph must be at least 1");
    e->graph_steps = steps_per_graph;
    drop_graph(*e);
    return BFVM_OK;
}

int bfvm_ensemble_get_graph_steps(bfvm_ensemble e, int* steps_per_graph)
{
    bfvm::RandGuard rand_guard;
    if (!e || !steps_per_graph) return BFVM_E_ARG;
    *steps_per_graph = e->graph_steps;
    return BFVM_OK;
}

int bfvm_ensemble_set_autotune_segments(bfvm_ensemble e, int on)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    e->tune_segments = on != 0;
    return BFVM_OK;
}

int bfvm_ensemble_get_segment(bfvm_ensemble e, int* nodes_per_segment)
{
    bfvm::RandGuard rand_guard;
    if (!e || !nodes_per_segment) return BFVM_E_ARG;
    *nodes_per_segment = e->L;
    return BFVM_OK;
}

int bfvm_ensemble_get_axis_segment(bfvm_ensemble e, int axis, int* nodes_per_segment)
{
    bfvm::RandGuard rand_guard;
    if (!e || !nodes_per_segment || axis < 0 || axis > 2) return BFVM_E_ARG;
    *nodes_per_segment = axis_segment(*e, axis);
    return BFVM_OK;
}

int bfvm_ensemble_set_segment(bfvm_ensemble e, int nodes_per_segment)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    if (nodes_per_segment != 0 && nodes_per_segment != 4 && nodes_per_segment != 8 && nodes_per_segment != 16 &&
        nodes_per_segment != 32)
        return efail(e, BFVM_E_ARG, "nodes per segment must be 0 (the members' own), 4, 8, 16 or 32");
    e->L = nodes_per_segment;
    e->tables_valid = false;
    drop_graph(*e);
    return BFVM_OK;
}

int bfvm_ensemble_set_x_chunks(bfvm_ensemble e, int on)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    e->xt_x = on != 0;
    e->tables_valid = false;
    drop_graph(*e);
    return BFVM_OK;
}

int bfvm_ensemble_get_x_chunks(bfvm_ensemble e, int* on)
{
    bfvm::RandGuard rand_guard;
    if (!e || !on) return BFVM_E_ARG;
    *on = e->xt_x;
    return BFVM_OK;
}

int bfvm_ensemble_set_row_tiles(bfvm_ensemble e, int on)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    e->rows_x = on ? 1 : 0;
    e->tables_valid = false;
    drop_graph(*e);
    return BFVM_OK;
}

int bfvm_ensemble_get_row_tiles(bfvm_ensemble e, int* on)
{
    bfvm::RandGuard rand_guard;
    if (!e || !on) return BFVM_E_ARG;
    *on = e->rows_x;
    return BFVM_OK;
}

int bfvm_ensemble_set_cache_policy(bfvm_ensemble e, int non_temporal)
{
    bfvm::RandGuard rand_guard;
    if (!e) return BFVM_E_ARG;
    e->nt = non_temporal ? 1 : 0;
    drop_graph(*e);
    return BFVM_OK;
}

int bfvm_ensemble_get_cache_policy(bfvm_ensemble e, int* non_temporal)
{
    bfvm::RandGuard rand_guard;
    if (!e || !non_temporal) return BFVM_E_ARG;
    *non_temporal = e->nt;
    return BFVM_OK;
}

const char* bfvm_ensemble_last_error(bfvm_ensemble e)
{
    bfvm::RandGuard rand_guard;
    return e ? e->err.c_str() : "null ensemble";
}

}  // extern "C"
