// C ABI of biofvm_amd: handle lifetime, host<->device transfers, the FP64
// Thomas factorisation, agent grouping and the per-step launch sequence.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "bfvm_internal.h"

using bfvm::Env;

namespace {

thread_local std::string g_thread_err;

int fail(Env* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    g_thread_err = msg;
    return code;
}

int hip_fail(Env* e, hipError_t st, const char* what)
{
    std::string m = std::string(what) + ": " + bfvm::hip_error_text(st);
    return fail(e, BFVM_E_DEVICE, m);
}

#define HIP_OR_FAIL(e, call)                               \
    do {                                                   \
        hipError_t st_ = (call);                           \
        if (st_ != hipSuccess) return hip_fail(e, st_, #call); \
    } while (0)

void dfree(void*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <typename P>
void dfree_t(P*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// Host values -> device buffer in the handle's precision.
hipError_t upload_values(Env& e, void* dst, const double* src, size_t n)
{
    if (e.precision == BFVM_FP64) return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice);
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    return hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice);
}

hipError_t alloc_values(Env& e, void** p, size_t n)
{
    return hipMalloc(p, std::max<size_t>(n, 1) * e.elem);
}

// FP64 tables (coefficients, source terms) are kept in FP64 whatever the
// store precision.
hipError_t upload_f64(void* dst, const double* src, size_t n)
{
    return hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice);
}

hipError_t alloc_f64(void** p, size_t n)
{
    return hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double));
}

#pragma clang fp contract(off)
// Thomas factorisation of one axis in FP64, in the reference's operation
// order (BioFVM_solvers.cpp:104-142).  denom/c are (n, S), element [i*S+s].
void thomas_axis(int n, int S, double dx, double dt, const std::vector<double>& D,
                 const std::vector<double>& lambda, int lod_dims, std::vector<double>& denom,
                 std::vector<double>& c, std::vector<double>& c1v)
{
    denom.assign((size_t)n * S, 0.0);
    c.assign((size_t)n * S, 0.0);
    c1v.assign(S, 0.0);
    for (int s = 0; s < S; ++s) {
        double c1 = D[s];
        c1 *= dt;
        c1 /= dx;
        c1 /= dx;
        double c1a = c1 * -1.0;
        const double c2 = bfvm::decay_split(lambda[s], dt, lod_dims);
        double c3 = 1.0;
        c3 += c1;
        c3 += c1;
        c3 += c2;
        double c3a = 1.0;
        c3a += c1;
        c3a += c2;
        c1v[s] = c1;
        for (int i = 0; i < n; ++i) {
            c[(size_t)i * S + s] = c1a;
            denom[(size_t)i * S + s] = c3;
        }
        denom[s] = c3a;
        denom[(size_t)(n - 1) * S + s] = c3a;
        if (n == 1) {
            denom[s] = 1.0;
            denom[s] += c2;
        }
        c[s] /= denom[s];
        for (int i = 1; i <= n - 1; ++i) {
            double prod = c1 * c[(size_t)(i - 1) * S + s];
            denom[(size_t)i * S + s] += prod;
            c[(size_t)i * S + s] /= denom[(size_t)i * S + s];
        }
    }
}

// temp1 = ((0 + S)*T)*c, temp2 = (1 + c*S) + c*U with c = dt*V/V_voxel
// (BioFVM_basic_agent.cpp:122-147).
void agent_constants(int S, double dt, double voxel_volume, double volume, const double* sec,
                     const double* sat, const double* upt, double* t1, double* t2)
{
    double c = dt * volume;
    c = c / voxel_volume;
    for (int s = 0; s < S; ++s) {
        double a = 0.0;
        a += sec[s];
        a *= sat[s];
        a *= c;
        double b = 1.0;
        double cs = c * sec[s];
        b += cs;
        double cu = c * upt[s];
        b += cu;
        t1[s] = a;
        t2[s] = b;
    }
}
#pragma clang fp contract(on)

// Constant-coefficient solve A_inf^-1 x (see AxisCoef), in FP64.
void solve_const(int n, double r, double a, const std::vector<double>& x, std::vector<double>& z)
{
    std::vector<double> y(n);
    double prev = 0.0;
    for (int i = 0; i < n; ++i) y[i] = prev = a * prev + x[i] * r;
    z.assign(n, 0.0);
    double next = 0.0;
    for (int i = n - 1; i >= 0; --i) z[i] = next = y[i] + (i < n - 1 ? a * next : 0.0);
}

// Per-axis tables of the segmented sweep (see AxisCoef).
int build_woodbury(Env& e, bfvm::AxisCoef& ax, int n, double dt)
{
    const int S = e.S;
    std::vector<double> wb((size_t)S * 8, 0.0), g0all((size_t)n * S), g1all((size_t)n * S);
    int kg0 = 1, kg1 = 1;
    for (int s = 0; s < S; ++s) {
        double c1 = e.D[s];
        c1 *= dt;
        c1 /= e.dx;
        c1 /= e.dx;
        const double c2 = bfvm::decay_split(e.lambda[s], dt, e.lod_dims);
        const double c3 = 1.0 + c1 + c1 + c2, c3a = 1.0 + c1 + c2;
        // fixed point of the reference's recurrence denom_i = c3 + c1*c_{i-1},
        // c_i = -c1/denom_i, iterated from its first row
        double d = c3a;
        for (int it = 0; it < 100000; ++it) {
            double dn = c3 + c1 * (-c1 / d);
            if (dn == d) break;
            d = dn;
        }
        const double r = 1.0 / d, a = c1 / d;
        const double delta0 = d - c3a, delta1 = (d + c1 * (c1 / d)) - c3a;
        std::vector<double> unit(n, 0.0), g0, g1;
        unit[0] = delta0;
        solve_const(n, r, a, unit, g0);
        unit.assign(n, 0.0);
        unit[n - 1] = delta1;
        solve_const(n, r, a, unit, g1);
        const double m00 = g0[0], m01 = g1[0], m10 = g0[n - 1], m11 = g1[n - 1];
        const double det = (1.0 - m00) * (1.0 - m11) - m01 * m10;
        double* w = &wb[(size_t)s * 8];
        w[0] = r;
        w[1] = a;
        w[2] = (1.0 - m11) / det;
        w[3] = m01 / det;
        w[4] = m10 / det;
        w[5] = (1.0 - m00) / det;
        double p0 = 0, p1 = 0;
        for (int i = 0; i < n; ++i) {
            p0 = std::max(p0, std::fabs(g0[i]));
            p1 = std::max(p1, std::fabs(g1[i]));
        }
        for (int i = 0; i < n; ++i) {
            g0all[(size_t)i * S + s] = g0[i];
            g1all[(size_t)i * S + s] = g1[n - 1 - i];
            if (std::fabs(g0[i]) > 1e-18 * p0) kg0 = std::max(kg0, i + 1);
            if (std::fabs(g1[n - 1 - i]) > 1e-18 * p1) kg1 = std::max(kg1, i + 1);
        }
    }
    dfree(ax.wb);
    dfree(ax.g0);
    dfree(ax.g1);
    HIP_OR_FAIL(&e, alloc_f64(&ax.wb, wb.size()));
    HIP_OR_FAIL(&e, alloc_f64(&ax.g0, (size_t)kg0 * S));
    HIP_OR_FAIL(&e, alloc_f64(&ax.g1, (size_t)kg1 * S));
    HIP_OR_FAIL(&e, upload_f64(ax.wb, wb.data(), wb.size()));
    HIP_OR_FAIL(&e, upload_f64(ax.g0, g0all.data(), (size_t)kg0 * S));
    HIP_OR_FAIL(&e, upload_f64(ax.g1, g1all.data(), (size_t)kg1 * S));
    ax.kg0 = kg0;
    ax.kg1 = kg1;
    return BFVM_OK;
}

// Waits for everything enqueued on the handle: its own stream and, for a
// member of an ensemble, the ensemble's stream.
hipError_t quiesce(Env& e)
{
    if (e.ensemble_stream) {
        hipError_t st = hipStreamSynchronize(e.ensemble_stream);
        if (st != hipSuccess) return st;
    }
    return hipStreamSynchronize(e.stream);
}

int ensure_coefficients_impl(Env& e, double dt)
{
    if (e.coef_valid && e.coef_dt == dt) return BFVM_OK;
    // kernels in flight may still read the tables about to be replaced
    HIP_OR_FAIL(&e, quiesce(e));
    const int n_axis[3] = {e.nx, e.ny, e.nz};
    const int S = e.S;
    std::vector<double> denom, c, c1v;
    for (int a = 0; a < 3; ++a) {
        thomas_axis(n_axis[a], S, e.dx, dt, e.D, e.lambda, e.lod_dims, denom, c, c1v);
        bfvm::AxisCoef& ax = e.ax[a];
        const size_t m = (size_t)n_axis[a] * S;
        HIP_OR_FAIL(&e, upload_f64(ax.denom, denom.data(), m));
        HIP_OR_FAIL(&e, upload_f64(ax.c, c.data(), m));
        int rc = build_woodbury(e, ax, n_axis[a], dt);
        if (rc) return rc;
    }
    HIP_OR_FAIL(&e, upload_f64(e.c1, c1v.data(), S));
    e.coef_valid = true;
    e.coef_dt = dt;
    ++e.coef_epoch;
    ++e.coef_builds;
    return BFVM_OK;
}


// Launch wrapper: optional per-kernel HIP events for bfvm_timed_run.
template <typename F>
int run_kernel(Env& e, int kind, F&& launch)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (e.kernel_timing) {
        while (e.ev_pool.size() < e.ev_next + 2) {
            hipEvent_t ev;
            HIP_OR_FAIL(&e, hipEventCreate(&ev));
            e.ev_pool.push_back(ev);
        }
        a = e.ev_pool[e.ev_next++];
        b = e.ev_pool[e.ev_next++];
        HIP_OR_FAIL(&e, hipEventRecord(a, e.stream));
    }
    hipError_t st = launch();
    if (st != hipSuccess) return hip_fail(&e, st, "kernel launch");
    if (e.kernel_timing) {
        HIP_OR_FAIL(&e, hipEventRecord(b, e.stream));
        e.ev_marks.push_back({kind, {a, b}});
    }
    return BFVM_OK;
}

int step_impl(Env& e, double dt, int flags)
{
    int rc;
    const bool bulk = (flags & BFVM_BULK) && e.have_bulk;
    const bool cells = (flags & BFVM_CELLS) && e.n_agents > 0;
    const bool diffusion = (flags & BFVM_DIFFUSION) != 0;
    if (diffusion) {
        if (!e.uniform)
            return fail(&e, BFVM_E_MESH,
                        "This algorithm is written for uniform Cartesian meshes (dx == dy == dz).");
        if (e.D.empty()) return fail(&e, BFVM_E_STATE, "coefficients not set (bfvm_set_coefficients)");
        if ((rc = ensure_coefficients_impl(e, dt))) return rc;
    }
    if (bulk) HIP_OR_FAIL(&e, bfvm::ensure_bulk(e, dt));
    const int chunk = e.xy_chunk;
    const bool chunked = diffusion && chunk > 0 && chunk < e.nz && bfvm::sweep_chunkable(e, 0) &&
                         bfvm::sweep_chunkable(e, 1);
    // the bulk update and the first two Dirichlet overwrites inside the x sweep
    // (the agent kernel applies the bulk update of the voxels with agents)
    const bool fused = diffusion && e.fused_sources && !chunked && (bulk || e.n_dir > 0) &&
                       bfvm::sweep_x_fusable(e);
    if (bulk && !fused) {
        if ((rc = run_kernel(e, BFVM_K_BULK, [&] { return bfvm::launch_bulk(e); }))) return rc;
    }
    if (cells) {
        if ((rc = run_kernel(e, BFVM_K_AGENTS, [&] { return bfvm::launch_agents(e, bulk && fused); }))) return rc;
    }
    if (diffusion) {
        auto dirichlet = [&] {
            return e.n_dir > 0 ? run_kernel(e, BFVM_K_DIRICHLET, [&] { return bfvm::launch_dirichlet(e); })
                               : BFVM_OK;
        };
        if (fused) {
            const bool dir_after = e.lod_dims >= 2;
            if ((rc = run_kernel(e, BFVM_K_SWEEP_X,
                                 [&] { return bfvm::launch_sweep_x_fused(e, bulk, cells, dir_after); })))
                return rc;
            for (int axis = 1; axis < e.lod_dims; ++axis) {
                if (axis > 1 && (rc = dirichlet())) return rc;
                if ((rc = run_kernel(e, BFVM_K_SWEEP_X + axis, [&] { return bfvm::launch_sweep(e, axis); })))
                    return rc;
            }
            return BFVM_OK;
        }
        int axis0 = 0;
        if (chunked) {
            // x then y per chunk of z planes.  Every Dirichlet launch writes all
            // nodes; the overwrite is idempotent, so each plane still sees it
            // before its x, y and z sweeps and never in the middle of one.
            for (int k0 = 0; k0 < e.nz; k0 += chunk) {
                const int k1 = std::min(e.nz, k0 + chunk);
                for (int axis = 0; axis < 2; ++axis) {
                    if ((rc = dirichlet())) return rc;
                    if ((rc = run_kernel(e, BFVM_K_SWEEP_X + axis,
                                         [&] { return bfvm::launch_sweep_planes(e, axis, k0, k1); })))
                        return rc;
                }
            }
            axis0 = 2;
        }
        for (int axis = axis0; axis < e.lod_dims; ++axis) {
            if ((rc = dirichlet())) return rc;
            if ((rc = run_kernel(e, BFVM_K_SWEEP_X + axis, [&] { return bfvm::launch_sweep(e, axis); })))
                return rc;
        }
    }
    return BFVM_OK;
}

}  // namespace

// A HIP graph of `steps` consecutive steps, and what its launches have baked
// in: the store, the tables and sources (through their epochs), the launch
// shapes, dt and the phases.
struct bfvm::StepGraph {
    hipGraphExec_t exec = nullptr;
    int steps = 0, flags = 0, lod_dims = 0, xy_chunk = 0;
    double dt = 0;
    bool fused = false;
    unsigned coef_epoch = 0, plan_epoch = 0, source_epoch = 0, agent_epoch = 0;
    const void *rho = nullptr, *out = nullptr, *bulk_A = nullptr, *bulk_B = nullptr;

    bool matches(const Env& e, double dt_, int flags_, int steps_) const
    {
        return exec && steps == steps_ && flags == flags_ && dt == dt_ && lod_dims == e.lod_dims &&
               xy_chunk == e.xy_chunk && fused == e.fused_sources && coef_epoch == e.coef_epoch &&
               plan_epoch == e.plan_epoch && source_epoch == e.source_epoch && agent_epoch == e.agent_epoch &&
               rho == e.rho && out == e.out && bulk_A == e.bulk_A && bulk_B == e.bulk_B;
    }
};

namespace {

using bfvm::StepGraph;

void free_graph(Env& e)
{
    if (!e.graph) return;
    if (e.graph->exec) (void)hipGraphExecDestroy(e.graph->exec);
    delete e.graph;
    e.graph = nullptr;
}

// Captures `g` steps into the handle's graph.  Every step before it ran
// eagerly with the same arguments, so the tables and bulk terms are built and
// the kernels' attributes set: the capture holds kernel launches only.
int capture_steps(Env& e, double dt, int flags, int g)
{
    free_graph(e);
    HIP_OR_FAIL(&e, hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    int rc = BFVM_OK;
    for (int n = 0; n < g && rc == BFVM_OK; ++n) rc = step_impl(e, dt, flags);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(e.stream, &graph);
    if (rc != BFVM_OK || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != BFVM_OK ? rc : hip_fail(&e, end, "hipStreamEndCapture");
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (st != hipSuccess) return hip_fail(&e, st, "hipGraphInstantiate");
    // stage the graph's launch resources now rather than at its first replay
    const hipError_t up = hipGraphUpload(exec, e.stream);
    if (up != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        return hip_fail(&e, up, "hipGraphUpload");
    }
    StepGraph* G = new StepGraph;
    G->exec = exec;
    G->steps = g;
    G->flags = flags;
    G->dt = dt;
    G->lod_dims = e.lod_dims;
    G->xy_chunk = e.xy_chunk;
    G->fused = e.fused_sources;
    G->coef_epoch = e.coef_epoch;
    G->plan_epoch = e.plan_epoch;
    G->source_epoch = e.source_epoch;
    G->agent_epoch = e.agent_epoch;
    G->rho = e.rho;
    G->out = e.out;
    G->bulk_A = e.bulk_A;
    G->bulk_B = e.bulk_B;
    e.graph = G;
    return BFVM_OK;
}

// `steps` steps on the handle's stream: the first eagerly, then whole graphs
// of e.graph_steps steps, then the rest eagerly.  The graph replays exactly
// the launches of the eager steps (same kernels, arguments and order), so
// the results are the same bits either way.
int run_steps(Env& e, double dt, int steps, int flags)
{
    int done = 0, rc = BFVM_OK;
    const int g = e.graph_steps;
    if (steps >= 2 && g >= 2 && !e.kernel_timing && steps - 1 >= g) {
        if ((rc = step_impl(e, dt, flags))) return rc;
        done = 1;
        if (!(e.graph && e.graph->matches(e, dt, flags, g)) && (rc = capture_steps(e, dt, flags, g))) return rc;
        for (; steps - done >= g; done += g) HIP_OR_FAIL(&e, hipGraphLaunch(e.graph->exec, e.stream));
    }
    for (; done < steps && rc == BFVM_OK; ++done) rc = step_impl(e, dt, flags);
    return rc;
}

void free_agents(Env& e)
{
    dfree_t(e.agent_seg_voxel);
    dfree_t(e.agent_seg_start);
    dfree_t(e.agent_bits);
    dfree(e.agent_t1);
    dfree(e.agent_t2);
    dfree_t(e.agent_c);
    dfree_t(e.agent_sat);
    e.agent_rates = false;
    e.agent_order.clear();
    e.n_agents = 0;
    e.n_agent_voxels = 0;
}

// c (n) and sat (n*S) in list order, or NULL: kept in grouped order for the
// adjoint's rate gradients.
int set_agents_impl(Env& e, int n, const int* voxel, const double* t1, const double* t2,
                    const double* c = nullptr, const double* sat = nullptr, double dt = 0)
{
    free_agents(e);
    ++e.agent_epoch;
    if (n == 0) return BFVM_OK;
    const int S = e.S;
    for (int a = 0; a < n; ++a)
        if (voxel[a] < 0 || (size_t)voxel[a] >= e.nvox)
            return fail(&e, BFVM_E_ARG, "agent voxel index out of range");
    // Group by voxel, keeping list order inside a voxel (stable sort).
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return voxel[x] < voxel[y]; });
    std::vector<int> gvox, gstart;
    std::vector<double> s1((size_t)n * S), s2((size_t)n * S);
    for (int q = 0; q < n; ++q) {
        int a = order[q];
        if (q == 0 || voxel[a] != voxel[order[q - 1]]) {
            gvox.push_back(voxel[a]);
            gstart.push_back(q);
        }
        std::memcpy(&s1[(size_t)q * S], t1 + (size_t)a * S, sizeof(double) * S);
        std::memcpy(&s2[(size_t)q * S], t2 + (size_t)a * S, sizeof(double) * S);
    }
    gstart.push_back(n);
    int ng = (int)gvox.size();
    HIP_OR_FAIL(&e, hipMalloc((void**)&e.agent_seg_voxel, sizeof(int) * ng));
    HIP_OR_FAIL(&e, hipMalloc((void**)&e.agent_seg_start, sizeof(int) * (ng + 1)));
    HIP_OR_FAIL(&e, alloc_f64(&e.agent_t1, (size_t)n * S));
    HIP_OR_FAIL(&e, alloc_f64(&e.agent_t2, (size_t)n * S));
    HIP_OR_FAIL(&e, hipMemcpy(e.agent_seg_voxel, gvox.data(), sizeof(int) * ng, hipMemcpyHostToDevice));
    HIP_OR_FAIL(&e, hipMemcpy(e.agent_seg_start, gstart.data(), sizeof(int) * (ng + 1),
                              hipMemcpyHostToDevice));
    HIP_OR_FAIL(&e, upload_f64(e.agent_t1, s1.data(), (size_t)n * S));
    HIP_OR_FAIL(&e, upload_f64(e.agent_t2, s2.data(), (size_t)n * S));
    e.agent_order = order;
    if (c && sat) {
        std::vector<double> cg(n), sg((size_t)n * S);
        for (int q = 0; q < n; ++q) {
            cg[q] = c[order[q]];
            std::memcpy(&sg[(size_t)q * S], sat + (size_t)order[q] * S, sizeof(double) * S);
        }
        HIP_OR_FAIL(&e, hipMalloc((void**)&e.agent_c, sizeof(double) * n));
        HIP_OR_FAIL(&e, hipMalloc((void**)&e.agent_sat, sizeof(double) * n * S));
        HIP_OR_FAIL(&e, hipMemcpy(e.agent_c, cg.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIP_OR_FAIL(&e, hipMemcpy(e.agent_sat, sg.data(), sizeof(double) * n * S, hipMemcpyHostToDevice));
        e.agent_rates = true;
        e.agent_dt = dt;
    }
    e.n_agents = n;
    e.n_agent_voxels = ng;
    HIP_OR_FAIL(&e, bfvm::build_agent_bits(e));
    return BFVM_OK;
}

}  // namespace

namespace bfvm {
int ensure_coefficients(Env& e, double dt) { return ensure_coefficients_impl(e, dt); }

hipError_t ensure_bulk(Env& e, double dt)
{
    if (!e.have_bulk || (e.bulk_valid && e.bulk_dt == dt)) return hipSuccess;
    hipError_t st = launch_bulk_prepare(e, dt);
    if (st == hipSuccess) {
        e.bulk_valid = true;
        e.bulk_dt = dt;
    }
    return st;
}
}  // namespace bfvm

extern "C" {

int bfvm_create_mesh(int nx, int ny, int nz, int n_substrates, double dx, double dy, double dz,
                     int precision, bfvm_handle* out)
{
    bfvm::RandGuard rand_guard;
    if (!out) return fail(nullptr, BFVM_E_ARG, "out is NULL");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || n_substrates < 1)
        return fail(nullptr, BFVM_E_ARG, "grid sizes and substrate count must be >= 1");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(nullptr, BFVM_E_ARG, "voxel edges must be > 0");
    if (precision != BFVM_FP64 && precision != BFVM_FP32)
        return fail(nullptr, BFVM_E_ARG, "precision must be BFVM_FP64 or BFVM_FP32");
    size_t nvox = (size_t)nx * ny * nz;
    if (nvox * (size_t)n_substrates >= ((size_t)1 << 31))
        return fail(nullptr, BFVM_E_ARG, "nx*ny*nz*S must stay below 2^31");
    bfvm_env* e = new bfvm_env;
    e->nx = nx; e->ny = ny; e->nz = nz; e->S = n_substrates;
    e->dx = dx; e->dy = dy; e->dz = dz;
    const double tol = 1e-16;  // BioFVM_mesh.cpp:811-815
    e->uniform = !(std::fabs(dx - dy) > tol || std::fabs(dy - dz) > tol || std::fabs(dx - dz) > tol);
    e->precision = precision;
    e->elem = precision == BFVM_FP64 ? 8 : 4;
    e->nvox = nvox;
    e->nvs = nvox * n_substrates;
    e->D.clear();
    e->lambda.clear();
    hipError_t st;
    if ((st = hipGetDevice(&e->device)) != hipSuccess ||
        (st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess ||
        (st = hipMalloc(&e->rho, e->nvs * e->elem)) != hipSuccess ||
        (st = hipMemset(e->rho, 0, e->nvs * e->elem)) != hipSuccess ||
        (st = hipMalloc(&e->c1, e->S * sizeof(double))) != hipSuccess ||
        (st = hipMalloc((void**)&e->dir_mask, e->S)) != hipSuccess) {
        int rc = hip_fail(nullptr, st, "bfvm_create");
        bfvm_destroy(e);
        return rc;
    }
    const int n_axis[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        size_t m = (size_t)n_axis[a] * e->S;
        e->ax[a].n = n_axis[a];
        if ((st = alloc_f64(&e->ax[a].denom, m)) != hipSuccess ||
            (st = alloc_f64(&e->ax[a].c, m)) != hipSuccess) {
            int rc = hip_fail(nullptr, st, "bfvm_create");
            bfvm_destroy(e);
            return rc;
        }
    }
    *out = e;
    return BFVM_OK;
}

int bfvm_create(int nx, int ny, int nz, int n_substrates, double dx, int precision, bfvm_handle* out)
{
    bfvm::RandGuard rand_guard;
    return bfvm_create_mesh(nx, ny, nz, n_substrates, dx, dx, dx, precision, out);
}

int bfvm_destroy(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_OK;
    if (h->ensemble_stream)
        return fail(h, BFVM_E_STATE, "the handle belongs to an ensemble: destroy the ensemble first");
    bfvm_env* e = h;
    bfvm::free_tape(*e);
    bfvm::free_writer(*e);
    free_graph(*e);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    dfree(e->rho);
    dfree(e->c1);
    for (auto& a : e->ax) {
        dfree(a.denom);
        dfree(a.c);
        dfree(a.wb);
        dfree(a.g0);
        dfree(a.g1);
    }
    dfree_t(e->dir_idx);
    dfree(e->dir_val);
    dfree_t(e->dir_mask);
    dfree_t(e->dir_bits);
    dfree_t(e->dir_rank);
    dfree_t(e->agent_bits);
    dfree_t(e->bulk_supply);
    dfree_t(e->bulk_target);
    dfree_t(e->bulk_uptake);
    dfree(e->bulk_A);
    dfree(e->bulk_B);
    dfree(e->grad);
    dfree(e->red);
    free_agents(*e);
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return BFVM_OK;
}

int bfvm_set_coefficients(bfvm_handle h, const double* D, const double* lambda)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (!D || !lambda) return fail(h, BFVM_E_ARG, "D and lambda must be non-NULL");
    // the same values as before keep the tables (a caller that sets them
    // before every step rebuilds nothing)
    const size_t bytes = sizeof(double) * h->S;
    if (h->D.size() == (size_t)h->S && std::memcmp(h->D.data(), D, bytes) == 0 &&
        std::memcmp(h->lambda.data(), lambda, bytes) == 0)
        return BFVM_OK;
    h->D.assign(D, D + h->S);
    h->lambda.assign(lambda, lambda + h->S);
    h->coef_valid = false;
    return BFVM_OK;
}

int bfvm_upload_density(bfvm_handle h, const double* host)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (!host) return fail(h, BFVM_E_ARG, "host is NULL");
    HIP_OR_FAIL(h, quiesce(*h));
    HIP_OR_FAIL(h, upload_values(*h, h->rho, host, h->nvs));
    return BFVM_OK;
}

int bfvm_download_density(bfvm_handle h, double* host)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (!host) return fail(h, BFVM_E_ARG, "host is NULL");
    HIP_OR_FAIL(h, quiesce(*h));
    if (h->precision == BFVM_FP64) {
        HIP_OR_FAIL(h, hipMemcpy(host, h->rho, h->nvs * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(h->nvs);
        HIP_OR_FAIL(h, hipMemcpy(tmp.data(), h->rho, h->nvs * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < h->nvs; ++i) host[i] = tmp[i];
    }
    return BFVM_OK;
}

int bfvm_set_dirichlet(bfvm_handle h, int n, const int* voxel_index, const double* values,
                       const unsigned char* mask)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (n < 0) return fail(h, BFVM_E_ARG, "n < 0");
    HIP_OR_FAIL(h, quiesce(*h));
    ++h->source_epoch;
    dfree_t(h->dir_idx);
    dfree(h->dir_val);
    dfree_t(h->dir_bits);
    dfree_t(h->dir_rank);
    h->n_dir = 0;
    if (n == 0) return BFVM_OK;
    if (!voxel_index || !values) return fail(h, BFVM_E_ARG, "voxel_index and values must be non-NULL");
    for (int d = 0; d < n; ++d)
        if (voxel_index[d] < 0 || (size_t)voxel_index[d] >= h->nvox)
            return fail(h, BFVM_E_ARG, "Dirichlet voxel index out of range");
    std::vector<unsigned char> m(h->S, 1);
    if (mask)
        for (int s = 0; s < h->S; ++s) m[s] = mask[s] ? 1 : 0;
    // One node per voxel, sorted by voxel: where a voxel is listed more than
    // once its last entry wins, as the reference's in-order overwrite leaves it.
    const int S = h->S;
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return voxel_index[a] < voxel_index[b]; });
    std::vector<int> idx;
    std::vector<double> val;
    for (int q = 0; q < n; ++q) {
        if (q + 1 < n && voxel_index[ord[q + 1]] == voxel_index[ord[q]]) continue;  // a later entry follows
        idx.push_back(voxel_index[ord[q]]);
        val.insert(val.end(), values + (size_t)ord[q] * S, values + (size_t)ord[q] * S + S);
    }
    const int u = (int)idx.size();
    const size_t words = h->nvox / 32 + 2;
    std::vector<unsigned> bits(words, 0u);
    std::vector<int> rank(words, 0);
    for (int v : idx) bits[(size_t)v / 32] |= 1u << (v % 32);
    for (size_t w = 1; w < words; ++w) rank[w] = rank[w - 1] + __builtin_popcount(bits[w - 1]);
    HIP_OR_FAIL(h, hipMalloc((void**)&h->dir_idx, sizeof(int) * u));
    HIP_OR_FAIL(h, alloc_values(*h, &h->dir_val, (size_t)u * S));
    HIP_OR_FAIL(h, hipMalloc((void**)&h->dir_bits, sizeof(unsigned) * words));
    HIP_OR_FAIL(h, hipMalloc((void**)&h->dir_rank, sizeof(int) * words));
    HIP_OR_FAIL(h, hipMemcpy(h->dir_idx, idx.data(), sizeof(int) * u, hipMemcpyHostToDevice));
    HIP_OR_FAIL(h, upload_values(*h, h->dir_val, val.data(), (size_t)u * S));
    HIP_OR_FAIL(h, hipMemcpy(h->dir_bits, bits.data(), sizeof(unsigned) * words, hipMemcpyHostToDevice));
    HIP_OR_FAIL(h, hipMemcpy(h->dir_rank, rank.data(), sizeof(int) * words, hipMemcpyHostToDevice));
    HIP_OR_FAIL(h, hipMemcpy(h->dir_mask, m.data(), h->S, hipMemcpyHostToDevice));
    h->n_dir = u;
    return BFVM_OK;
}

int bfvm_set_bulk(bfvm_handle h, const double* supply_rate, const double* supply_target,
                  const double* uptake_rate)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    HIP_OR_FAIL(h, quiesce(*h));
    bool none = !supply_rate && !supply_target && !uptake_rate;
    bool all = supply_rate && supply_target && uptake_rate;
    if (!none && !all) return fail(h, BFVM_E_ARG, "give all three bulk fields or none");
    ++h->source_epoch;
    h->have_bulk = false;
    h->bulk_valid = false;
    if (none) {
        dfree_t(h->bulk_supply);
        dfree_t(h->bulk_target);
        dfree_t(h->bulk_uptake);
        dfree(h->bulk_A);
        dfree(h->bulk_B);
        return BFVM_OK;
    }
    size_t bytes = h->nvs * sizeof(double);
    if (!h->bulk_supply) {
        HIP_OR_FAIL(h, hipMalloc((void**)&h->bulk_supply, bytes));
        HIP_OR_FAIL(h, hipMalloc((void**)&h->bulk_target, bytes));
        HIP_OR_FAIL(h, hipMalloc((void**)&h->bulk_uptake, bytes));
        HIP_OR_FAIL(h, alloc_f64(&h->bulk_A, h->nvs));
        HIP_OR_FAIL(h, alloc_f64(&h->bulk_B, h->nvs));
    }
    HIP_OR_FAIL(h, hipMemcpy(h->bulk_supply, supply_rate, bytes, hipMemcpyHostToDevice));
    HIP_OR_FAIL(h, hipMemcpy(h->bulk_target, supply_target, bytes, hipMemcpyHostToDevice));
    HIP_OR_FAIL(h, hipMemcpy(h->bulk_uptake, uptake_rate, bytes, hipMemcpyHostToDevice));
    h->have_bulk = true;
    return BFVM_OK;
}

int bfvm_set_agents(bfvm_handle h, int n, const int* voxel_index, const double* temp1,
                    const double* temp2)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (n < 0) return fail(h, BFVM_E_ARG, "n < 0");
    if (n > 0 && (!voxel_index || !temp1 || !temp2))
        return fail(h, BFVM_E_ARG, "voxel_index, temp1 and temp2 must be non-NULL");
    HIP_OR_FAIL(h, quiesce(*h));
    return set_agents_impl(*h, n, voxel_index, temp1, temp2);
}

int bfvm_set_agent_rates(bfvm_handle h, int n, const double* position, const double* volume,
                         const double* secretion, const double* saturation, const double* uptake,
                         double dt, int* out_voxel)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (n < 0) return fail(h, BFVM_E_ARG, "n < 0");
    if (n > 0 && (!position || !volume || !secretion || !saturation || !uptake))
        return fail(h, BFVM_E_ARG, "agent arrays must be non-NULL");
    const int S = h->S;
    std::vector<int> vox(n);
    std::vector<double> t1((size_t)n * S), t2((size_t)n * S), cv(n);
    const double ext[3] = {h->nx * h->dx, h->ny * h->dy, h->nz * h->dz};
    const double edge[3] = {h->dx, h->dy, h->dz};
    const int cnt[3] = {h->nx, h->ny, h->nz};
    const double dV = h->dx * h->dy * h->dz;  // BioFVM_mesh.cpp:839
    for (int a = 0; a < n; ++a) {
        int ijk[3];
        for (int d = 0; d < 3; ++d) {
            double p = position[3 * a + d];
            if (!(p >= 0.0 && p <= ext[d]))  // BioFVM_mesh.cpp:208-217
                return fail(h, BFVM_E_ARG, "agent " + std::to_string(a) + " lies outside the domain");
            int q = (int)std::floor(p / edge[d]);  // BioFVM_mesh.cpp:932-946
            ijk[d] = std::min(std::max(q, 0), cnt[d] - 1);
        }
        vox[a] = (ijk[2] * h->ny + ijk[1]) * h->nx + ijk[0];
        agent_constants(S, dt, dV, volume[a], secretion + (size_t)a * S, saturation + (size_t)a * S,
                        uptake + (size_t)a * S, &t1[(size_t)a * S], &t2[(size_t)a * S]);
        cv[a] = dt * volume[a] / dV;  // as agent_constants
    }
    if (out_voxel) std::copy(vox.begin(), vox.end(), out_voxel);
    HIP_OR_FAIL(h, quiesce(*h));
    return set_agents_impl(*h, n, vox.data(), t1.data(), t2.data(), cv.data(), saturation, dt);
}

int bfvm_step(bfvm_handle h, double dt, int flags)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (h->ensemble_stream) HIP_OR_FAIL(h, hipStreamSynchronize(h->ensemble_stream));
    return step_impl(*h, dt, flags);
}

int bfvm_run(bfvm_handle h, double dt, int steps, int flags)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (steps < 0) return fail(h, BFVM_E_ARG, "steps < 0");
    if (h->ensemble_stream) HIP_OR_FAIL(h, hipStreamSynchronize(h->ensemble_stream));
    return run_steps(*h, dt, steps, flags);
}

int bfvm_sync(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    HIP_OR_FAIL(h, quiesce(*h));
    return BFVM_OK;
}

int bfvm_timed_run(bfvm_handle h, double dt, int steps, int flags, double* ms_total,
                   double* per_kernel_ms, int* launches)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (steps < 0) return fail(h, BFVM_E_ARG, "steps < 0");
    Env& e = *h;
    // Factorisation and bulk terms are prepared outside the timed region.
    int rc;
    if ((flags & BFVM_DIFFUSION) && e.uniform && !e.D.empty())
        if ((rc = ensure_coefficients_impl(e, dt))) return rc;
    if (flags & BFVM_BULK) HIP_OR_FAIL(&e, bfvm::ensure_bulk(e, dt));
    hipEvent_t t0, t1;
    HIP_OR_FAIL(&e, hipEventCreate(&t0));
    HIP_OR_FAIL(&e, hipEventCreate(&t1));
    e.kernel_timing = per_kernel_ms != nullptr;
    e.ev_next = 0;
    e.ev_marks.clear();
    if (e.kernel_timing) {
        // the events of a step's launches (up to 8, more with plane chunks),
        // created before the timed region
        const size_t want = (size_t)steps * 2 * 16;
        while (e.ev_pool.size() < want) {
            hipEvent_t ev;
            HIP_OR_FAIL(&e, hipEventCreate(&ev));
            e.ev_pool.push_back(ev);
        }
    }
    HIP_OR_FAIL(&e, quiesce(e));
    HIP_OR_FAIL(&e, hipEventRecord(t0, e.stream));
    rc = run_steps(e, dt, steps, flags);
    HIP_OR_FAIL(&e, hipEventRecord(t1, e.stream));
    HIP_OR_FAIL(&e, hipEventSynchronize(t1));
    e.kernel_timing = false;
    if (rc) return rc;
    float ms = 0;
    HIP_OR_FAIL(&e, hipEventElapsedTime(&ms, t0, t1));
    if (ms_total) *ms_total = ms;
    if (per_kernel_ms) {
        for (int k = 0; k < BFVM_KERNEL_KINDS; ++k) {
            per_kernel_ms[k] = 0;
            if (launches) launches[k] = 0;
        }
        for (auto& m : e.ev_marks) {
            float km = 0;
            HIP_OR_FAIL(&e, hipEventElapsedTime(&km, m.second.first, m.second.second));
            per_kernel_ms[m.first] += km;
            if (launches) launches[m.first] += 1;
        }
    }
    e.ev_marks.clear();
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return BFVM_OK;
}

int bfvm_autotune(bfvm_handle h, double dt)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    Env& e = *h;
    if (e.D.empty()) return fail(&e, BFVM_E_STATE, "coefficients not set (bfvm_set_coefficients)");
    if (!e.uniform) return fail(&e, BFVM_E_MESH, "non-uniform mesh: the LOD solver is not used");
    HIP_OR_FAIL(&e, quiesce(e));
    int rc = ensure_coefficients_impl(e, dt);
    if (rc) return rc;
    void* scratch = nullptr;
    HIP_OR_FAIL(&e, hipMalloc(&scratch, e.nvs * e.elem));
    HIP_OR_FAIL(&e, hipMemcpyAsync(scratch, e.rho, e.nvs * e.elem, hipMemcpyDeviceToDevice, e.stream));
    void* keep = e.rho;
    e.rho = scratch;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    hipError_t st = hipSuccess;
    for (int axis = 0; axis < 3 && st == hipSuccess; ++axis) {
        bfvm::SweepPlan best{};
        float best_ms = 1e30f;
        float best_block_ms = 1e30f;  // the fastest one-block-per-tile x shape, for the fused x sweep
        for (const bfvm::SweepPlan& cand : bfvm::sweep_candidates(e, axis, e.tune_segments)) {
            e.plan[axis] = cand;
            if ((st = bfvm::launch_sweep(e, axis)) != hipSuccess) break;  // warm-up
            (void)hipEventRecord(t0, e.stream);
            for (int rep = 0; rep < 3 && st == hipSuccess; ++rep) st = bfvm::launch_sweep(e, axis);
            (void)hipEventRecord(t1, e.stream);
            if (st != hipSuccess || (st = hipEventSynchronize(t1)) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
            if (axis == 0 && cand.stream == 0 && ms < best_block_ms) {
                best_block_ms = ms;
                e.fused_plan = cand;
            }
        }
        e.plan[axis] = best;
        // then how that shape's tiles are spread over the XCDs
        int best_spread = e.spread[axis];
        best_ms = 1e30f;
        for (int spread : {0, 16, 64, 256, -1}) {
            e.spread[axis] = spread;
            if ((st = bfvm::launch_sweep(e, axis)) != hipSuccess) break;
            (void)hipEventRecord(t0, e.stream);
            for (int rep = 0; rep < 3 && st == hipSuccess; ++rep) st = bfvm::launch_sweep(e, axis);
            (void)hipEventRecord(t1, e.stream);
            if (st != hipSuccess || (st = hipEventSynchronize(t1)) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best_spread = spread;
            }
        }
        e.spread[axis] = best_spread;
    }
    // the whole-cache-line x sweep's cache policy, timed with the y sweep
    // that follows it: with ordinary stores the x sweep leaves its lines in
    // the caches, which slows the next kernel more than it slows itself
    if (st == hipSuccess && e.plan[0].stream == 3 && bfvm::sweep_nt_allowed(e, 0, e.plan[0])) {
        int best_nt = e.plan[0].nt;
        float best_ms = 1e30f;
        for (int nt : {0, 1}) {
            e.plan[0].nt = nt;
            auto pair = [&] {
                hipError_t q = bfvm::launch_sweep(e, 0);
                if (q == hipSuccess && e.lod_dims >= 2) q = bfvm::launch_sweep(e, 1);
                return q;
            };
            if ((st = pair()) != hipSuccess) break;  // warm-up
            (void)hipEventRecord(t0, e.stream);
            for (int rep = 0; rep < 3 && st == hipSuccess; ++rep) st = pair();
            (void)hipEventRecord(t1, e.stream);
            if (st != hipSuccess || (st = hipEventSynchronize(t1)) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best_nt = nt;
            }
        }
        e.plan[0].nt = best_nt;
    }
    // the x sweep with the sources fused in loads the bulk fields beside the
    // store, so its best shape is timed on its own (on the scratch store,
    // with the handle's bulk fields, agents and Dirichlet nodes)
    if (st == hipSuccess && e.fused_sources && (e.have_bulk || e.n_dir > 0)) {
        if (e.have_bulk) st = bfvm::ensure_bulk(e, dt);
        bfvm::SweepPlan best_fused = e.fused_plan;
        float best_ms = 1e30f;
        for (const bfvm::SweepPlan& cand : bfvm::sweep_candidates(e, 0, e.tune_segments)) {
            if (st != hipSuccess) break;
            if (cand.stream != 0) continue;
            e.fused_plan = cand;
            if (!bfvm::sweep_x_fusable(e)) continue;
            auto fused_x = [&] {
                return bfvm::launch_sweep_x_fused(e, e.have_bulk, e.n_agents > 0, e.lod_dims >= 2);
            };
            if ((st = fused_x()) != hipSuccess) break;  // warm-up
            (void)hipEventRecord(t0, e.stream);
            for (int rep = 0; rep < 3 && st == hipSuccess; ++rep) st = fused_x();
            (void)hipEventRecord(t1, e.stream);
            if (st != hipSuccess || (st = hipEventSynchronize(t1)) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best_fused = cand;
            }
        }
        e.fused_plan = best_fused;
    }
    // plane chunking of the x and y sweeps: whole diffusion phases, timed
    if (st == hipSuccess && bfvm::sweep_chunkable(e, 0) && bfvm::sweep_chunkable(e, 1)) {
        int best_chunk = 0;
        float best_ms = 1e30f;
        for (int chunk : {0, 4, 8, 16, 32, 64}) {
            if (chunk >= e.nz) break;
            e.xy_chunk = chunk;
            int rc2 = step_impl(e, dt, BFVM_DIFFUSION);  // warm-up
            (void)hipEventRecord(t0, e.stream);
            for (int rep = 0; rep < 2 && rc2 == BFVM_OK; ++rep) rc2 = step_impl(e, dt, BFVM_DIFFUSION);
            (void)hipEventRecord(t1, e.stream);
            if (rc2 != BFVM_OK || (st = hipEventSynchronize(t1)) != hipSuccess) {
                if (st == hipSuccess) st = hipErrorUnknown;
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, t0, t1);
            if (ms < best_ms) {
                best_ms = ms;
                best_chunk = chunk;
            }
        }
        e.xy_chunk = best_chunk;
    }
    (void)hipStreamSynchronize(e.stream);
    e.rho = keep;
    (void)hipFree(scratch);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    ++e.plan_epoch;
    if (st != hipSuccess) {
        for (auto& p : e.plan) p = bfvm::SweepPlan{};
        e.fused_plan = bfvm::SweepPlan{};
        return hip_fail(&e, st, "bfvm_autotune");
    }
    return BFVM_OK;
}

int bfvm_set_autotune_segments(bfvm_handle h, int on)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    h->tune_segments = on != 0;
    return BFVM_OK;
}

int bfvm_get_autotune_segments(bfvm_handle h, int* on)
{
    bfvm::RandGuard rand_guard;
    if (!h || !on) return fail(h, BFVM_E_ARG, "null argument");
    *on = h->tune_segments ? 1 : 0;
    return BFVM_OK;
}

int bfvm_get_coefficient_builds(bfvm_handle h, unsigned* builds)
{
    bfvm::RandGuard rand_guard;
    if (!h || !builds) return fail(h, BFVM_E_ARG, "null argument");
    *builds = h->coef_builds;
    return BFVM_OK;
}

int bfvm_get_sweep_plan(bfvm_handle h, int axis, int* lines_per_block, int* nodes_per_segment,
                        int* streaming)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    bfvm::SweepPlan p = bfvm::effective_plan(*h, axis);
    if (lines_per_block) *lines_per_block = p.cols;
    if (nodes_per_segment) *nodes_per_segment = p.L;
    if (streaming) *streaming = p.stream;
    return BFVM_OK;
}

int bfvm_set_sweep_plan(bfvm_handle h, int axis, int lines_per_block, int nodes_per_segment,
                        int streaming)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    if (axis == 0) h->fused_plan = bfvm::SweepPlan{};  // the fused x sweep follows the x shape set here
    ++h->plan_epoch;
    if (lines_per_block == 0) {
        h->plan[axis] = bfvm::SweepPlan{};
        return BFVM_OK;
    }
    for (const bfvm::SweepPlan& c : bfvm::sweep_candidates(*h, axis, true))
        if (c.cols == lines_per_block && c.L == nodes_per_segment && c.stream == streaming) {
            h->plan[axis] = c;
            return BFVM_OK;
        }
    return fail(h, BFVM_E_ARG, "launch shape not valid for this axis length");
}

int bfvm_get_sweep_cache_policy(bfvm_handle h, int axis, int* non_temporal)
{
    bfvm::RandGuard rand_guard;
    if (!h || !non_temporal) return fail(h, BFVM_E_ARG, "null argument");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    const bfvm::SweepPlan p = bfvm::effective_plan(*h, axis);
    *non_temporal = p.nt != 0 && bfvm::sweep_nt_allowed(*h, axis, p) ? 1 : 0;
    return BFVM_OK;
}

int bfvm_set_sweep_cache_policy(bfvm_handle h, int axis, int non_temporal)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    bfvm::SweepPlan p = bfvm::effective_plan(*h, axis);
    if (non_temporal && (p.cols == 0 || !bfvm::sweep_nt_allowed(*h, axis, p)))
        return fail(h, BFVM_E_ARG,
                    "non-temporal loads and stores are for y and z sweeps whose tiles' row pieces are "
                    "whole 128-byte lines and for whole-cache-line x sweeps");
    if (p.cols == 0) return BFVM_OK;  // the one-thread-per-line fallback has no policy
    p.nt = non_temporal ? 1 : 0;
    h->plan[axis] = p;
    ++h->plan_epoch;
    return BFVM_OK;
}

int bfvm_set_fused_sources(bfvm_handle h, int on)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    HIP_OR_FAIL(h, quiesce(*h));
    h->fused_sources = on != 0;
    return BFVM_OK;
}

int bfvm_get_fused_sources(bfvm_handle h, int* on)
{
    bfvm::RandGuard rand_guard;
    if (!h || !on) return fail(h, BFVM_E_ARG, "null argument");
    *on = h->fused_sources ? 1 : 0;
    return BFVM_OK;
}

int bfvm_get_sweep_spread(bfvm_handle h, int axis, int* tiles_per_chunk)
{
    bfvm::RandGuard rand_guard;
    if (!h || !tiles_per_chunk) return fail(h, BFVM_E_ARG, "null argument");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    *tiles_per_chunk = h->spread[axis];
    return BFVM_OK;
}

int bfvm_set_sweep_spread(bfvm_handle h, int axis, int tiles_per_chunk)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (axis < 0 || axis > 2) return fail(h, BFVM_E_ARG, "axis must be 0, 1 or 2");
    if (tiles_per_chunk < -1) return fail(h, BFVM_E_ARG, "tiles_per_chunk must be -1, 0 or positive");
    HIP_OR_FAIL(h, quiesce(*h));
    h->spread[axis] = tiles_per_chunk;
    ++h->plan_epoch;
    return BFVM_OK;
}

int bfvm_compute_gradients(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (h->ensemble_stream) HIP_OR_FAIL(h, hipStreamSynchronize(h->ensemble_stream));
    if (!h->grad) HIP_OR_FAIL(h, hipMalloc(&h->grad, h->nvs * 3 * h->elem));
    HIP_OR_FAIL(h, bfvm::launch_gradient(*h, h->grad));
    return BFVM_OK;
}

int bfvm_download_gradients(bfvm_handle h, double* host)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (!host) return fail(h, BFVM_E_ARG, "host is NULL");
    if (!h->grad) return fail(h, BFVM_E_STATE, "no gradients computed (bfvm_compute_gradients)");
    HIP_OR_FAIL(h, quiesce(*h));
    const size_t n = h->nvs * 3;
    if (h->precision == BFVM_FP64) {
        HIP_OR_FAIL(h, hipMemcpy(host, h->grad, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        HIP_OR_FAIL(h, hipMemcpy(tmp.data(), h->grad, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    }
    return BFVM_OK;
}

int bfvm_gradient_device_ptr(bfvm_handle h, void** ptr)
{
    bfvm::RandGuard rand_guard;
    if (!h || !ptr) return fail(h, BFVM_E_ARG, "null argument");
    *ptr = h->grad;
    return BFVM_OK;
}

int bfvm_set_lod_dimensions(bfvm_handle h, int dims)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (dims != 2 && dims != 3) return fail(h, BFVM_E_ARG, "dims must be 2 or 3");
    if (dims == 2 && h->nz != 1) return fail(h, BFVM_E_MESH, "LOD_2D needs a mesh with nz == 1");
    HIP_OR_FAIL(h, quiesce(*h));
    if (h->lod_dims != dims) {
        h->lod_dims = dims;
        h->coef_valid = false;  // the decay split changes
    }
    return BFVM_OK;
}

int bfvm_set_graph_steps(bfvm_handle h, int steps_per_graph)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (steps_per_graph < 0 || steps_per_graph > 4096)
        return fail(h, BFVM_E_ARG, "steps_per_graph must be 0 to 4096");
    HIP_OR_FAIL(h, quiesce(*h));
    h->graph_steps = steps_per_graph;
    free_graph(*h);
    return BFVM_OK;
}

int bfvm_get_graph_steps(bfvm_handle h, int* steps_per_graph)
{
    bfvm::RandGuard rand_guard;
    if (!h || !steps_per_graph) return fail(h, BFVM_E_ARG, "null argument");
    *steps_per_graph = h->graph_steps;
    return BFVM_OK;
}

int bfvm_get_sweep_chunk(bfvm_handle h, int* planes)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (planes) *planes = h->xy_chunk;
    return BFVM_OK;
}

int bfvm_set_sweep_chunk(bfvm_handle h, int planes)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (planes < 0) return fail(h, BFVM_E_ARG, "planes < 0");
    h->xy_chunk = planes;
    return BFVM_OK;
}

int bfvm_get_info(bfvm_handle h, int* nx, int* ny, int* nz, int* n_substrates, int* precision,
                  size_t* density_bytes)
{
    bfvm::RandGuard rand_guard;
    if (!h) return fail(nullptr, BFVM_E_ARG, "null handle");
    if (nx) *nx = h->nx;
    if (ny) *ny = h->ny;
    if (nz) *nz = h->nz;
    if (n_substrates) *n_substrates = h->S;
    if (precision) *precision = h->precision;
    if (density_bytes) *density_bytes = h->nvs * h->elem;
    return BFVM_OK;
}

int bfvm_density_device_ptr(bfvm_handle h, void** ptr)
{
    bfvm::RandGuard rand_guard;
    if (!h || !ptr) return fail(h, BFVM_E_ARG, "null argument");
    *ptr = h->rho;
    return BFVM_OK;
}

const char* bfvm_last_error(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    return h ? h->err.c_str() : g_thread_err.c_str();
}

}  // extern "C"
