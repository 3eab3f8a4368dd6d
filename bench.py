"""Benchmark of the diffusion-decay step on one MI355X.

    python bench.py [--steps K] [--warmup W]
                    [--workload precision_400|drop_in_200|lod_50|ensemble_64|differentiation_128]

Default workload (the configuration the headline metric is quoted on):
``precision_400`` — 400^3 voxels x 8 substrates (dx=10 um, dt=0.01 min), D
log-uniform in [1e2,1e5], lambda log-uniform in [1e-3,1], 1e6 agents in a
Gaussian cluster (sigma=600 um); each step = agent secretion/uptake + LOD
diffusion-decay.  The FP32 store is timed, then the FP64 store on the same
inputs; the FP32 result is compared with the FP64 one after the same steps.

Prints ONE JSON line.  ``value`` is voxel-substrate updates/s of the FP32
full step (nx*ny*nz*S*K / wall seconds of K steps, inputs resident in HBM).
``roofline`` prices the dominant kernel's algorithmic HBM bytes per launch
against the 8 TB/s HBM3E peak, using its average launch time from HIP event
pairs recorded around every launch, on the library's stream, in up to 10
more steps run right after the timed ones (pairs inside the timed steps
would slow them: 1.5 % at 400^3 x 8).  ``cpu_baseline`` times BioFVM itself (compiled from
source, oracle/_ref) on the whole workload on this host's cores (a child process set up while
the GPU timings run, its steps timed after them; see cpu_child), ``cpu_baseline_1core`` the
same on one thread.  The line also carries compact ``ensemble_64`` and
``differentiation_128`` objects (ensemble_result, differentiation_result) and
compact ``lod_50`` and ``drop_in_200`` objects (config_result).

``--workload ensemble_64`` times the batching form instead: 64 independent
uneven grids (FP32 stores) advanced together, one step = one HIP graph replay
of batched agent and sweep launches; its roofline prices the three batched
sweeps of a diffusion-only step.  ``--workload differentiation_128`` times the
adjoint (see differentiation_result).  ``--ceiling`` measures the bandwidth
ceilings beside the sweeps (ceiling_bench).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(case, w):
    """Per-step algorithmic HBM bytes: 3 sweeps x (1R+1W) of the store (w bytes
    per value), bulk terms A/B read once (FP64), agents (voxel key + FP64
    temp1/temp2), Dirichlet nodes (index + values) before each sweep."""
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    S = case["S"]
    ph = case["phases"]
    b = nvs * 6 * w
    if ph & 1 and case.get("bulk_supply") is not None:
        b += nvs * 2 * 8  # read A and B (FP64); the store pass is shared with the x-sweep when fused
    if ph & 2 and case.get("agent_volume") is not None:
        b += len(case["agent_volume"]) * (4 + 2 * S * 8)  # voxel key + FP64 temp1/temp2
    if ph & 4 and case.get("dir_idx") is not None:
        b += 3 * len(case["dir_idx"]) * (4 + S * w)
    return b


def survey_bytes(case, w):
    """The same per-step figure with every term at the store's width w (agent
    constants counted as w-byte values, bulk A/B as w-byte values): the
    formula the HBM targets for this path are stated in."""
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    S, ph = case["S"], case["phases"]
    b = nvs * 6 * w
    if ph & 1 and case.get("bulk_supply") is not None:
        b += nvs * 2 * w
    if ph & 2 and case.get("agent_volume") is not None:
        b += len(case["agent_volume"]) * (4 + 2 * S * w)
    if ph & 4 and case.get("dir_idx") is not None:
        b += 3 * len(case["dir_idx"]) * (4 + S * w)
    return b


def source_bytes(case, w):
    """Algorithmic bytes of the sources the fused x sweep applies: bulk A/B
    (FP64) and two Dirichlet overwrites (the agents keep their own kernel)."""
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    S, ph, b = case["S"], case["phases"], 0
    if ph & 1 and case.get("bulk_supply") is not None:
        b += nvs * 2 * 8
    if ph & 4 and case.get("dir_idx") is not None:
        b += 2 * len(case["dir_idx"]) * (4 + S * w)
    return b


def kernel_bytes(case, kernel, w, fused=False):
    """Algorithmic bytes of one step's launches of `kernel` (for the sweeps: one
    whole sweep, whether it runs as one launch or plane chunk by chunk; the x
    sweep with the sources fused in adds theirs)."""
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    if kernel == "sweep_x" and fused:
        return 2 * nvs * w + source_bytes(case, w)
    if kernel.startswith("sweep"):
        return 2 * nvs * w
    if kernel == "bulk":
        return nvs * (2 * w + 2 * 8)
    if kernel == "agents":
        return len(case["agent_volume"]) * (4 + 2 * case["S"] * 8)
    if kernel == "dirichlet":
        return len(case["dir_idx"]) * (4 + case["S"] * w)
    return 0


def measured_traffic(workload, kernel_name):
    """HBM bytes per launch of `kernel_name` from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"k_sweep_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel_name)
    return k["traffic_bytes_per_launch"] if k else None


def lod_only(per, nvs, w):
    """The three sweeps alone (HIP-event launch times per step): updates/s and HBM fraction."""
    ms = sum(per[k][0] for k in ("sweep_x", "sweep_y", "sweep_z"))
    return {"value": nvs / (ms * 1e-3), "ms_per_step": ms,
            "hbm_frac": 6 * nvs * w / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def time_precision(case, precision, steps, warmup, fused=True, plan_x=""):
    from biofvm_amd import FP32, setup_case
    m = setup_case(case, precision)
    dt, ph = case["dt"], case["phases"]
    m.fused_sources = fused
    m.autotune(dt)  # launch shapes, picked on a scratch copy outside the timed region
    if plan_x and precision == FP32:  # a fixed x shape (for counter runs of one kernel)
        m.set_sweep_plan(0, *[int(v) for v in plan_x.split(",")])
    m.run(dt, warmup, ph)
    m.sync()
    t0 = time.perf_counter()
    ev_ms, _ = m.timed_run(dt, steps, ph)
    m.sync()
    wall = time.perf_counter() - t0
    # right after, the same step again with an event pair around every launch
    # (on the library's stream): recorded inside the timed steps, the pairs
    # cost 1.5 % of a 400^3 x 8 step and 9 % of a 200^3 x 4 one
    prof_steps = max(1, min(steps, 10))
    _, per = m.timed_run(dt, prof_steps, ph, per_kernel=True)
    # per kernel: (ms per step, launches per step)
    per = {k: (v[0] / prof_steps, v[1] / prof_steps) for k, v in per.items()}
    m.sync()
    total_steps = warmup + steps + prof_steps
    return m, wall, ev_ms, per, total_steps


def sources_fused(case, per):
    """Whether the bulk update and the first two Dirichlet overwrites ran
    inside the x sweep: no bulk launch, one Dirichlet launch per step (before z)."""
    ph = case["phases"]
    has_bulk = bool(ph & 1 and case.get("bulk_supply") is not None)
    has_dir = bool(ph & 4 and case.get("dir_idx") is not None and len(case["dir_idx"]) > 0)
    return bool((has_bulk or has_dir) and per.get("bulk", (0, 0))[1] == 0
                and (not has_dir or per.get("dirichlet", (0, 0))[1] < 2))


def config_result(workload, steps, warmup):
    """lod_50 or drop_in_200 (FP64 stores), compact, for the default line: the
    C ABI path timed as the main workload is (autotuned, graph replays), the
    step's HBM fraction and the dominant kernel's roofline (its own
    --workload gives the full line, with the facade and the CPU baseline)."""
    from biofvm_amd import FP64
    from biofvm_amd import workloads as W
    case = getattr(W, workload)()
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    m, wall, ev_ms, per, _ = time_precision(case, FP64, steps, warmup)
    plans = m.sweep_plans()
    m.close()
    fused = sources_fused(case, per)
    dom = max(per, key=lambda k: per[k][0])
    dom_ms, dom_n = per[dom]
    dom_bytes = kernel_bytes(case, dom, 8, fused)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    n_agents = 0 if case.get("agent_volume") is None else len(case["agent_volume"])
    n_dir = 0 if case.get("dir_idx") is None else len(case["dir_idx"])
    return {"value": nvs * steps / wall, "unit": "voxel-substrate updates/s", "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * wall / steps, "events_ms_per_step": ev_ms / steps, "store_dtype": "f64",
            "config": {"workload": f"{workload}: {case['nx']}^3 voxels x {case['S']} substrates, "
                                   f"{n_agents} agents, {n_dir} Dirichlet nodes, phases={case['phases']}"},
            "hbm_frac_step": algorithmic_bytes(case, 8) / (wall / steps) / 1e9 / HBM_PEAK_GBS,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_step": dom_bytes,
                         "ms_per_step": dom_ms, "launches_per_step": dom_n},
            "sources_fused_in_x_sweep": fused,
            "kernels_ms_per_step": {k: v[0] for k, v in per.items()},
            "sweep_plans_lines_nodes_streaming": dict(zip("xyz", plans))}


def error_curve(case, total, every):
    """FP32 store against the FP64 store on the same inputs, from step 0, every
    `every` steps up to `total`: max normwise, max elementwise (values above
    1e-3 of the largest) and L2 relative errors."""
    from biofvm_amd import FP32, FP64, setup_case
    dt, ph = case["dt"], case["phases"]
    m32 = setup_case(case, FP32)
    m64 = setup_case(case, FP64)
    m32.autotune(dt)
    m64.autotune(dt)
    curve = []
    for n in range(every, total + 1, every):
        m32.run(dt, every, ph)
        m64.run(dt, every, ph)
        a, b = m32.download(), m64.download()
        diff = np.abs(a - b)
        scale = np.abs(b)
        big = scale > 1e-3 * scale.max()
        curve.append({"step": n, "max_rel_normwise": float(diff.max() / scale.max()),
                      "max_rel_elementwise": float((diff[big] / scale[big]).max()),
                      "l2_rel": float(np.linalg.norm(a - b) / np.linalg.norm(b))})
        del a, b, diff, scale, big
        print(f"[bench] error curve step {n}: {curve[-1]['max_rel_normwise']:.3e}", file=sys.stderr, flush=True)
    m32.close()
    m64.close()
    return curve


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_case(workload):
    """The workload the CPU baseline runs: the whole configuration (the
    drop-in one with whole-vector Dirichlet values, the reference's only
    kind)."""
    from biofvm_amd import workloads as W
    if workload == "drop_in_200":
        return W.drop_in_200(mask_substrate0=False)
    return getattr(W, workload)()


def cpu_child(workload):
    """Child process of the CPU baseline: sets the compiled reference
    (oracle/_ref) up on the whole workload, says so on stdout, waits for a
    line on stdin (sent once the GPU timings are done, so the two never share
    the host), then times, on the OpenMP threads its environment names: one
    untimed warm-up step (the reference builds its Thomas coefficients in its
    first diffusion call), K full steps, K/2 LOD-only steps, then full steps
    on one thread.  Prints one JSON line and exits without tearing the
    reference's 64e6 voxel objects down."""
    import oracle
    t0 = time.perf_counter()
    case = cpu_case(workload)
    print(f"[bench cpu] {workload} generated ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sess = oracle.RefSession(case)
    setup = time.perf_counter() - t0
    print(f"[bench cpu] reference set up ({setup:.1f} s)", file=sys.stderr, flush=True)
    print(json.dumps({"ready": True}), flush=True)
    sys.stdin.readline()
    n = int(os.environ["OMP_NUM_THREADS"])
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    ph = case["phases"]
    warm = sess.steps(1, ph, n)
    k = int(min(200, max(5, round(5.0 / max(warm, 1e-6)))))
    full = sess.steps(k, ph, n)
    print(f"[bench cpu] {k} steps on {n} threads: {full:.2f} s", file=sys.stderr, flush=True)
    k_lod = max(3, k // 2)
    lod = sess.steps(k_lod, 4, n)
    k1 = int(max(1, min(200, round(5.0 / max(full / k * n, 1e-6)))))
    one = sess.steps(k1, ph, 1)
    print(f"[bench cpu] {k1} steps on 1 thread: {one:.2f} s", file=sys.stderr, flush=True)
    print(json.dumps({"nvs": nvs, "threads": n, "setup_s": setup, "warmup_s": warm,
                      "steps": k, "seconds": full, "lod_steps": k_lod, "lod_seconds": lod,
                      "steps_1core": k1, "seconds_1core": one}), flush=True)
    os._exit(0)


class CpuBaseline:
    """The compiled reference on the whole workload in a child process
    (cpu_child), started before the GPU is touched so that its set-up
    overlaps the GPU timings; its steps run after them."""

    def __init__(self, workload):
        import subprocess
        import oracle
        self.workload = workload
        self.proc = None
        if not oracle.have_ref():
            return
        self.threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        self.threads = min(self.threads, os.cpu_count() or self.threads)
        env = dict(os.environ, OMP_NUM_THREADS=str(self.threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        self.proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child", workload],
                                     stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, env=env)

    def _json_line(self):
        for line in self.proc.stdout:  # the reference prints lines of its own
            if line.startswith("{"):
                return json.loads(line)
        raise RuntimeError(f"CPU baseline child ended (code {self.proc.wait()})")

    def result(self, case):
        if self.proc is None:
            return port_baseline(case, self.workload), None
        self._json_line()  # ready
        self.proc.stdin.write("go\n")
        self.proc.stdin.flush()
        r = self._json_line()
        self.proc.wait()
        nvs = r["nvs"]
        build = "g++ -O3 -march=x86-64-v3 -fopenmp (oracle/Makefile)"
        common = {"unit": "voxel-substrate updates/s", "kind": "reference", "cpu_model": cpu_model(),
                  "timer": "std::chrono::steady_clock around the step loop"}
        out = dict(common, value=nvs * r["steps"] / r["seconds"], cores=r["threads"],
                   ms_per_step=1e3 * r["seconds"] / r["steps"],
                   sample=f"the whole {self.workload} workload, {r['steps']} full steps after one warm-up "
                          f"step, {r['threads']} threads, OMP_PROC_BIND=close OMP_PLACES=cores; {build}; "
                          f"set-up {r['setup_s']:.0f} s, outside the timing",
                   lod_only={"value": nvs * r["lod_steps"] / r["lod_seconds"], "steps": r["lod_steps"],
                             "ms_per_step": 1e3 * r["lod_seconds"] / r["lod_steps"]})
        one = dict(common, value=nvs * r["steps_1core"] / r["seconds_1core"], cores=1,
                   ms_per_step=1e3 * r["seconds_1core"] / r["steps_1core"],
                   sample=f"the same session, {r['steps_1core']} full steps on one thread")
        return out, one

    def cancel(self):
        if self.proc is not None and self.proc.poll() is None:
            self.proc.kill()
            self.proc.wait()


def port_baseline(case, workload):
    """Without the compiled reference: the C restatement (one thread) on a
    16-plane z-slab of the workload."""
    import oracle
    from biofvm_amd import workloads as W
    nz_s = min(16, case["nz"])
    nx, ny, S, dx = case["nx"], case["ny"], case["S"], case["dx"]
    slab = W.make_case(nx, ny, nz_s, S, dx, case["D"], case["lam"], case["dt"], 2, 4,
                       case["rho0"][: nx * ny * nz_s * S])
    t0 = time.perf_counter()
    oracle.oracle_run(slab)
    secs = time.perf_counter() - t0
    return {"value": nx * ny * nz_s * S * 2 / secs, "unit": "voxel-substrate updates/s", "cores": 1,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nx}x{ny}x{nz_s} slab of {workload}, LOD only, 2 steps, {secs:.2f} s wall"}


def facade_bench(steps=50, warmup=5, every_step_steps=3):
    """The drop-in configuration (200^3 x 4, Dirichlet faces, bulk callbacks,
    1e5 agents, FP64) driven through the reference's own API: the C++ façade
    (biofvm_amd/facade) under tests/cpp/facade_driver, a program written like
    the reference's examples (bulk -> cells -> diffusion per step).  Dirichlet
    nodes overwrite the whole value vector {1, 0, 0, 0}, the reference's
    semantics (the C-ABI drop_in_200 run masks substrates 1-3 instead).  Timed
    by the driver's steady clock over `steps` steps after `warmup`, until the
    device has finished; once with the bulk fields kept on the device (the
    default), once with the callbacks evaluated at every step as the
    reference does (`every_step_steps` steps).  The result after
    warmup + steps steps is compared with the C ABI's on the same inputs."""
    import subprocess
    import tempfile
    from biofvm_amd import FP64, setup_case
    from biofvm_amd import workloads as W
    driver = os.path.join(ROOT, "tests", "cpp", "facade_driver")
    case = W.drop_in_200(steps=steps, mask_substrate0=False)
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    out = {"workload": "drop_in_200 through the reference API (C++ facade): 200^3 x 4, FP64, "
                       "Dirichlet faces, bulk callbacks, 1e5 agents", "warmup": warmup}
    with tempfile.TemporaryDirectory() as d:
        runs = {}
        for mode, n, extra in (("fields_on_device", steps, []),
                               ("callbacks_every_step", every_step_steps, ["--bulk-every-step"])):
            c = dict(case, steps=n)
            cin, cout = os.path.join(d, f"{mode}.bin"), os.path.join(d, f"{mode}.out")
            W.write_driver_case(cin, c, FP64)
            r = subprocess.run([driver, cin, cout, "--time", str(warmup), *extra], capture_output=True,
                               text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"facade driver failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
            line = json.loads(r.stdout.strip().splitlines()[-1])
            line["value"] = nvs / (line["ms_per_step"] * 1e-3)
            runs[mode] = (line, cout)
            print(f"[bench] facade {mode}: {line['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
        main_line, main_out = runs["fields_on_device"]
        got = np.fromfile(main_out, dtype="<f8")
        with setup_case(dict(case, steps=warmup + steps), FP64) as m:
            m.autotune(case["dt"])
            m.run(case["dt"], warmup + steps, case["phases"])
            ref = m.download()
    out.update({
        "steps": steps, "ms_per_step": main_line["ms_per_step"], "value": main_line["value"],
        "unit": "voxel-substrate updates/s", "coefficient_builds": main_line["coefficient_builds"],
        "bulk_host_seconds_per_step": main_line["bulk_rates_host_seconds"] / steps,
        "max_rel_diff_vs_c_abi": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))),
        "callbacks_every_step": {
            "steps": every_step_steps, "ms_per_step": runs["callbacks_every_step"][0]["ms_per_step"],
            "bulk_host_seconds_per_step":
                runs["callbacks_every_step"][0]["bulk_rates_host_seconds"] / every_step_steps},
    })
    return out


def calibrate_copy(mib):
    """Copies `mib` MiB device to device with hipMemcpy (the runtime's copy
    kernel), three times, so rocprofv3 counters of a kernel with known bytes
    (read mib, write mib) can scale those of the sweeps."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_mib": mib, "bytes_read": n, "bytes_written": n}), flush=True)


def ceiling_bench(args):
    """`--ceiling`: the bandwidth ceilings the sweeps are judged against,
    from libbfvm_ceiling.so (biofvm_amd/csrc/ceiling.hip): streaming passes
    over a buffer the size of the FP32 400^3 x 8 store (in place, copy, read,
    write, each with plain and non-temporal accesses, the faster one as
    `gbps`; the in-place pass also over a 128 MB region that stays in the
    Infinity Cache), then per store precision the three sweeps of the
    library (autotuned, diffusion-only steps, HIP events per launch) beside
    the same sweeps' access pattern alone (same tiles, block shapes, strides,
    XCD numbering and cache policy; loads, one barrier, stores; no solve)."""
    import ctypes
    from biofvm_amd import DIFFUSION, FP32, FP64, setup_case
    from biofvm_amd import build as B
    from biofvm_amd import workloads as W
    lib = ctypes.CDLL(B.CEILING_LIB)
    dpp = ctypes.POINTER(ctypes.c_double)
    lib.bfc_stream_pass.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                    dpp, dpp]
    lib.bfc_sweep_pattern.argtypes = [ctypes.c_int] * 9 + [dpp, dpp]
    lib.bfc_sweep_pattern_xt.argtypes = [ctypes.c_int] * 6 + [dpp, dpp]
    n, S = 400, 8
    store = n ** 3 * S * 4
    out = {"mode": "ceiling", "unit": "GB/s", "peak": HBM_PEAK_GBS, "reps": 10}

    def stream(kind, buf, region):
        res = {"region_bytes": region}
        for nt, tag in ((0, "plain"), (1, "non_temporal")):
            g, ms = ctypes.c_double(), ctypes.c_double()
            rc = lib.bfc_stream_pass(kind, nt, buf, region, 10, ctypes.byref(g), ctypes.byref(ms))
            if rc:
                raise RuntimeError(f"bfc_stream_pass({kind}, {nt}) failed: {rc}")
            res[tag] = {"gbps": g.value, "ms": ms.value}
        res["gbps"] = max(res["plain"]["gbps"], res["non_temporal"]["gbps"])
        return res

    out["streams"] = {
        "in_place_read_write": stream(0, store, store),
        "in_place_read_write_128MB_region": stream(0, store, 128 << 20),
        "copy": stream(1, store // 2, store // 2),
        "read": stream(2, store, store),
        "write": stream(3, store, store),
    }
    print(f"[bench] streams: {json.dumps(out['streams'])}", file=sys.stderr, flush=True)
    case = W.precision_400(steps=1)
    case = dict(case, phases=DIFFUSION, agent_pos=None, agent_volume=None)
    for prec, name, elem in ((FP32, "fp32", 4), (FP64, "fp64", 8)):
        m = setup_case(case, prec)
        m.autotune(case["dt"])
        plans = m.sweep_plans()
        spreads = [m.sweep_spread(a) for a in range(3)]
        policies = m.sweep_cache_policies()
        m.run(case["dt"], 2, DIFFUSION)
        _, per = m.timed_run(case["dt"], 10, DIFFUSION, per_kernel=True)
        m.close()
        res = {}
        for a, k in enumerate(("x", "y", "z")):
            cols, L, streaming = plans[a]
            ms = per[f"sweep_{k}"][0] / 10
            sweep = 2 * n ** 3 * S * elem / (ms * 1e-3) / 1e9
            g, pms = ctypes.c_double(), ctypes.c_double()
            if streaming == 3:  # the whole-cache-line x kernel has its own pattern
                rc = lib.bfc_sweep_pattern_xt(elem, n, cols, L, spreads[a], 10, ctypes.byref(g), ctypes.byref(pms))
            else:
                rc = lib.bfc_sweep_pattern(elem, n, S, a, cols, L, spreads[a], policies[a], 10, ctypes.byref(g),
                                           ctypes.byref(pms))
            if rc:
                raise RuntimeError(f"bfc_sweep_pattern({k}) failed: {rc}")
            res[k] = {"plan_lines_nodes_streaming": plans[a], "spread": spreads[a],
                      "non_temporal": policies[a], "sweep_ms": ms,
                      "sweep_gbps": sweep, "pattern_ms": pms.value, "pattern_gbps": g.value,
                      "sweep_over_pattern": sweep / g.value,
                      "sweep_over_in_place": sweep / out["streams"]["in_place_read_write"]["gbps"]}
        out[name] = res
        print(f"[bench] {name}: {json.dumps(res)}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def ensemble_result(steps, warmup, with_fp64):
    """ensemble_64: 64 uneven grids advanced together, one HIP graph replay per
    step (FP32 stores), then the FP64 stores on the same inputs.  Returns the
    line's fields (value, timings, the batched sweeps' roofline)."""
    from biofvm_amd import FP32, FP64, Ensemble, setup_case
    from biofvm_amd import workloads as W
    cases = W.ensemble_64()
    dt, ph = cases[0]["dt"], cases[0]["phases"]
    nvs = sum(c["nx"] * c["ny"] * c["nz"] * c["S"] for c in cases)
    n_agents = sum(len(c["agent_volume"]) for c in cases)

    def run(precision):
        members = [setup_case(c, precision) for c in cases]
        ens = Ensemble(members)
        ens.autotune(dt)
        L = ens.axis_segments()  # nodes per segment of the batched x, y, z sweeps
        # the warm-up replays the multi-step graph the timed run replays, and
        # runs long enough (about 20 ms) for the first timed step to run at
        # the steady state's clocks (shorter warm-ups timed the first run a
        # few % slow, graph or eager alike)
        wu = max(warmup, 4 * ens.graph_steps)
        ens.run(dt, wu, ph, graph=True)
        ens.sync()
        t0 = time.perf_counter()
        ev_ms = ens.timed_run(dt, steps, ph, graph=True)
        wall = time.perf_counter() - t0
        eager_ms = ens.timed_run(dt, max(1, steps // 4), ph, graph=False) / max(1, steps // 4)
        return ens, members, L, wall, ev_ms, eager_ms, wu

    ens, members, L, wall, ev_ms, eager_ms, wu = run(FP32)
    out = [m.download() for m in members]
    # diffusion-only steps (after the result was taken) price the three batched
    # sweeps: whole multi-step graphs (each replay boundary costs a few us)
    sweep_steps = 3 * ens.graph_steps
    # the diffusion-only graphs captured and replayed for ~20 ms before the timing
    ens.run(dt, 4 * ens.graph_steps, 4, graph=True)
    sweep_ms = ens.timed_run(dt, sweep_steps, 4, graph=True) / sweep_steps
    knobs = {"segment_nodes_xyz": L, "x_chunks": ens.x_chunks, "x_row_tiles": ens.row_tiles,
             "yz_non_temporal": ens.non_temporal}
    ens.close()
    for m in members:
        m.close()
    sweep_bytes = 3 * 2 * nvs * 4
    achieved = sweep_bytes / (sweep_ms * 1e-3) / 1e9  # three launches' bytes over their summed time
    res = {
        "value": nvs * steps / wall, "unit": "voxel-substrate updates/s", "steps": steps,
        "warmup": wu,
        "ms_per_step": 1e3 * wall / steps, "events_ms_per_step": ev_ms / steps, "eager_ms_per_step": eager_ms,
        "store_dtype": "f32",
        "config": {"workload": f"ensemble_64: {len(cases)} grids, {nvs} voxel-substrate values, "
                               f"{n_agents} agents, dt={dt}, phases={ph}, one HIP graph per step"},
        "roofline": {"bound": "hbm", "kernel": "k_sweep_xt_batch or k_sweep_rows_batch (x), k_sweep_batch (y, z): "
                                               "3 per step, average",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": sweep_bytes / 3, "avg_launch_ms": sweep_ms / 3},
    }
    res.update(knobs)
    if with_fp64:
        ens64, mem64, L64, wall64, ev64, eager64, _ = run(FP64)
        err_n, err_l2n, err_l2d = 0.0, 0.0, 0.0
        for m, a in zip(mem64, out):
            b = m.download()
            err_n = max(err_n, float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
            err_l2n += float(np.sum((a - b) ** 2))
            err_l2d += float(np.sum(b ** 2))
        ens64.close()
        for m in mem64:
            m.close()
        res["fp64"] = {"value": nvs * steps / wall64, "ms_per_step": 1e3 * wall64 / steps, "segment_nodes_xyz": L64}
        res["error_vs_fp64"] = {"steps": warmup + steps + max(1, steps // 4),
                                "max_rel_normwise_per_member": err_n, "l2_rel": (err_l2n / err_l2d) ** 0.5}
    return res


def ensemble_bench(args):
    res = ensemble_result(args.steps, args.warmup, not args.no_fp64)
    line = {
        "metric": "voxel-substrate updates/s (ensemble step: agents + LOD diffusion-decay, all members)",
        "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded numpy default_rng; see biofvm_amd/workloads.py)",
    }
    line.update(res)
    print(json.dumps(line), flush=True)


def adjoint_bytes(case, w, kept=True):
    """Algorithmic HBM bytes of one backward step of the adjoint: the step
    recomputed from the segment's checkpoint once when every step's states
    are kept (else twice: the step starts, then the step again keeping its
    states; 3 sweeps of the store, w bytes per value, read and written; the
    agents' FP64 constants), the transposed sweeps of the FP64 adjoint state
    (3, read and written) with each axis's D/lambda sums fused in (the
    forward state read once per axis), the agents' adjoint (their constants
    once more)."""
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    agents = len(case["agent_volume"]) * (4 + 2 * case["S"] * 8)
    return (1 if kept else 2) * (6 * nvs * w + agents) + 6 * nvs * 8 + 3 * nvs * w + agents


def differentiation_result(full):
    """Adjoint of 500 steps at 128^3 x 2 with 2e4 agents (FP32 store,
    checkpoints every 50 steps).  The target fields come from the hidden
    parameters (FP64 run); the gradient is taken at the start of a fit (D,
    lambda and secretion rates off by a factor exp(U[-0.3, 0.3])).  Times the
    recorded forward and the backward pass and checks the gradient against
    FP64 differences of J (summed with math.fsum) on random directions in the
    joint parameter space (initial field, D, lambda, secretion, uptake):
    4th-order central differences (Richardson, h = 1e-4 relative,
    truncation O(h^4)).  full: 32 directions and the FP64 adjoint beside the
    FP32 one; otherwise 2 directions, FP32 adjoint only."""
    import math
    from biofvm_amd import FP32, FP64, setup_case
    from biofvm_amd import workloads as W
    hidden = W.differentiation_128()
    nvs = hidden["nx"] * hidden["ny"] * hidden["nz"] * hidden["S"]
    dt, ph, steps = hidden["dt"], hidden["phases"], hidden["steps"]
    n_agents = len(hidden["agent_volume"])
    rng = np.random.default_rng(12)
    with setup_case(hidden, FP64) as m:
        m.run(dt, steps, ph)
        target = m.download()
    case = dict(hidden)
    for k, shape in (("D", hidden["S"]), ("lam", hidden["S"]), ("agent_secretion", hidden["agent_secretion"].shape)):
        case[k] = np.asarray(hidden[k]) * np.exp(rng.uniform(-0.3, 0.3, shape))

    def adjoint(precision):
        m = setup_case(case, precision)
        m.autotune(dt)
        m.sync()
        t0 = time.perf_counter()
        m.tape_record(dt, steps, ph, 50)
        m.sync()
        t_fwd = time.perf_counter() - t0
        loss = m.tape_seed_l2(target)
        t0 = time.perf_counter()
        g = m.tape_backward(n_agents)
        t_bwd = time.perf_counter() - t0
        kept = m.tape_states_kept()
        m.close()
        return g, t_fwd, t_bwd, loss, kept

    g, t_fwd, t_bwd, loss, kept = adjoint(FP32)
    step_bytes = adjoint_bytes(case, 4, kept)
    achieved = step_bytes / (t_bwd / steps) / 1e9
    out = {"value": nvs * steps / t_bwd, "unit": "voxel-substrate updates/s", "steps": steps,
           "ms_per_step": 1e3 * t_bwd / steps, "forward_s": t_fwd, "backward_s": t_bwd, "loss": loss,
           "store_dtype": "f32", "states_kept": kept,
           "config": {"workload": f"differentiation_128: {case['nx']}^3 voxels x {case['S']} substrates, "
                                  f"{n_agents} agents, {steps} steps, checkpoints every 50"},
           "roofline": {"bound": "hbm", "kernel": "backward step (recomputed forward, transposed sweeps, "
                                                  "D/lambda sums, agents' adjoint)",
                        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_step": step_bytes}}
    keys = ["rho0", "D", "lam", "agent_secretion", "agent_uptake"]
    gk = {"rho0": "grad_rho0", "D": "grad_D", "lam": "grad_lambda", "agent_secretion": "grad_secretion",
          "agent_uptake": "grad_uptake"}
    grads = {"fp32": g}
    if full:
        g64, _, t_bwd64, _, _ = adjoint(FP64)
        grads["fp64"] = g64
        out["fp64_backward_s"] = t_bwd64
        out["fp32_vs_fp64_adjoint"] = {
            k: {"max_abs_over_max": float(np.abs(g[v] - g64[v]).max() / np.abs(g64[v]).max()),
                "l2_rel": float(np.linalg.norm(g[v] - g64[v]) / np.linalg.norm(g64[v]))}
            for k, v in gk.items()}

    def J(ds, h):
        c = dict(case)
        for k, d in ds.items():
            c[k] = np.asarray(case[k]) + h * d
        with setup_case(c, FP64) as mm:
            mm.run(dt, steps, ph)
            r = mm.download() - target
        return 0.5 * math.fsum(r * r)

    h, n_dir = 1e-4, (32 if full else 2)
    worst = {k: 0.0 for k in grads}
    for _ in range(n_dir):
        ds = {k: rng.normal(size=np.shape(case[k])) * np.asarray(case[k]) for k in keys}
        fd = (8 * (J(ds, h) - J(ds, -h)) - (J(ds, 2 * h) - J(ds, -2 * h))) / (12 * h)
        for name, gr in grads.items():
            ad = math.fsum(np.concatenate([np.asarray(gr[gk[k]]).reshape(-1) * ds[k].reshape(-1) for k in keys]))
            worst[name] = max(worst[name], abs(ad - fd) / abs(fd))
    out["fd_check"] = {"directions": n_dir, "h": h, "difference": "4th-order central (Richardson), J by fsum",
                       "space": "joint (rho0, D, lambda, secretion, uptake)",
                       "max_rel_err": worst["fp32"]}
    if "fp64" in worst:
        out["fd_check"]["fp64_adjoint_max_rel_err"] = worst["fp64"]
    return out


def differentiation_bench(args):
    res = differentiation_result(not args.no_fp64)
    line = {
        "metric": "voxel-substrate adjoint step updates/s (backward pass incl. checkpoint recomputation)",
        "n_gpus": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded numpy default_rng; see biofvm_amd/workloads.py)",
    }
    line.update(res)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="precision_400",
                    choices=["precision_400", "drop_in_200", "lod_50", "ensemble_64",
                             "differentiation_128"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="leave the compact ensemble_64, differentiation_128, lod_50 and drop_in_200 objects "
                         "out of the default line")
    ap.add_argument("--plan-x", default="", metavar="LINES,NODES,KERNEL",
                    help="after autotuning, run the FP32 x sweep with this shape (profiling aid)")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="FP32-vs-FP64 error curve every 100 steps up to this step (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device hipMemcpy of MIB MiB (a plain copy "
                         "kernel of known size, to scale HBM counters) and exit")
    ap.add_argument("--ceiling", action="store_true",
                    help="only measure the bandwidth ceilings (streaming passes, the sweeps' access "
                         "pattern) beside the sweeps, and exit")
    ap.add_argument("--cpu-child", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the z-slab decomposition is a design, not code). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's child starts before this process touches the GPU
    cpu = None
    if not args.no_cpu_baseline and not args.ceiling and args.workload in ("precision_400", "drop_in_200", "lod_50"):
        cpu = CpuBaseline(args.workload)
    try:
        gpu_main(args, cpu)
    finally:
        if cpu is not None:
            cpu.cancel()


def gpu_main(args, cpu):
    from biofvm_amd import FP32, FP64, _capi
    from biofvm_amd import workloads as W
    _capi.load()
    if args.ceiling:
        ceiling_bench(args)
        return
    if args.workload == "ensemble_64":
        ensemble_bench(args)
        return
    if args.workload == "differentiation_128":
        differentiation_bench(args)
        return

    case = getattr(W, args.workload)()
    nvs = case["nx"] * case["ny"] * case["nz"] * case["S"]
    steps, warmup = args.steps, args.warmup
    # the CPU-runnable and drop-in workloads are FP64 by definition
    main_prec = FP32 if args.workload == "precision_400" else FP64
    w_main = 4 if main_prec == FP32 else 8

    m32, wall, ev_ms, per, n_done = time_precision(case, main_prec, steps, warmup, plan_x=args.plan_x)
    m32_chunk = m32.sweep_chunk
    plans = m32.sweep_plans()
    policies = m32.sweep_cache_policies()
    spreads = [m32.sweep_spread(ax) for ax in range(3)]
    value = nvs * steps / wall
    step_ms = 1e3 * wall / steps
    fused = sources_fused(case, per)
    dom = max(per, key=lambda k: per[k][0])
    dom_ms, dom_n = per[dom]  # per step
    dom_avg = dom_ms / dom_n  # per launch
    dom_bytes = kernel_bytes(case, dom, w_main, fused)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9  # one sweep's bytes over its time per step
    traffic = None
    if dom.startswith("sweep"):
        cols, L, streaming = plans["xyz".index(dom[-1])]
        kname = ("k_sweep_x_src" if dom == "sweep_x" and fused else
                 "k_sweep_rows" if streaming == 2 else "k_sweep_xt" if streaming == 3 else
                 "k_sweep_stream" if streaming else "k_sweep")
        tname = "float" if main_prec == FP32 else "double"
        # the row-tile kernel's template has no lines-per-block parameter, the
        # whole-cache-line x kernel's has the substrate count instead; the
        # column and whole-cache-line kernels carry a cache policy (2:
        # non-temporal)
        pol = 2 if policies["xyz".index(dom[-1])] else 0
        if kname == "k_sweep_rows":
            targs = f"{tname}, {L}"
        elif kname == "k_sweep_xt":
            targs = f"{tname}, {L}, {case['S']}, {pol}"
        elif kname == "k_sweep_x_src":
            targs = f"{tname}, {L}, {cols}"
        else:
            targs = f"{tname}, {L}, {cols}, {pol}"
        traffic = measured_traffic(args.workload, f"{kname}<{targs}>")
    step_bytes = algorithmic_bytes(case, w_main)

    out_extra = {}
    if main_prec == FP32 and not args.no_fp64:
        m64, wall64, ev64, per64, n64 = time_precision(case, FP64, steps, warmup)
        assert n64 == n_done
        a = m32.download()
        b = m64.download()
        plans64, policies64 = m64.sweep_plans(), m64.sweep_cache_policies()
        spreads64 = [m64.sweep_spread(ax) for ax in range(3)]
        m64.close()
        diff = np.abs(a - b)
        scale = np.abs(b)
        big = scale > 1e-3 * scale.max()
        out_extra["fp64"] = {"value": nvs * steps / wall64, "ms_per_step": 1e3 * wall64 / steps,
                             "kernels_ms_per_step": {k: v[0] for k, v in per64.items()},
                             "sweep_plans_lines_nodes_streaming": dict(zip("xyz", plans64)),
                             "sweep_non_temporal": dict(zip("xyz", policies64)),
                             "sweep_spread": dict(zip("xyz", spreads64)),
                             "lod_only": lod_only(per64, nvs, 8),
                             "hbm_frac_step": algorithmic_bytes(case, 8) / (wall64 / steps) / 1e9 / HBM_PEAK_GBS}
        out_extra["fp32_speedup_vs_fp64"] = wall64 / wall
        out_extra["error_vs_fp64"] = {
            "steps": n_done, "max_rel_normwise": float(diff.max() / scale.max()),
            "max_rel_elementwise": float((diff[big] / scale[big]).max()),
            "l2_rel": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
        del a, b
    m32.close()
    if fused:
        # the same step with bulk, agents and Dirichlet as kernels of their own
        msep, wsep, _, persep, _ = time_precision(case, main_prec, max(1, steps // 2), warmup, fused=False)
        msep.close()
        out_extra["separate_source_kernels"] = {
            "ms_per_step": 1e3 * wsep / max(1, steps // 2),
            "kernels_ms_per_step": {k: v[0] for k, v in persep.items()}}
    if main_prec == FP32 and not args.no_fp64 and args.error_steps > 0:
        out_extra["error_vs_fp64_curve"] = error_curve(case, args.error_steps, 100)

    line = {
        "metric": "voxel-substrate updates/s (full step: agents + LOD diffusion-decay)",
        "value": value,
        "unit": "voxel-substrate updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": step_ms,
        "events_ms_per_step": ev_ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the arithmetic type (every store precision solves in FP64 registers
        # with FP64 coefficients); the store's value type beside it
        "dtype": "f64",
        "store_dtype": "f32" if main_prec == FP32 else "f64",
        "data": "synthetic (seeded numpy default_rng; see biofvm_amd/workloads.py)",
        "config": {"workload": f"{args.workload}: {case['nx']}x{case['ny']}x{case['nz']} voxels x "
                               f"{case['S']} substrates, dx={case['dx']}, dt={case['dt']}, "
                               f"{0 if case.get('agent_volume') is None else len(case['agent_volume'])} agents, "
                               f"phases={case['phases']}"},
        "hbm_frac_step": step_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        # the same with every term at the store's width (agent constants too)
        "hbm_frac_step_store_width": survey_bytes(case, w_main) / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_step": dom_bytes, "ms_per_step": dom_ms, "launches_per_step": dom_n,
                     "avg_launch_ms": dom_avg},
        "sources_fused_in_x_sweep": fused,
        "kernels_ms_per_step": {k: v[0] for k, v in per.items()},
        "kernel_launches_per_step": {k: v[1] for k, v in per.items()},
        "sweep_xy_chunk_planes": m32_chunk,
        "lod_only": lod_only(per, nvs, w_main),
        "sweep_plans_lines_nodes_streaming": {"x": plans[0], "y": plans[1], "z": plans[2]},
        "sweep_non_temporal": {"x": policies[0], "y": policies[1], "z": policies[2]},
        "sweep_spread": dict(zip("xyz", spreads)),
    }
    line.update(out_extra)
    if args.workload == "drop_in_200":
        line["facade"] = facade_bench()
    if args.workload == "precision_400" and not args.no_forms:
        # the batching and adjoint forms, compact (their own --workload gives the full lines)
        print("[bench] ensemble_64", file=sys.stderr, flush=True)
        line["ensemble_64"] = ensemble_result(20, 5, False)
        print("[bench] differentiation_128", file=sys.stderr, flush=True)
        line["differentiation_128"] = differentiation_result(False)
        # the FP64 configurations (the CPU-runnable grid and the drop-in step)
        print("[bench] lod_50", file=sys.stderr, flush=True)
        line["lod_50"] = config_result("lod_50", 100, 10)
        print("[bench] drop_in_200", file=sys.stderr, flush=True)
        line["drop_in_200"] = config_result("drop_in_200", 50, 5)
    if cpu is not None:
        line["cpu_baseline"], one_core = cpu.result(case)
        if one_core:
            line["cpu_baseline_1core"] = one_core
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
