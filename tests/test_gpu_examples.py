"""The reference's own example programs, compiled unmodified against the façade.

biofvm_amd/facade/examples/Makefile builds every examples/*.cpp of the
reference against the C++ façade (their `#include "../BioFVM.h"` finds the
façade's header); oracle/Makefile builds the same sources against the
reference's own objects (oracle/_ref/examples).  Both builds run with the
same arguments in separate directories and every MAT file they write is
compared: mesh columns (centres, volumes) bit for bit, densities within
1e-12 of the file's largest density (FP64 store).  The programs call
omp_set_num_threads(8); OMP_THREAD_LIMIT=1 caps the reference at one thread,
whose agent loop is then serial (with more threads it races when agents
share a voxel).

- convergence tests 1, 1_analytical, 2, 3 and 4_1 (examples/convergence_test*.cpp:
  1-D cosine, bulk supply strip, hex-lattice uptake agents, agents at voxel
  centres) at dx = 20 um, run live beside the compiled reference, and the
  agent listings conv 3 and 4_1 write;
- tutorials 1-3 (agents from rand(), bulk callbacks, Dirichlet strip; 500
  steps on 100^3 voxels) against samples of the compiled reference's outputs
  (tests/golden/gen_examples.py): every 997th voxel's column and every row's
  sum over all voxels;
- convergence_test5 seeds its agents from std::random_device, so only its
  façade build's run is checked (it finishes, every output finite).
The convergence_test4_2 input file and the performance programs' timing
outputs have nothing to compare and are not run.

Also: the C library's rand() sequence of a program is the same with and
without calls into the library in between (the HIP runtime seeds that
generator from the clock when it loads a code object; tutorial 1 draws its
agent positions from rand() after its first device use).
"""
import glob
import json
import math
import os
import subprocess

import numpy as np
import pytest

from biofvm_amd.matlab import read_matlab

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "biofvm_amd", "facade", "examples", "_build")
REF = os.path.join(ROOT, "oracle", "_ref", "examples")
GOLDEN = os.path.join(ROOT, "tests", "golden", "examples")

LIVE = [("convergence_test1", ["0.01", "2", "20"]),
        ("convergence_test1_analytical", ["2", "20"]),
        ("convergence_test2", ["0.01", "1", "20"]),
        ("convergence_test3", ["0.01", "1", "20"]),
        ("convergence_test4_1", ["0.01", "1", "20"])]


def run(exe, args, cwd, threads=None, timeout=300):
    env = dict(os.environ)
    if threads:
        env["OMP_THREAD_LIMIT"] = str(threads)
    os.makedirs(cwd, exist_ok=True)
    r = subprocess.run([exe, *args], cwd=cwd, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{exe}: {r.stdout[-3000:]} {r.stderr[-3000:]}"
    return r.stdout


def compare_mat(got_path, ref_path):
    name_g, G = read_matlab(got_path)
    name_r, R = read_matlab(ref_path)
    assert name_g == name_r and G.shape == R.shape, (got_path, name_g, name_r, G.shape, R.shape)
    np.testing.assert_array_equal(G[:4], R[:4])  # centres and volumes
    scale = max(float(np.max(np.abs(R[4:]))), 1e-300)
    err = float(np.max(np.abs(G[4:] - R[4:]))) / scale
    assert err <= 1e-12, (os.path.basename(got_path), err)


def built(prog):
    return os.path.exists(os.path.join(FACADE, prog))


@pytest.mark.parametrize("prog,args", LIVE, ids=[p for p, _ in LIVE])
def test_convergence_example_matches_compiled_reference(tmp_path, prog, args):
    if not os.path.exists(os.path.join(REF, prog)):
        pytest.skip("compiled reference examples not built (make -C oracle examples)")
    assert built(prog), "façade examples not built (make -C biofvm_amd/facade/examples)"
    run(os.path.join(REF, prog), args, str(tmp_path / "ref"), threads=1)
    run(os.path.join(FACADE, prog), args, str(tmp_path / "gpu"))
    ref_mats = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / "ref" / "*.mat")))
    got_mats = sorted(os.path.basename(p) for p in glob.glob(str(tmp_path / "gpu" / "*.mat")))
    assert ref_mats and got_mats == ref_mats
    for f in ref_mats:
        compare_mat(str(tmp_path / "gpu" / f), str(tmp_path / "ref" / f))
    for txt in glob.glob(str(tmp_path / "ref" / "cells_*.txt")):  # the agent listings
        with open(txt) as a, open(str(tmp_path / "gpu" / os.path.basename(txt))) as b:
            assert a.read() == b.read()


@pytest.mark.parametrize("prog", ["tutorial1_BioFVM", "tutorial2_BioFVM", "tutorial3_BioFVM"])
def test_tutorial_matches_reference_samples(tmp_path, prog):
    assert built(prog), "façade examples not built (make -C biofvm_amd/facade/examples)"
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        manifest = json.load(f)
    stride = manifest["stride"]
    spec = manifest["programs"][prog]
    gold = np.load(os.path.join(GOLDEN, f"{prog}.npz"))
    run(os.path.join(FACADE, prog), spec["args"], str(tmp_path))
    for fname, info in spec["files"].items():
        _, M = read_matlab(str(tmp_path / fname))
        assert list(M.shape) == info["shape"]
        key = fname.replace(".mat", "")
        sample, rowsum = gold[f"{key}_sample"], gold[f"{key}_rowsum"]
        got = M[:, ::stride]
        np.testing.assert_array_equal(got[:4], sample[:4])
        scale = max(float(np.max(np.abs(sample[4:]))), 1e-300)
        assert float(np.max(np.abs(got[4:] - sample[4:]))) / scale <= 1e-12, fname
        sums = np.array([math.fsum(row) for row in M])
        np.testing.assert_array_equal(sums[:4], rowsum[:4])
        assert np.all(np.abs(sums[4:] - rowsum[4:]) <= 1e-12 * np.abs(rowsum[4:])), (fname, sums, rowsum)


def test_convergence_test5_runs(tmp_path):
    # agents from std::random_device: no reference run to compare with
    prog = "convergence_test5"
    assert built(prog)
    run(os.path.join(FACADE, prog), ["0.01", "1", "20"], str(tmp_path))
    mats = glob.glob(str(tmp_path / "*.mat"))
    assert len(mats) >= 2
    for p in mats:
        _, M = read_matlab(p)
        assert np.all(np.isfinite(M))


def test_library_calls_leave_the_programs_rand_sequence_alone():
    # in a fresh process, so that the HIP runtime loads the library's code
    # objects during the calls (it seeds the C library's generator then)
    import sys
    code = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from biofvm_amd import ALL, FP64, DeviceMicroenvironment, load
load()  # the library (and the torch it loads beside it) before the sequence starts
libc = ctypes.CDLL(None)
libc.srand(7)
want = [libc.rand() for _ in range(6)]
libc.srand(7)
got = [libc.rand() for _ in range(3)]
with DeviceMicroenvironment(12, 10, 8, 2, 10.0, FP64) as m:
    m.set_coefficients([1e3, 1e4], [0.1, 0.5])
    m.upload(np.ones(12 * 10 * 8 * 2))
    m.set_bulk(np.full(1920, 0.1), np.ones(1920), np.full(1920, 0.05))
    m.run(0.01, 3, ALL)
    m.mass_balance()
    m.download()
got += [libc.rand() for _ in range(3)]
print("OK" if got == want else f"MISMATCH {got} {want}")
"""
    r = subprocess.run([sys.executable, "-c", code, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "OK", r.stdout
