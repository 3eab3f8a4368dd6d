"""MultiCellDS snapshots of the façade (save_microenvironment_to_MultiCellDS_xml_pugi,
BioFVM_MultiCellDS.cpp:394-720) against the compiled reference's, byte for byte:
the XML document of a first call and of its update (new time, run time, data),
the densities' MAT file and the mesh MAT file.

The CSV-data and mesh-as-MAT forms need no GPU (no step runs, the densities
stay in the host mirror); the default form, whose MAT file is packed on the
GPU after a few steps, is marked gpu.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from biofvm_amd import FP64
from biofvm_amd import workloads as W

from test_gpu_facade import DRIVER

pytestmark = pytest.mark.skipif(not (oracle.have_ref() and os.path.exists(DRIVER)),
                                reason="compiled reference or façade driver not built")


def _case(steps, phases):
    rng = np.random.default_rng(44)
    nx, ny, nz, S = 5, 4, 3, 3
    case = W.make_case(nx, ny, nz, S, 20.0, [1e5, 2e3, 7.5e2], [0.1, 0.25, 1e-3], 0.01, steps, phases,
                       rng.uniform(0, 1, nx * ny * nz * S))
    case.update(W.random_agents(rng, 6, S, nx, ny, nz, 20.0))
    return case


def _files(base, flags):
    names = [base + "_a.xml", base + "_b.xml"]
    if "x" not in flags:
        names += [base + "_a_microenvironment0.mat", base + "_b_microenvironment0.mat"]
    if "m" in flags:
        names += [base + "_a_mesh0.mat"]
    return names


def _snapshots(tmp_path, case, flags):
    base = str(tmp_path / "snap")
    cin, cout = tmp_path / "case.bin", tmp_path / "out.bin"
    W.write_driver_case(cin, case, FP64)
    r = subprocess.run([DRIVER, str(cin), str(cout), "-", "-", base, flags], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rho = np.fromfile(cout, dtype="<f8")
    got = {}
    for f in _files(base, flags):
        with open(f, "rb") as h:
            got[f] = h.read()
        os.remove(f)
    # the reference on the densities the façade ended with, same base names
    t = case["steps"] * case["dt"]
    args = (case["nx"], case["ny"], case["nz"], case["S"], case["dx"], case["D"], case["lam"], rho)
    kw = dict(mesh_as_matlab="m" in flags, data_as_matlab="x" not in flags)
    oracle.ref_write_multicellds(*args, base + "_a", current_time=t, current_runtime=1.25, fresh=True, **kw)
    oracle.ref_write_multicellds(*args, base + "_b", current_time=t + 1.0, current_runtime=2.5, fresh=False, **kw)
    for f in _files(base, flags):
        with open(f, "rb") as h:
            assert h.read() == got[f], f


@pytest.mark.parametrize("flags", ["x", "mx"], ids=["csv_data", "mesh_mat_csv_data"])
def test_multicellds_document_without_gpu(tmp_path, flags):
    _snapshots(tmp_path, _case(0, 0), flags)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", ["-", "m"], ids=["mat_data", "mesh_mat"])
def test_multicellds_document_with_gpu_packed_data(tmp_path, flags):
    _snapshots(tmp_path, _case(3, W.ALL), flags)
