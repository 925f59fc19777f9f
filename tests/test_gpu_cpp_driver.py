"""The C++ task layer and its `sw` executable on the device: the reference's
own CTest (apps/sw/CMakeLists.txt:28-31: test case 1, 5 steps of 40 s on a
10,242-cell sphere, L2 gate 0.0015 in apps/sw/finalize.cc:24-25) through the
23-task-per-step driver and through the fused step, both checked against
the oracle."""
import hashlib
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

SW = os.path.join(ROOT, "mpas-o-flecsi-2.0_amd", "bin", "sw")


def _fnv1a(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a, dtype=np.float64).tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


def _sw(*args):
    r = subprocess.run([SW, *args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout.strip(), r.stderr


def test_reference_ctest_case1_both_drivers():
    rc, out, err = _sw("--level", "5", "--unrotated", "-t", "1", "-s", "5", "-d", "40")
    assert rc == 0, err
    tasks = json.loads(out)
    rc, out, err = _sw("--level", "5", "--unrotated", "-t", "1", "-s", "5", "-d", "40", "--fused")
    assert rc == 0, err
    fused = json.loads(out)
    assert tasks["cells"] == 10242
    assert 0 < tasks["l2_error"] <= 0.0015
    assert tasks["h_fnv1a"] == fused["h_fnv1a"]
    assert tasks["l2_error"] == fused["l2_error"]

    # the same run on the oracle
    from mpas_hip import Mesh
    from oracle.oracle import OracleMesh, OracleSW
    m = Mesh(5, radius=6371000.0, rotate=False, omega=7.29212e-5)
    u, h = m.sw_case1(a=6371000.0)
    o = OracleSW(OracleMesh(m), freeze_u=True)
    o.set("h", h)
    o.set("u", u)
    o.set("h_s", np.full(m.nCells, 4500.0))
    o.init_diagnostics()
    o.run(5, 40.0)
    assert _fnv1a(o.get("h")) == tasks["h_fnv1a"]
    assert abs(o.compute_error() - tasks["l2_error"]) <= 1e-12 * tasks["l2_error"]


def test_case5_with_tracers_tasks_equal_fused():
    args = ["--level", "4", "-t", "5", "-s", "3", "-d", "600", "--tracers", "2"]
    rc, out, err = _sw(*args)
    assert rc == 0, err
    rc2, out2, err2 = _sw(*args, "--fused")
    assert rc2 == 0, err2
    assert json.loads(out)["h_fnv1a"] == json.loads(out2)["h_fnv1a"]


def test_gate_failure_is_reported():
    """A run far beyond the gate's 5 steps must fail it like the reference
    (its finalize throws): case 1 on a coarse mesh for a day."""
    rc, out, err = _sw("--level", "3", "--unrotated", "-t", "1", "-s", "200", "-d", "432")
    assert rc == 1 and "exceeds" in err
