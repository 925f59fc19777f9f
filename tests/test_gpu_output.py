"""The run's output file (apps/sw/run.cc:73-113) written by the fused Python
driver (mpas_hip.output.run_sw) and by the `sw` executable in both modes
(23 tasks per step and fused): rows hold the state after exactly the steps
output_flagger names, bit for bit, and the two C++ modes write identical
files."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, bitwise_equal, mesh

pytestmark = [pytest.mark.gpu]

SW = os.path.join(ROOT, "mpas-o-flecsi-2.0_amd", "bin", "sw")
FIELDS = ("h", "h_s", "vorticity", "potential_vorticity", "tracers")


def _io_built():
    from mpas_hip.meshfile import IO_LIB
    return os.path.exists(IO_LIB)


needs_io = pytest.mark.skipif(not _io_built(), reason="libmpas_io.so not built (no HDF5)")


@needs_io
def test_fused_driver_rows_equal_direct_runs(tmp_path):
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    from mpas_hip.output import OutputFile, output_steps, read_output, run_sw
    m = mesh(4)
    nT, nsteps, T, dt = 2, 12, 5, 300.0
    u, h, tr = sw_workload_state(m, nT, seed=3)
    hs = np.full(m.nCells, 4500.0)
    c = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    p = str(tmp_path / "output.ncdf")
    c.sw_upload(h, u, tr.ravel(), hs)
    with OutputFile(p, m, T, nLevels=1, nTracers=nT) as f:
        run_sw(c, dt, nsteps, f)
    final = c.sw_download()
    steps = output_steps(nsteps, T)
    assert steps == [2, 5, 8, 11]
    # the same states by direct runs
    c.sw_upload(h, u, tr.ravel(), hs)
    want, done = [], 0
    for i in steps:
        c.rk4_step(dt, i - done)
        done = i
        want.append({"h": c.sw_field("h"), "vorticity": c.sw_field("vorticity"),
                     "potential_vorticity": c.sw_field("pv_vertex"),
                     "tracers": c.sw_field("tracers")})
    c.rk4_step(dt, nsteps - done)
    for a, b in zip(final, c.sw_download()):  # run_sw also ran the steps after the last row
        assert bitwise_equal(a, b)
    c.close()
    got = {k: read_output(p, k) for k in FIELDS}
    assert got["tracers"].shape == (T, m.nCells, 1, nT)
    assert bitwise_equal(got["h"][0].ravel(), h) and bitwise_equal(got["h_s"][0], hs)
    assert bitwise_equal(got["tracers"][0].ravel(), tr.ravel())
    assert not got["vorticity"][0].any() and not got["potential_vorticity"][0].any()
    for t, w in enumerate(want, start=1):
        for k, v in w.items():
            assert bitwise_equal(got[k][t].ravel(), v), (k, t)
        assert bitwise_equal(got["h_s"][t], hs)


def _sw(*args):
    r = subprocess.run([SW, *args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout.strip(), r.stderr


@needs_io
def test_sw_executable_tasks_and_fused_write_the_same_file(tmp_path):
    from mpas_hip.output import output_steps, read_output, scale_name
    args = ["--level", "4", "-t", "5", "-s", "3", "-d", "600", "--tracers", "2",
            "--num_output", "5"]
    pt, pf = str(tmp_path / "tasks.ncdf"), str(tmp_path / "fused.ncdf")
    rc, out_t, err = _sw(*args, "--output", pt)
    assert rc == 0, err
    rc, out_f, err = _sw(*args, "--fused", "--output", pf)
    assert rc == 0, err
    assert output_steps(3, 5) == [1, 2, 3]  # rows 0..3 written, row 4 never
    for k in FIELDS:
        a, b = read_output(pt, k), read_output(pf, k)
        assert a.shape[0] == 5 and bitwise_equal(a, b), k
        assert not a[4].any(), k
        assert scale_name(pf, k, 0) == "Time"
    # row 3 holds the final h the run reports
    from test_gpu_cpp_driver import _fnv1a
    assert _fnv1a(read_output(pf, "h")[3].ravel()) == json.loads(out_f)["h_fnv1a"]
    for name in ("areaCell", "latCell", "dvEdge", "cellsOnVertex", "mDensity"):
        assert bitwise_equal(read_output(pt, name), read_output(pf, name)), name
