"""Writes sw_fct4_L8.json: the oracle's run of the single-instance
workload (icosahedral level 8, 655,362 cells; FP64 shallow-water RK4,
dt = 60 s, 4 FCT tracers, del2 mixing kappa2 = 1e3 and viscosity
visc2 = 1e5; the seeded state of seed 1) after 2 and after 1,000 steps.

For each checkpoint the file holds the SHA-256 of the float64 bytes of h, u
and the tracers (bit equality), their L2 and max norms, total mass
sum(h * areaCell) and tracer mass, and the values at 64 seeded indices of
each field.  The oracle (the CPU restatement of the reference's tasks)
takes about 1.3 s per step on one core: about 25 minutes in all.

    python tests/golden/make_sw_fct4_L8.py [--steps 1000]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mpas-o-flecsi-2.0_amd"))

LEVEL, N_TR, DT, SEED, VISC2, KAPPA2 = 8, 4, 60.0, 1, 1e5, 1e3
N_SAMPLE = 64


def summary(m, h, u, tr):
    rng = np.random.default_rng(2024)
    area = m["areaCell"]
    out = {}
    for name, x in (("h", h), ("u", u), ("tracers", tr)):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        idx = np.sort(rng.choice(x.size, N_SAMPLE, replace=False))
        out[name] = {"sha256": hashlib.sha256(x.tobytes()).hexdigest(),
                     "l2": float(np.sqrt(np.sum(x * x))), "max": float(np.abs(x).max()),
                     "index": idx.tolist(), "value": x[idx].tolist()}
    out["mass"] = float(np.sum(h * area))
    out["tracer_mass"] = np.sum(tr.reshape(-1, N_TR) * (h * area)[:, None], axis=0).tolist()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "sw_fct4_L8.json"))
    args = ap.parse_args()
    from mpas_hip import Mesh, sw_workload_state
    from oracle.oracle import OracleMesh, OracleSW
    m = Mesh(LEVEL)
    u, h, tr = sw_workload_state(m, N_TR, seed=SEED)
    h_s = np.full(m.nCells, 4500.0)
    o = OracleSW(OracleMesh(m), nT=N_TR, tracer_adv=2, visc2=VISC2, kappa2=KAPPA2)
    o.set("h", h)
    o.set("u", u)
    o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    doc = {"workload": {"level": LEVEL, "cells": m.nCells, "tracers": N_TR, "dt": DT,
                        "seed": SEED, "visc2": VISC2, "kappa2": KAPPA2, "h_s": 4500.0},
           "initial": summary(m, h, u, tr), "checkpoints": {}}
    done, t0 = 0, time.time()
    for target in sorted({2, args.steps}):
        o.run(target - done, DT)
        done = target
        doc["checkpoints"][str(done)] = summary(m, o.get("h"), o.get("u"), o.get("tracers"))
        print(f"step {done}: {time.time() - t0:.0f} s", flush=True)
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
