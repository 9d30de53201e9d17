"""Re-measures the per-level launch-configuration table of csrc/tune.cpp on the current GPU.

For each workload (3D Laplacian 64^3, 12 levels, tol 1e-2; with ``--all`` also the
variable-kappa diffusion 96^3, 13 levels, tol 1e-3) and each factor precision, every knob
of the table is swept with the others at the table's choice (environment overrides:
SPAND_RRQR_REG, SPAND_RRQR_STEPPED_MIN_ELEMS, SPAND_RRQR_STEP_WAVES, SPAND_TRSM_CW), and the
per-level device time of repeated factorizations (level statistics column 11, the best of
two after one warmup) is written with the level shapes to a JSON file.

Usage: python -m spand_public_amd.tune_sweep [--all] [--out sweep.json]
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from . import FP32, FP64, Tree
from .problems import DiffusionFV, gaussian_field, linspace_nd, neglapl, symmetric_graph

KNOBS = {"SPAND_RRQR_REG": [0, 1], "SPAND_RRQR_STEPPED_MIN_ELEMS": [4096, 10000, 24000],
         "SPAND_RRQR_STEP_WAVES": [0, 4, 8, 16], "SPAND_TRSM_CW": [0, 1, 2]}


def _tree(A, X, nl, tol, prec):
    t = Tree(nl)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.set_tol(tol)
    t.set_precision(prec)
    t.partition(symmetric_graph(A))
    t.assemble(A)
    return t


def workloads(include_96: bool):
    yield "laplace64", (lambda: (neglapl(64, 3), linspace_nd(64, 3))), 12, 1e-2
    if include_96:
        def kappa96():
            fv = DiffusionFV(96, 3)
            return fv.matrix(np.exp(gaussian_field(96, 3, 2.0, 8.0, 2019))), linspace_nd(96, 3)
        yield "kappa96", kappa96, 13, 1e-3


def sweep(include_96: bool = False, reps: int = 3):
    out = []
    for name, make, nl, tol in workloads(include_96):
        A, X = make()
        for prec in (FP32, FP64):
            t = _tree(A, X, nl, tol, prec)
            t.factorize()
            for knob, vals in KNOBS.items():
                for v in vals:
                    for k in KNOBS:
                        os.environ.pop(k, None)
                    os.environ[knob] = str(v)
                    times = []
                    for rep in range(reps):
                        t.factorize()
                        if rep > 0:
                            times.append(t.level_stats()[:, 11].tolist())
                    per_level = np.min(np.array(times), axis=0).tolist()
                    out.append({"work": name, "prec": prec, "knob": knob, "value": v, "level_s": per_level,
                                "shape": t.level_tune()})
                    print(name, "fp32" if prec == FP32 else "fp64", knob, v, f"{1e3 * sum(per_level):.2f} ms",
                          flush=True)
            for k in KNOBS:
                os.environ.pop(k, None)
            del t
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true", help="also the 96^3 variable-kappa workload")
    ap.add_argument("--out", default="tune_sweep.json")
    a = ap.parse_args()
    json.dump(sweep(a.all), open(a.out, "w"))
