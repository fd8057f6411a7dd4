"""Runs the same seeded world through the HIP engine and the CPU checker."""
from __future__ import annotations

import numpy as np

from lenia_amd import workloads as WL


def run_pair(world: dict, steps: int, fields=("A",), keep_velocity=True, keep_usum=False,
             graph=True, prec=32, schedule=None):
    from oracle import oracle as O
    from lenia_amd.engine import Engine

    ow = O.World(world["shape"], world["C"], world["kernels"], world["cfg"], prec)
    ow.set(O.F_A, world["A"])
    if "V" in world:
        ow.set(O.F_V, world["V"])
    eng = Engine(keep_velocity=keep_velocity, keep_usum=keep_usum, graph=graph, schedule=schedule,
                 keep_mu="MU" in fields)
    wid = eng.add_world(world["shape"], world["C"], world["kernels"], world["cfg"])
    eng.commit()
    eng.set(wid, "A", world["A"])
    if "V" in world:
        eng.set(wid, "V", world["V"])
    ow.step(steps)
    eng.step(steps)
    eng.synchronize()
    omap = {"A": O.F_A, "GA": O.F_GA, "S": O.F_S, "V": O.F_V, "H": O.F_H, "B": O.F_B,
            "MU": O.F_MU, "USUM": O.F_USUM}
    out = {}
    for f in fields:
        out[f] = (eng.get(wid, f), ow.get(omap[f]).astype(np.float32))
    eng.close()
    return out


def mismatch(a: np.ndarray, b: np.ndarray):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    return int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32))), float(d.max() if d.size else 0.0)
