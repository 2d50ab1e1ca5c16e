"""Project autotuner for the Casida matvec launch shapes.

For one operator and precision it times, with HIP events on the launch stream,
the exchange part of the matvec restricted to an atom slice (CasidaOperator
field ``atom_limit``) for every candidate (atoms per slice, expand tile, contract
tile), normalises by the atoms processed and keeps the fastest.  Results are
cached per (precision, shape bucket) in a JSON file so later operators of a
similar size reuse them without measuring.
"""
import json
import os

import numpy as np
import torch

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dftbaby_amd", "tuning.json")


def _bucket(op, precision):
    b = lambda x: int(2 ** np.round(np.log2(max(x, 1))))
    return "p%d_nat%d_o%d_v%d" % (precision, b(op.nat), b(op.nocc), b(op.nvirt))


def _load(path):
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return {}


def autotune(op, precision, nvec=16, atom_chunks=(8, 16, 32, 64), tiles=(64, 128), reps=2, cache=CACHE,
             use_cache=True, log=None):
    """Pick atom_chunk / tile_expand / tile_contract for op at this precision; store in op.tuning."""
    key = _bucket(op, precision)
    table = _load(cache) if use_cache else {}
    if use_cache and key in table:
        op.tuning[precision] = table[key]
        return table[key]
    op.prepare(precision)
    dev = op.device
    V = torch.zeros((nvec, op.opad, op.vpad), dtype=torch.float64, device=dev)
    V[:, :op.nocc, :op.nvirt] = torch.randn((nvec, op.nocc, op.nvirt), dtype=torch.float64, device=dev)
    V = V.reshape(nvec, -1)
    P = torch.empty_like(V)
    M = torch.empty_like(V)
    best = None
    results = []
    for ac in atom_chunks:
        ac = min(ac, op.nat)
        for te in tiles:
            for tc in tiles:
                op.matvec(V, precision, P, M, atom_chunk=ac, tiles=(te, tc), atom_limit=ac)
                ts = []
                for _ in range(reps):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    op.matvec(V, precision, P, M, atom_chunk=ac, tiles=(te, tc), atom_limit=ac)
                    e1.record()
                    e1.synchronize()
                    ts.append(e0.elapsed_time(e1))
                per_atom = float(np.min(ts)) / ac
                results.append((per_atom, ac, te, tc))
                if log:
                    log("autotune %s ac=%d tiles=(%d,%d): %.3f ms/atom" % (key, ac, te, tc, per_atom))
                if best is None or per_atom < best[0]:
                    best = (per_atom, ac, te, tc)
    cfg = {"atom_chunk": best[1], "tile_expand": best[2], "tile_contract": best[3], "ms_per_atom": best[0],
           "nvec": nvec}
    op.tuning[precision] = cfg
    if cache:
        table = _load(cache)
        table[key] = cfg
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as fh:
            json.dump(table, fh, indent=1)
    return cfg
