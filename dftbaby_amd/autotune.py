"""Project autotuner for the Casida matvec launch shapes.

For one operator and precision it times candidate launch shapes with HIP events
on the launch stream and keeps the fastest in ``op.tuning[precision]``, which
the operator passes to every later matvec (CasidaOperator fields ``atom_chunk``,
``tile_expand``, ``tile_contract``):

* AO-basis exchange in BF16x3 (the default exchange of the mixed modes): the
  pipelined tile of the first two GEMMs of each exchange chain (B1 and the
  Hadamard GEMM, ``tile_expand``) and of the last two (B3 and the output GEMM,
  ``tile_contract``), each 256 x 128 (``256``) or 256 x 256 (``512``), timed on
  full matvecs of a 2k-vector block (the Davidson's expansion block);
* transition-charge exchange: atoms per slice and the expand / contract tiles,
  timed on the exchange restricted to an atom slice (``atom_limit``) and
  normalised per atom.

Results are cached per (precision, exchange, shape bucket) in a JSON file so
later operators of a similar size reuse them without measuring.
"""
import json
import os

import numpy as np
import torch

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dftbaby_amd", "tuning.json")


def _bucket(op, precision):
    b = lambda x: int(2 ** np.round(np.log2(max(x, 1))))
    return "p%d_%s_nat%d_o%d_v%d" % (precision, op.exchange, b(op.nat), b(op.nocc), b(op.nvirt))


def _time(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.min(ts))


def _random_block(op, nvec):
    V = torch.zeros((nvec, op.opad, op.vpad), dtype=torch.float64, device=op.device)
    V[:, :op.nocc, :op.nvirt] = torch.randn((nvec, op.nocc, op.nvirt), dtype=torch.float64, device=op.device)
    return V.reshape(nvec, -1)


def _tune_ao(op, precision, nvec, reps, log, key):
    V = _random_block(op, nvec)
    P = torch.empty_like(V)
    M = torch.empty_like(V)
    best = None
    for te in (256, 512):
        for tc in (256, 512):
            ms = _time(lambda: op.matvec(V, precision, P, M, tiles=(te, tc)), reps)
            if log:
                log("autotune %s tiles=(%d,%d): %.3f ms per %d vectors" % (key, te, tc, ms, nvec))
            if best is None or ms < best[0]:
                best = (ms, te, tc)
    return {"atom_chunk": 0, "tile_expand": best[1], "tile_contract": best[2], "ms_per_block": best[0],
            "nvec": nvec}


def _load(path):
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return {}


def autotune(op, precision, nvec=40, atom_chunks=(8, 16, 32, 64), tiles=(64, 128), reps=2, cache=CACHE,
             use_cache=True, log=None):
    """Pick the launch shapes of op's matvec at this precision (0 fp64, 1 fp32,
    2 bf16x3); store them in op.tuning (used by every later matvec) and the cache."""
    key = _bucket(op, precision)
    table = _load(cache) if use_cache else {}
    if use_cache and key in table:
        op.tuning[precision] = table[key]
        return table[key]
    op.prepare(precision)
    if op.lc and op.exchange == "ao":
        if precision != 2:
            return {}                  # FP64 / FP32 AO chains take the automatic tiles
        cfg = _tune_ao(op, precision, nvec, reps, log, key)
        return _store(op, precision, cfg, cache, key)
    if not op.lc:
        return {}                      # no exchange products to tune
    V = _random_block(op, nvec)
    P = torch.empty_like(V)
    M = torch.empty_like(V)
    best = None
    for ac in atom_chunks:
        ac = min(ac, op.nat)
        for te in tiles:
            for tc in tiles:
                ms = _time(lambda: op.matvec(V, precision, P, M, atom_chunk=ac, tiles=(te, tc), atom_limit=ac),
                           reps)
                per_atom = ms / ac
                if log:
                    log("autotune %s ac=%d tiles=(%d,%d): %.3f ms/atom" % (key, ac, te, tc, per_atom))
                if best is None or per_atom < best[0]:
                    best = (per_atom, ac, te, tc)
    cfg = {"atom_chunk": best[1], "tile_expand": best[2], "tile_contract": best[3], "ms_per_atom": best[0],
           "nvec": nvec}
    return _store(op, precision, cfg, cache, key)


def _store(op, precision, cfg, cache, key):
    op.tuning[precision] = cfg
    if cache:
        table = _load(cache)
        table[key] = cfg
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            with open(cache, "w") as fh:
                json.dump(table, fh, indent=1)
        except OSError:
            pass                       # an unwritable cache only costs a re-measurement next time
    return cfg
