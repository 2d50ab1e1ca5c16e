"""dftb_trans_charges_noalloc (fixed device scratch, no workspace) against
dftb_trans_charges (caller workspace) through the C ABI: two calls in a row,
the second on a smaller molecule, bit-identical outputs (same kernels)."""
import ctypes

import numpy as np
import pytest

from conftest import ground_state

torch = pytest.importorskip("torch")


def _ref_and_noalloc(lib, gs, nocc_act=None, nvirt_act=None):
    from oracle import tddftb_oracle as O
    dev = torch.device("cuda")
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe, nocc_act, nvirt_act)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    Co, Cv, St = t(C[:, ao]), t(C[:, av]), t(S)
    nat, norb, o, v = len(gs.orbsPerAtom), S.shape[0], len(ao), len(av)
    opa = np.ascontiguousarray(gs.orbsPerAtom, dtype=np.int32)
    popa = opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    outs = []
    for noalloc in (False, True):
        # outputs prefilled with NaN: every element must be written
        q = [torch.full(s, float("nan"), dtype=torch.float64, device=dev) for s in ((nat, o, o), (nat, o, v),
                                                                                      (nat, v, v))]
        if noalloc:
            rc = lib.dftb_trans_charges_noalloc(p(Co), p(Cv), p(St), popa, nat, norb, o, v, p(q[0]), p(q[1]),
                                                p(q[2]), st)
        else:
            nb = lib.dftb_trans_charges_workspace(norb, o, v, nat)
            ws = torch.empty(nb, dtype=torch.uint8, device=dev)
            rc = lib.dftb_trans_charges(p(Co), p(Cv), p(St), popa, nat, norb, o, v, p(q[0]), p(q[1]), p(q[2]),
                                        p(ws), nb, st)
        assert rc == 0
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in q])
    ref = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    return outs[0], outs[1], ref


@pytest.mark.gpu
def test_trans_charges_noalloc_matches_workspace_entry_point():
    from dftbaby_amd import _lib
    lib = _lib.load()
    # first call: a larger molecule; second: a smaller one (the scratch keeps the
    # first call's contents, which the second must not depend on)
    for name, lc, no, nv in (("C20H22", 1, None, None), ("benzene", 1, 4, 5)):
        gs = ground_state(name, lc, 0.05, 7)
        print("noalloc check:", name, no, nv, flush=True)
        ws_out, na_out, ref = _ref_and_noalloc(lib, gs, no, nv)
        for a, b, r in zip(ws_out, na_out, ref):
            assert np.isfinite(b).all()
            np.testing.assert_array_equal(a, b)
            np.testing.assert_allclose(b, r, rtol=0, atol=1e-13)
