"""Learned part: GPU MLP gradients vs a CPU reference, and the learned start on held-out molecules (-m gpu)."""
import numpy as np
import pytest

from conftest import ground_state
from oracle import mlp_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_mlp_forward_backward_match_cpu(rng):
    from dftbaby_amd.learned import PairMLP
    m = PairMLP(seed=3)
    X = rng.standard_normal((300, 6)).astype(np.float32)
    T = rng.standard_normal((300, 2)).astype(np.float32)
    w = np.array([1.0, 0.5], dtype=np.float32)
    Xd = torch.as_tensor(X, device="cuda")
    part = m.loss_and_grad(Xd, torch.as_tensor(T, device="cuda"), torch.as_tensor(w, device="cuda"))
    params = m.params.cpu().numpy().astype(np.float64)
    loss_ref, grad_ref = mlp_oracle.loss_and_grad(params, m.shapes, X.astype(np.float64), T.astype(np.float64),
                                                  w.astype(np.float64))
    assert abs(float(part.sum().cpu()) - loss_ref) < 1e-4 * max(1.0, abs(loss_ref))
    g = m.grads.cpu().numpy()
    assert np.abs(g - grad_ref).max() < 2e-4 * max(1.0, np.abs(grad_ref).max())
    Y = m.forward(Xd).cpu().numpy()
    np.testing.assert_allclose(Y, mlp_oracle.forward(params, m.shapes, X.astype(np.float64)), rtol=1e-4, atol=1e-4)


def test_learned_start_converges_to_plain_roots():
    """The learned initial space changes the trajectory, not the converged roots,
    and on a held-out C12H14 it converges in fewer iterations than the plain start."""
    from dftbaby_amd.engine import CasidaOperatorGPU, nonhermitian_davidson_gpu
    from dftbaby_amd.learned import PairMLP, learned_start, training_rows
    from oracle import tddftb_oracle as O

    def op_for(seed):
        gs = ground_state("C12H14", 1, 0.05, seed)
        f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
        _, _, ao, av = O.active_orbitals(f, orbe)
        om, _ = O.orbital_differences(orbe, f, ao, av)
        return CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], S, gs.orbsPerAtom, om, gs.getGamma(),
                                               gs.getGamma_lr(), 1)
    k = 6
    Xs, Ts = [], []
    for seed in range(30, 36):
        op = op_for(seed)
        w, C, XmY, XpY, st = nonhermitian_davidson_gpu(op, k, conv=1e-7, maxiter=100)
        X, T = training_rows(op, w, C, n_top=300, n_rand=600)
        Xs.append(X)
        Ts.append(T)
    model = PairMLP(seed=0)
    hist = model.fit(torch.cat(Xs), torch.cat(Ts), steps=400, batch=1024, lr=3e-3)
    assert hist[-1][1] < hist[0][1]
    op = op_for(77)
    w0, _, _, _, st0 = nonhermitian_davidson_gpu(op, k, conv=1e-7, maxiter=100)
    V, pom, wg = learned_start(op, model, k)
    w1, _, _, _, st1 = nonhermitian_davidson_gpu(op, k, conv=1e-7, maxiter=100, initial_space=V, precond_omega=pom)
    assert np.abs(w0.cpu().numpy() - w1.cpu().numpy()).max() < 1e-8
    print("plain iterations", st0.iterations, "learned iterations", st1.iterations)
    # what the trained network gains on this held-out geometry: fewer Davidson
    # iterations than the plain start (19 -> 15 measured)
    assert st1.iterations < st0.iterations
