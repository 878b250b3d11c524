"""Adjoint phase gradients (differentiation workload: Clements rectangle on 12
modes, 132 phases, 6 photons in modes 0..5, batch 256, loss = summed
probability of 32 seeded target outputs).

Checked against (a) probabilities from the oracle's permanents and from the
SLOS backend on the same unitaries, (b) central finite differences (h = 1e-6)
of the oracle-computed loss, and (c) torch autograd wiring."""
import math

import numpy as np
import pytest

import oracle
from perceval_amd.components import clements_mesh

pytestmark = pytest.mark.gpu

M, NPH = 12, 6
IN = [1] * NPH + [0] * (M - NPH)


def _setup(batch=256, seed=4):
    from perceval_amd.gradient import MeshProgram
    mesh = clements_mesh(M)
    prog = MeshProgram(mesh)
    rng = np.random.RandomState(seed)
    phases = rng.uniform(0, 2 * math.pi, size=(batch, prog.nparams))
    states = oracle.fock_enumerate(M, NPH)
    targets = states[np.sort(rng.choice(len(states), 32, replace=False))]
    return mesh, prog, phases, targets


def _oracle_loss(prog, ph, targets):
    U = prog.unitary(ph)
    return sum(abs(oracle.perm(oracle.submatrix(U, IN, t))) ** 2 / oracle.prodnfact(t) for t in targets)


def test_program_matches_circuit_unitary(gpu):
    mesh, prog, phases, _ = _setup(batch=2)
    for p, v in zip(prog.params, phases[0]):
        p.set_value(v)
    assert np.allclose(prog.unitary(phases[0]), mesh.compute_unitary(), atol=1e-13)


def test_loss_and_probs_vs_oracle_and_slos(gpu):
    from perceval_amd.gradient import mesh_adjoint
    from perceval_amd.backends import SLOSBackend
    from perceval_amd.components import Unitary
    _, prog, phases, targets = _setup()
    loss, grad, probs = mesh_adjoint(prog, phases, IN, targets, want_probs=True)
    loss = loss.cpu().numpy()
    probs = probs.cpu().numpy()
    for b in range(0, 256, 17):
        assert abs(loss[b] - _oracle_loss(prog, phases[b], targets)) < 1e-13
    slos = SLOSBackend()
    for b in (0, 101, 255):
        slos.set_circuit(Unitary(prog.unitary(phases[b])))
        slos.set_input_state(IN)
        for t, p in zip(targets, probs[b]):
            assert abs(slos.probability(list(t)) - p) < 1e-13


def test_gradient_vs_central_differences(gpu):
    from perceval_amd.gradient import mesh_adjoint
    _, prog, phases, targets = _setup(batch=256)
    _, grad = mesh_adjoint(prog, phases, IN, targets)
    grad = grad.cpu().numpy()
    h = 1e-6
    for b in (0, 77, 200):
        fd = np.empty(prog.nparams)
        for k in range(prog.nparams):
            pp = phases[b].copy(); pp[k] += h
            pm = phases[b].copy(); pm[k] -= h
            fd[k] = (_oracle_loss(prog, pp, targets) - _oracle_loss(prog, pm, targets)) / (2 * h)
        assert np.max(np.abs(grad[b] - fd)) < 1e-9, np.max(np.abs(grad[b] - fd))


def test_autograd_function(gpu):
    import torch
    from perceval_amd.gradient import mesh_adjoint, target_probability
    _, prog, phases, targets = _setup(batch=8)
    x = torch.from_numpy(phases).cuda().requires_grad_(True)
    loss = target_probability(x, prog, IN, targets)
    w = torch.arange(1, 9, dtype=torch.float64, device="cuda")
    (loss * w).sum().backward()
    _, grad = mesh_adjoint(prog, phases, IN, targets)
    assert torch.allclose(x.grad, grad * w[:, None], rtol=0, atol=1e-15)
