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


def _random_targets(m, n, count, rng):
    """`count` distinct n-photon occupation vectors of m modes, bunched ones included."""
    seen, out = set(), []
    while len(out) < count:
        occ = np.bincount(rng.randint(0, m, size=n), minlength=m)
        key = tuple(int(v) for v in occ)
        if key not in seen:
            seen.add(key)
            out.append(occ.astype(np.int32))
    return np.array(out, dtype=np.int32)


@pytest.mark.parametrize("m,inp", [(4, [0, 1, 0, 0]), (32, [2, 1, 1, 1, 1, 1, 1] + [0] * 25)])
def test_limits_bunched_and_long_program(gpu, m, inp):
    """One photon on 4 modes (two of its outputs as targets), and the limits: 32 modes,
    8 photons with a bunched input mode, 64 targets with bunched outputs, and a
    Clements program of 992 phases, too long for the kernel to stage in LDS (read
    from global memory instead).  Loss against the oracle, gradient against
    central differences of the oracle loss on a few phases."""
    from perceval_amd.gradient import MeshProgram, mesh_adjoint
    rng = np.random.RandomState(m)
    prog = MeshProgram(clements_mesh(m))
    phases = rng.uniform(0, 2 * math.pi, size=(3, prog.nparams))
    n = sum(inp)
    targets = oracle.fock_enumerate(m, n)[:2] if m == 4 else _random_targets(m, n, 64, rng)

    def loss_at(ph):
        U = prog.unitary(ph)
        return sum(abs(oracle.perm(oracle.submatrix(U, inp, t))) ** 2 /
                   (oracle.prodnfact(t) * oracle.prodnfact(inp)) for t in targets)

    loss, grad = mesh_adjoint(prog, phases, inp, targets)
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    h = 1e-6
    for b in range(3):
        assert abs(loss[b] - loss_at(phases[b])) < 1e-12
        for k in rng.choice(prog.nparams, 4, replace=False):
            pp = phases[b].copy(); pp[k] += h
            pm = phases[b].copy(); pm[k] -= h
            fd = (loss_at(pp) - loss_at(pm)) / (2 * h)
            assert abs(grad[b, k] - fd) < 1e-12 + 1e-6 * abs(fd), (b, k, grad[b, k], fd)


def test_target_blocks_and_split_plans(gpu):
    """512 bunched targets at 32 modes and 8 photons: their tables exceed the
    kernel's LDS in one plan (-4), so mesh_adjoint runs two plans of 256, each
    in 4 blocks of 64 targets.  Loss and gradient equal the sums over 8
    single-block plans of 64 targets (to rounding), and the loss the oracle's."""
    from perceval_amd.gradient import MeshPlan, MeshProgram, mesh_adjoint
    m, inp = 32, [2, 1, 1, 1, 1, 1, 1] + [0] * 25
    rng = np.random.RandomState(17)
    prog = MeshProgram(clements_mesh(m))
    phases = rng.uniform(0, 2 * math.pi, size=(2, prog.nparams))
    targets = _random_targets(m, sum(inp), 512, rng)
    loss, grad = mesh_adjoint(prog, phases, inp, targets)
    loss, grad = loss.cpu().numpy(), grad.cpu().numpy()
    assert sorted(k[1][0] for k in prog._plans) == [256, 256]
    ref_l, ref_g = 0.0, 0.0
    for t0 in range(0, 512, 64):
        plan = MeshPlan(prog, inp, targets[t0:t0 + 64])
        lp, gp = plan(phases)
        ref_l, ref_g = ref_l + lp.cpu().numpy(), ref_g + gp.cpu().numpy()
        plan.close()
    assert np.max(np.abs(loss - ref_l)) < 1e-13 * max(1.0, np.max(np.abs(ref_l)))
    assert np.max(np.abs(grad - ref_g)) < 1e-13 * max(1.0, np.max(np.abs(ref_g)))
    for b in range(2):
        U = prog.unitary(phases[b])
        want = sum(abs(oracle.perm(oracle.submatrix(U, inp, t))) ** 2 /
                   (oracle.prodnfact(t) * oracle.prodnfact(inp)) for t in targets)
        assert abs(loss[b] - want) < 1e-12


def test_one_shot_entry_matches_the_plan(gpu):
    """pcvl_mesh_adjoint (create, execute, synchronise, destroy in one call)
    against a MeshPlan executed twice on new phases: the same bits."""
    import ctypes
    import torch
    from perceval_amd import _native
    from perceval_amd.gradient import MeshPlan
    _, prog, phases, targets = _setup(batch=64, seed=9)
    plan = MeshPlan(prog, IN, targets)
    lib = _native.load()
    i32 = ctypes.POINTER(ctypes.c_int32)
    tg = np.ascontiguousarray(targets, dtype=np.int32)
    inp = np.ascontiguousarray(IN, dtype=np.int32)
    for shift in (0.0, 0.25):
        ph = torch.from_numpy(phases + shift).cuda()
        loss_p, grad_p, probs_p = plan(ph, want_probs=True)
        loss = torch.empty_like(loss_p)
        grad = torch.empty_like(grad_p)
        probs = torch.empty_like(probs_p)
        _native.check(lib.pcvl_mesh_adjoint(
            prog.kind.ctypes.data_as(i32), prog.mode.ctypes.data_as(i32), prog.param.ctypes.data_as(i32),
            prog.const.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), prog.nops, prog.m, _native.ptr(ph),
            prog.nparams, 64, inp.ctypes.data_as(i32), tg.ctypes.data_as(i32), len(tg), _native.ptr(loss),
            _native.ptr(grad), _native.ptr(probs), _native.stream_handle()), "pcvl_mesh_adjoint")
        assert torch.equal(loss, loss_p) and torch.equal(grad, grad_p) and torch.equal(probs, probs_p)
    plan.close()


def test_weighted_observable_over_every_output(gpu):
    """The objective of benchmark/benchmark_QML-DE-solver.py:61-84 on a phase
    mesh: 6 photons in 6 modes, every one of the 462 output states weighted by
    a seeded lambda (standardised and scaled by 200 as there), loss =
    sum_s lambda_s P(s).  The 462 targets run as one plan, in 8 blocks of at
    most 64 inside the kernel; the loss matches the oracle's SLOS
    probabilities, the gradient central
    differences of that loss (h = 1e-6), and the unweighted loss is the total
    probability 1."""
    import torch
    from perceval_amd.gradient import MeshProgram, mesh_adjoint, target_probability
    m = n = 6
    prog = MeshProgram(clements_mesh(m))
    rng = np.random.RandomState(0)
    targets = oracle.fock_enumerate(m, n)
    assert len(targets) == 462
    lam = rng.rand(len(targets))
    lam = 200 * (lam - lam.mean()) / lam.std()
    phases = rng.uniform(0, 2 * math.pi, size=(4, prog.nparams))
    inp = [1] * n
    loss, grad, probs = mesh_adjoint(prog, phases, inp, targets, want_probs=True, weights=lam)
    loss, grad, probs = loss.cpu().numpy(), grad.cpu().numpy(), probs.cpu().numpy()

    def f(ph):
        return float(np.dot(lam, oracle.slos_probs(prog.unitary(ph), inp)))

    for b in range(4):
        ref = oracle.slos_probs(prog.unitary(phases[b]), inp)
        assert np.max(np.abs(probs[b] - ref)) < 1e-13
        assert abs(loss[b] - f(phases[b])) < 1e-10
    h = 1e-6
    for b in (0, 3):
        fd = np.empty(prog.nparams)
        for k in range(prog.nparams):
            pp = phases[b].copy(); pp[k] += h
            pm = phases[b].copy(); pm[k] -= h
            fd[k] = (f(pp) - f(pm)) / (2 * h)
        assert np.max(np.abs(grad[b] - fd)) < 1e-6 * max(1.0, np.max(np.abs(fd))), np.max(np.abs(grad[b] - fd))
    ones, gones = mesh_adjoint(prog, phases, inp, targets)
    assert np.max(np.abs(ones.cpu().numpy() - 1.0)) < 1e-12
    assert np.max(np.abs(gones.cpu().numpy())) < 1e-11          # the total probability does not move
    x = torch.from_numpy(phases).cuda().requires_grad_(True)
    target_probability(x, prog, inp, targets, lam).sum().backward()
    assert np.max(np.abs(x.grad.cpu().numpy() - grad)) < 1e-12
