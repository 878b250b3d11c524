"""Pins of the proposal-network oracle (oracle/mlp.py), which checks the GPU
forward, backward and Adam kernels of the learned sampler.

Each function of the oracle is fixed here by something other than its own
formula: central finite differences of the loss, torch's CPU autograd and
torch.optim.Adam (independent library routines), the closed form of Adam's
first step, torch's bfloat16 cast, and a teacher-forcing case written out by
hand.  A plausible slip in the oracle -- a transposed weight, a dropped ReLU
mask, the mask bound off by one, a missing bias correction -- fails one of
them."""
import numpy as np
import pytest

from oracle import mlp

torch = pytest.importorskip("torch")


def _net(m, n, H, seed):
    rng = np.random.RandomState(seed)
    D = m + n
    W1 = rng.randn(H, D) * 0.4
    b1 = rng.randn(H) * 0.2
    W2 = rng.randn(m, H) * 0.3
    b2 = rng.randn(m) * 0.2
    return W1, b1, W2, b2


def _batch(m, n, B, seed):
    rng = np.random.RandomState(seed)
    seq = np.sort(rng.randint(0, m, size=(B, n)), axis=1).astype(np.int32)
    return seq, mlp.features(seq, m)


def test_features_hand_written_case():
    # two sorted mode sequences over m = 4 modes, n = 3 photons
    seq = np.array([[0, 0, 2], [1, 3, 3]], dtype=np.int32)
    X, lo, tgt = mlp.features(seq, 4)
    want = np.array([
        # counts of modes 0..3 so far | one-hot step 0..2
        [0, 0, 0, 0, 1, 0, 0],   # b0 p0
        [1, 0, 0, 0, 0, 1, 0],   # b0 p1 (mode 0 drawn)
        [2, 0, 0, 0, 0, 0, 1],   # b0 p2 (mode 0 twice)
        [0, 0, 0, 0, 1, 0, 0],   # b1 p0
        [0, 1, 0, 0, 0, 1, 0],   # b1 p1
        [0, 1, 0, 1, 0, 0, 1],   # b1 p2
    ], dtype=float)
    assert np.array_equal(X, want)
    assert lo.tolist() == [0, 0, 0, 0, 1, 3]
    assert tgt.tolist() == [0, 0, 2, 1, 3, 3]


@pytest.mark.parametrize("x", [1.0, -2.5, 3.0e38, -1.0e-39, 0.0, 1.00390625, 1.01171875, 1.0078125 + 2 ** -16,
                               65504.0, 2 ** -126, float("inf"), -float("inf")])
def test_bf16_round_special_values_vs_torch(x):
    got = mlp.bf16_round(np.array([x], dtype=np.float32))[0]
    ref = torch.tensor([x], dtype=torch.float32).to(torch.bfloat16).to(torch.float64).item()
    assert got == ref or (np.isinf(got) and np.isinf(ref) and np.sign(got) == np.sign(ref))


def test_bf16_round_random_values_vs_torch():
    rng = np.random.RandomState(0)
    # random bit patterns cover ties, carries into the exponent and every binade
    bits = rng.randint(0, 2 ** 32, size=200000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    got = mlp.bf16_round(x)
    ref = torch.from_numpy(x.copy()).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(got, ref)
    # exact ties round to even
    ties = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8], dtype=np.float32)
    assert mlp.bf16_round(ties).tolist() == [1.0, 1.0 + 2 ** -6]


def _loss_sum(W1, b1, W2, b2, X, lo, tgt, scale):
    loss, _ = mlp.loss_and_grads(W1, b1, W2, b2, X, lo, tgt, scale)
    return scale * loss.sum()


def test_loss_and_grads_vs_central_differences(monkeypatch):
    """With the bf16 rounding bypassed, the oracle's gradients are the
    derivatives of scale * sum(loss): central differences in float64."""
    monkeypatch.setattr(mlp, "bf16_round", lambda w: np.asarray(w, dtype=np.float64))
    m, n, H = 7, 3, 9
    W1, b1, W2, b2 = _net(m, n, H, seed=1)
    seq, (X, lo, tgt) = _batch(m, n, 5, seed=2)
    scale = 0.37
    _, grads = mlp.loss_and_grads(W1, b1, W2, b2, X, lo, tgt, scale)
    params = [W1, b1, W2, b2]
    h = 1e-6
    worst = 0.0
    for k, P in enumerate(params):
        fd = np.zeros_like(P)
        for idx in np.ndindex(P.shape):
            old = P[idx]
            P[idx] = old + h
            fp = _loss_sum(*params, X, lo, tgt, scale)
            P[idx] = old - h
            fm = _loss_sum(*params, X, lo, tgt, scale)
            P[idx] = old
            fd[idx] = (fp - fm) / (2 * h)
        err = np.max(np.abs(fd - grads[k]))
        worst = max(worst, err)
        assert err < 1e-7 * max(1.0, np.max(np.abs(fd))), (k, err)
    assert worst > 0.0   # the comparison is not vacuous


def test_loss_and_grads_vs_torch_autograd():
    """Same network in torch (float64, CPU autograd) on the bf16-rounded
    weights; the oracle's gradients are straight-through: the gradient with
    respect to the rounded copy, applied to the master weights."""
    m, n, H = 11, 4, 16
    W1, b1, W2, b2 = _net(m, n, H, seed=3)
    seq, (X, lo, tgt) = _batch(m, n, 6, seed=4)
    scale = 1.0 / 6
    loss, grads = mlp.loss_and_grads(W1, b1, W2, b2, X, lo, tgt, scale)

    tW1 = torch.tensor(mlp.bf16_round(W1), requires_grad=True)
    tb1 = torch.tensor(b1, requires_grad=True)
    tW2 = torch.tensor(mlp.bf16_round(W2), requires_grad=True)
    tb2 = torch.tensor(b2, requires_grad=True)
    tX = torch.tensor(X)
    Z = torch.relu(tX @ tW1.T + tb1) @ tW2.T + tb2
    col = torch.arange(m).unsqueeze(0)
    Zm = Z.masked_fill(col < torch.tensor(lo).unsqueeze(1), float("-inf"))
    ref_loss = -torch.log_softmax(Zm, dim=1).gather(1, torch.tensor(tgt).unsqueeze(1)).squeeze(1)
    (scale * ref_loss.sum()).backward()
    assert np.allclose(loss, ref_loss.detach().numpy(), rtol=1e-13, atol=1e-13)
    for g, t in zip(grads, (tW1, tb1, tW2, tb2)):
        assert np.allclose(g, t.grad.numpy(), rtol=1e-12, atol=1e-13)


def test_mask_excludes_modes_below_previous_draw():
    """A logit below the mask bound changes neither loss nor gradients."""
    m, n, H = 6, 3, 5
    W1, b1, W2, b2 = _net(m, n, H, seed=5)
    seq = np.array([[2, 3, 5], [1, 4, 4]], dtype=np.int32)
    X, lo, tgt = mlp.features(seq, m)
    base_loss, base_grads = mlp.loss_and_grads(W1, b1, W2, b2, X, lo, tgt)
    # raise mode 0's bias: every row with lo >= 1 must be unaffected
    b2x = b2.copy()
    b2x[0] += 5.0
    loss2, grads2 = mlp.loss_and_grads(W1, b1, W2, b2x, X, lo, tgt)
    rows = lo >= 1
    assert rows.sum() == 4
    assert np.allclose(loss2[rows], base_loss[rows], rtol=0, atol=1e-15)
    assert not np.allclose(loss2[~rows], base_loss[~rows])
    # the gradient of b2[0] gets contributions only from the unmasked rows
    assert abs(grads2[3][0]) > 0 and abs(base_grads[3][0]) > 0


def test_adam_first_step_closed_form():
    """At step 1 the bias-corrected moments are g and g^2, so every weight
    moves by lr * g / (|g| + eps): lr * sign(g) up to eps."""
    rng = np.random.RandomState(6)
    w = rng.randn(1000)
    g = rng.randn(1000)
    lr, eps = 1e-3, 1e-8
    w1, m1, m2 = mlp.adam(w, g, np.zeros_like(w), np.zeros_like(w), lr, 0.9, 0.999, eps, 1)
    assert np.allclose(m1, 0.1 * g, rtol=1e-12, atol=0)
    assert np.allclose(m2, 0.001 * g * g, rtol=1e-12, atol=0)
    assert np.allclose(w - w1, lr * g / (np.abs(g) + eps), rtol=1e-12, atol=0)
    assert np.allclose(w - w1, lr * np.sign(g), rtol=0, atol=lr * eps / np.min(np.abs(g)) * 1.01)


def test_adam_steps_vs_torch_optim():
    """Five steps with changing gradients against torch.optim.Adam (float64)."""
    rng = np.random.RandomState(7)
    w0 = rng.randn(300)
    tw = torch.tensor(w0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    w, m1, m2 = w0.copy(), np.zeros(300), np.zeros(300)
    for step in range(1, 6):
        g = rng.randn(300) * step
        w, m1, m2 = mlp.adam(w, g, m1, m2, 2e-3, 0.9, 0.999, 1e-8, step)
        tw.grad = torch.tensor(g)
        opt.step()
        assert np.allclose(w, tw.detach().numpy(), rtol=1e-13, atol=1e-15), step
