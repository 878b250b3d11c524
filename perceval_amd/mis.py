"""Learned Metropolis independence sampler (MIS) for boson sampling.

An autoregressive proposal network (ProposalNet: bf16 weights, fp32
arithmetic, trained on the GPU with this library's own forward, backward and
Adam kernels) proposes output states as non-decreasing mode sequences, so the
proposal probability q(state) is exact.  Target probabilities p(state) come
from FP64 permanents (the batched Glynn kernel), and the independence
Metropolis chain accepts a proposal y from state x with probability
min(1, p(y) q(x) / (p(x) q(y))): its stationary law is exactly the
boson-sampling law, whatever the network learned.  Plain Clifford-Clifford
sampling (perceval_amd.sampling) provides the training data, the chain's exact
starting state, and the reference the chain is judged against.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from .permanent import PermanentPlan


def _lib():
    return _native.load()


def _p(t):
    return _native.ptr(t)


def _s():
    return _native.stream_handle()


class ProposalNet:
    """x (counts so far | one-hot step) -> relu(x W1^T + b1) -> W2^T -> logits over modes."""

    def __init__(self, m: int, n: int, hidden: int = 256, seed: int = 0, device: str = "cuda"):
        import torch
        self.m, self.n, self.H, self.D = m, n, hidden, m + n
        g = torch.Generator().manual_seed(seed)
        W1 = torch.randn(hidden, self.D, generator=g, dtype=torch.float64) * math.sqrt(2.0 / self.D)
        W2 = torch.randn(m, hidden, generator=g, dtype=torch.float64) * (0.1 / math.sqrt(hidden))
        self.W1 = W1.float().to(device)
        self.b1 = torch.zeros(hidden, dtype=torch.float32, device=device)
        self.W2 = W2.float().to(device)
        self.b2 = torch.zeros(m, dtype=torch.float32, device=device)
        self.params = [self.W1, self.b1, self.W2, self.b2]
        self.W1b = torch.empty(self.W1.shape, dtype=torch.bfloat16, device=device)
        self.W2b = torch.empty(self.W2.shape, dtype=torch.bfloat16, device=device)
        self._m1 = [torch.zeros_like(p) for p in self.params]
        self._m2 = [torch.zeros_like(p) for p in self.params]
        self._step = 0
        self.refresh_bf16()

    def refresh_bf16(self):
        lib = _lib()
        _native.check(lib.pcvl_mlp_to_bf16(_p(self.W1), _p(self.W1b), self.W1.numel(), _s()), "to_bf16")
        _native.check(lib.pcvl_mlp_to_bf16(_p(self.W2), _p(self.W2b), self.W2.numel(), _s()), "to_bf16")

    # ----------------------------------------------------------------- forward
    def forward(self, x):
        """x: (R, D) fp32 CUDA -> (h (R, H), logits (R, m))."""
        import torch
        lib = _lib()
        R = x.shape[0]
        h = torch.empty((R, self.H), dtype=torch.float32, device=x.device)
        z = torch.empty((R, self.m), dtype=torch.float32, device=x.device)
        _native.check(lib.pcvl_mlp_gemm(_p(x), self.D, 1, _p(self.W1b), 1, 1, self.D, _p(h), self.H, R, self.H,
                                        self.D, _p(self.b1), 1, 0, _s()), "gemm1")
        _native.check(lib.pcvl_mlp_gemm(_p(h), self.H, 1, _p(self.W2b), 1, 1, self.H, _p(z), self.m, R, self.m,
                                        self.H, _p(self.b2), 0, 0, _s()), "gemm2")
        return h, z

    def loss_and_grads(self, x, lo, target, scale: float):
        """Cross-entropy per row and parameter gradients of scale * sum(loss)."""
        import torch
        lib = _lib()
        R = x.shape[0]
        h, z = self.forward(x)
        loss = torch.empty(R, dtype=torch.float32, device=x.device)
        dz = torch.empty_like(z)
        _native.check(lib.pcvl_mlp_ce(_p(z), R, self.m, _p(target), _p(lo), float(scale), _p(loss), _p(dz), _s()),
                      "ce")
        dW1, db1 = torch.empty_like(self.W1), torch.empty_like(self.b1)
        dW2, db2 = torch.empty_like(self.W2), torch.empty_like(self.b2)
        work = self._wgrad_workspace(R, x.device)
        # dW2 = dz^T h ; db2 = colsum dz  (row-split, deterministic)
        _native.check(lib.pcvl_mlp_wgrad(_p(dz), self.m, _p(h), self.H, R, self.m, self.H, _p(dW2), _p(db2),
                                         _p(work), work.numel(), _s()), "wgrad 2")
        # dh = dz W2 (bf16 copy, as in the forward), relu'
        dh = torch.empty_like(h)
        _native.check(lib.pcvl_mlp_gemm(_p(dz), self.m, 1, _p(self.W2b), 1, self.H, 1, _p(dh), self.H, R, self.H,
                                        self.m, None, 0, 0, _s()), "gemm dh")
        _native.check(lib.pcvl_mlp_relu_bwd(_p(dh), _p(h), dh.numel(), _s()), "relu_bwd")
        # dW1 = dh^T x ; db1 = colsum dh
        _native.check(lib.pcvl_mlp_wgrad(_p(dh), self.H, _p(x), self.D, R, self.H, self.D, _p(dW1), _p(db1),
                                         _p(work), work.numel(), _s()), "wgrad 1")
        return loss, [dW1, db1, dW2, db2]

    def _wgrad_workspace(self, R: int, device):
        import torch
        lib = _lib()
        need = 0
        for mo, no in ((self.m, self.H), (self.H, self.D)):
            f = ctypes.c_int64()
            _native.check(lib.pcvl_mlp_wgrad_workspace(R, mo, no, ctypes.byref(f)), "wgrad_workspace")
            need = max(need, f.value)
        if getattr(self, "_wwork", None) is None or self._wwork.numel() < need or self._wwork.device != device:
            self._wwork = torch.empty(max(1, need), dtype=torch.float32, device=device)
        return self._wwork

    def adam_step(self, grads, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        lib = _lib()
        self._step += 1
        bf = {0: self.W1b, 2: self.W2b}
        for k, (p, g) in enumerate(zip(self.params, grads)):
            _native.check(lib.pcvl_mlp_adam(_p(p), _p(g), _p(self._m1[k]), _p(self._m2[k]),
                                            _p(bf[k]) if k in bf else None, p.numel(), lr, beta1, beta2, eps,
                                            self._step, _s()), "adam")

    # ------------------------------------------------------------------ data
    def teacher_forcing(self, seqs):
        """Inputs, mask bounds and targets of every (sequence, step) row."""
        import torch
        B = seqs.shape[0]
        x = torch.empty((B * self.n, self.D), dtype=torch.float32, device=seqs.device)
        lo = torch.empty(B * self.n, dtype=torch.int32, device=seqs.device)
        tgt = torch.empty(B * self.n, dtype=torch.int32, device=seqs.device)
        _native.check(_lib().pcvl_mis_features_all(_p(seqs), B, self.n, self.m, _p(x), _p(lo), _p(tgt), _s()),
                      "features_all")
        return x, lo, tgt

    def train(self, seqs, epochs: int = 3, batch: int = 2048, lr: float = 2e-3, seed: int = 0, log=None,
              lr_end: float | None = None) -> list:
        """Maximum likelihood on sorted mode sequences (N x n int32 CUDA).
        With lr_end, the step size follows a cosine from lr down to lr_end
        over the run (per batch)."""
        import torch
        g = torch.Generator(device="cpu").manual_seed(seed)
        N = seqs.shape[0]
        history = []
        steps = epochs * ((N + batch - 1) // batch)
        t = 0
        for ep in range(epochs):
            order = torch.randperm(N, generator=g).to(seqs.device)
            tot = torch.zeros((), dtype=torch.float64, device=seqs.device)
            cnt = 0
            for i0 in range(0, N, batch):
                sb = seqs[order[i0:i0 + batch]].contiguous()
                x, lo, tgt = self.teacher_forcing(sb)
                loss, grads = self.loss_and_grads(x, lo, tgt, 1.0 / sb.shape[0])
                step_lr = lr if lr_end is None else lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * t / steps))
                self.adam_step(grads, lr=step_lr)
                tot += loss.sum(dtype=torch.float64)
                cnt += sb.shape[0]
                t += 1
            history.append(float(tot.item()) / cnt)   # mean -log q per sequence
            if log:
                log(f"epoch {ep}: mean -log q = {history[-1]:.4f}")
        return history

    def log_q(self, seqs):
        """Exact log q of sequences (N x n), by teacher forcing."""
        import torch
        x, lo, tgt = self.teacher_forcing(seqs)
        _, z = self.forward(x)
        loss = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        _native.check(_lib().pcvl_mlp_ce(_p(z), x.shape[0], self.m, _p(tgt), _p(lo), 1.0, _p(loss), None, _s()),
                      "ce")
        return -loss.double().view(-1, self.n).sum(dim=1)

    def fused_ok(self) -> bool:
        """Whether the one-launch sampler (pcvl_mis_sample_fused) covers this shape."""
        return self.m <= 64 and 8 <= self.H <= 256 and self.H % 8 == 0

    def sample(self, B: int, seed: int, first: int = 0, fused: bool | None = None):
        """B sequences drawn autoregressively; returns (seq (B x n) int32, log q (B,) float64).
        The fused sampler runs all n steps of a tile of proposals in one launch
        (first layer updated in place); the stepwise path runs one
        features / GEMM / GEMM / draw sequence per step.  Both draw step p of
        proposal b with the uniform (seed, first + b, p) and return the exact
        log q of what they drew."""
        import torch
        lib = _lib()
        seq = torch.zeros((B, self.n), dtype=torch.int32, device="cuda")
        logq = torch.zeros(B, dtype=torch.float64, device="cuda")
        if fused is None:
            fused = self.fused_ok()
        if fused:
            _native.check(lib.pcvl_mis_sample_fused(_p(self.W1b), _p(self.b1), _p(self.W2b), _p(self.b2), B, self.m,
                                                    self.n, self.H, int(seed), int(first), _p(seq), _p(logq), _s()),
                          "sample_fused")
            return seq, logq
        x = torch.empty((B, self.D), dtype=torch.float32, device="cuda")
        lo = torch.empty(B, dtype=torch.int32, device="cuda")
        for p in range(self.n):
            _native.check(lib.pcvl_mis_features(_p(seq), B, self.n, self.m, p, _p(x), _p(lo), _s()), "features")
            _, z = self.forward(x)
            _native.check(lib.pcvl_mis_sample_step(_p(z), B, self.m, self.n, p, _p(lo), int(seed), int(first),
                                                   _p(seq), _p(logq), _s()), "sample_step")
        return seq, logq


def occ_to_seq(occ, n: int):
    import torch
    B, m = occ.shape
    seq = torch.empty((B, n), dtype=torch.int32, device=occ.device)
    _native.check(_lib().pcvl_mis_occ_to_seq(_p(occ.contiguous()), B, m, n, _p(seq), _s()), "occ_to_seq")
    return seq


def seq_to_occ(seq, m: int):
    import torch
    B, n = seq.shape
    occ = torch.empty((B, m), dtype=torch.int32, device=seq.device)
    _native.check(_lib().pcvl_mis_seq_to_occ(_p(seq.contiguous()), B, n, m, _p(occ), _s()), "seq_to_occ")
    return occ


class LearnedMIS:
    """Independence Metropolis chain with a ProposalNet proposal and FP64
    permanent targets, for a fixed unitary and single-photon input.  chunk:
    proposals scored per round of launches (the chain's draws do not depend
    on it; larger rounds leave fewer tails where the device drains: at
    m=60, n=20, rounds of 65,536 take about 5% less time than rounds of
    16,384, and one round of 131,072 for 10^5 samples 1.3% less than two of
    65,536; a round's workspace is about 1.8 KB per proposal)."""

    def __init__(self, U, in_occ, net: ProposalNet, chunk: int = 131072, multiset: bool = True):
        import torch
        self.in_occ = list(in_occ)
        self.m = len(self.in_occ)
        self.n = int(sum(self.in_occ))
        assert net.m == self.m and net.n == self.n
        self.net = net
        self.U = torch.from_numpy(np.ascontiguousarray(U, dtype=np.complex128)).cuda()
        cols = [i for i, v in enumerate(self.in_occ) for _ in range(v)]
        self.cols = torch.tensor(cols, dtype=torch.int32, device="cuda")
        self.log_prod_in = float(sum(math.lgamma(v + 1) for v in self.in_occ))
        self.chunk = chunk
        self.multiset = multiset
        self._perms = torch.empty(chunk, dtype=torch.complex128, device="cuda")
        if not multiset:
            self._plan = PermanentPlan(np.full(chunk, self.n, dtype=np.int32))
            self._mats = torch.empty(chunk * self.n * self.n, dtype=torch.complex128, device="cuda")

    def permanents(self, seqs):
        """perm(U[seq, cols]) of sorted sequences (B x n), complex128 CUDA tensor."""
        import torch
        B = seqs.shape[0]
        out = torch.empty(B, dtype=torch.complex128, device="cuda")
        _native.check(_lib().pcvl_mis_permanents(_p(self.U), self.m, _p(self.cols), self.n, _p(seqs.contiguous()),
                                                 B, _p(out), _s()), "mis_permanents")
        return out

    def log_p(self, seqs):
        """Exact log p of sorted sequences (B x n), by FP64 permanents."""
        import torch
        lib = _lib()
        B = seqs.shape[0]
        out = torch.empty(B, dtype=torch.float64, device="cuda")
        for i0 in range(0, B, self.chunk):
            sb = seqs[i0:i0 + self.chunk].contiguous()
            b = sb.shape[0]
            if self.multiset:
                # repeated rows summed by multiplicity (multiset Glynn)
                _native.check(lib.pcvl_mis_permanents(_p(self.U), self.m, _p(self.cols), self.n, _p(sb), b,
                                                      _p(self._perms), _s()), "mis_permanents")
            else:
                _native.check(lib.pcvl_mis_submatrices(_p(self.U), self.m, _p(self.cols), self.n, _p(sb), b,
                                                       _p(self._mats), _s()), "submatrices")
                if b < self.chunk:   # the plan covers a full chunk; zero the unused matrices
                    self._mats[b * self.n * self.n:].zero_()
                self._plan(self._mats, self._perms)
            _native.check(lib.pcvl_mis_logp(_p(self._perms), _p(sb), self.n, b, self.log_prod_in, _p(out[i0:i0 + b]),
                                            _s()), "logp")
        return out

    def run(self, count: int, seed: int, x0_seq, first: int = 0):
        """count chain samples starting from x0 (a sorted sequence, e.g. one
        exact Clifford-Clifford draw).  Returns dict with 'seq' (count x n
        CUDA int32), 'acceptance', 'logw' of the proposals."""
        import torch
        lib = _lib()
        props, logq = self.net.sample(count, seed, first)
        x0 = x0_seq.reshape(1, self.n).to(device="cuda", dtype=torch.int32).contiguous()
        # x0 is scored in the proposals' launches, at index 0.  The scan starts
        # from weight -inf, so it takes x0 at step 0 whatever that step's
        # uniform, and starting its stream at first - 1 leaves proposal t on
        # (seed, first + t): the same draws as a scan from x0 with x0's weight,
        # with no host round trip for that weight.
        seqs = torch.cat([x0, props])
        logw_all = (self.log_p(seqs) - torch.cat([self.net.log_q(x0), logq])).contiguous()
        state = torch.empty(count + 1, dtype=torch.int64, device="cuda")
        acc = torch.zeros(1, dtype=torch.int64, device="cuda")
        _native.check(lib.pcvl_mis_chain(_p(logw_all), count + 1, float("-inf"), int(seed) ^ 0x5DEECE66D,
                                         (int(first) - 1) & 0xFFFFFFFFFFFFFFFF, _p(state), _p(acc), _s()), "chain")
        out = torch.empty((count, self.n), dtype=torch.int32, device="cuda")
        _native.check(lib.pcvl_mis_gather(_p(seqs), _p(state[1:]), count, self.n, _p(x0), _p(out), _s()), "gather")
        moves = acc - (state[0] >= 0).to(torch.int64)   # step 0 (taking x0) is not a move
        return {"seq": out, "acceptance": float(moves.item()) / max(1, count), "logw": logw_all[1:],
                "proposals": props}

    def close(self):
        if not self.multiset:
            self._plan.close()


# --------------------------------------------------------------- diagnostics

def marginal_tvds(occ_a: np.ndarray, occ_b: np.ndarray, pairs: int = 200, seed: int = 0) -> dict:
    """TVD of 1-mode marginals (distribution of n_j, every mode) and of 2-mode
    marginals (joint (n_j, n_k), `pairs` seeded mode pairs) between two sample
    sets (rows = occupation vectors)."""
    m = occ_a.shape[1]
    one = []
    for j in range(m):
        va = np.bincount(occ_a[:, j], minlength=occ_a.max() + occ_b.max() + 2) / len(occ_a)
        vb = np.bincount(occ_b[:, j], minlength=len(va)) / len(occ_b)
        L = max(len(va), len(vb))
        va = np.pad(va, (0, L - len(va)))
        vb = np.pad(vb, (0, L - len(vb)))
        one.append(0.5 * np.abs(va - vb).sum())
    rng = np.random.RandomState(seed)
    two = []
    K = int(max(occ_a.max(), occ_b.max())) + 1
    for _ in range(pairs):
        j, k = rng.choice(m, 2, replace=False)
        ia = occ_a[:, j] * K + occ_a[:, k]
        ib = occ_b[:, j] * K + occ_b[:, k]
        va = np.bincount(ia, minlength=K * K) / len(ia)
        vb = np.bincount(ib, minlength=K * K) / len(ib)
        two.append(0.5 * np.abs(va - vb).sum())
    return {"tvd1_max": float(max(one)), "tvd1_mean": float(np.mean(one)),
            "tvd2_max": float(max(two)), "tvd2_mean": float(np.mean(two))}
