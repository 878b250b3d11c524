"""Phase gradients of output probabilities (host side of pcvl_mesh_adjoint),
exposed to torch autograd.

A circuit made of phase shifters and fixed 2-mode components (beam splitters,
2-mode permutations) is compiled into a mesh program; its named phase
parameters become the trainable inputs.  For a batch of phase vectors the
kernel returns the summed probability of a set of target output states and its
exact gradient (adjoint method), so an optimiser can replace the reference's
finite-difference loop (benchmark/benchmark_QML-DE-solver.py:140-142).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .components import ACircuit, Circuit, PS, Parameter


class MeshProgram:
    """Circuit -> (kind, mode, param, const) arrays; phases ordered as
    circuit.get_parameters()."""

    def __init__(self, circuit: Circuit):
        self.m = circuit.m
        self.params: list[Parameter] = circuit.get_parameters()
        index = {id(p): i for i, p in enumerate(self.params)}
        kind, mode, param, const = [], [], [], []
        for r, comp in _flatten(circuit):
            c8 = np.zeros(8)
            if isinstance(comp, PS):
                p = comp.phi
                kind.append(0)
                mode.append(r[0])
                if id(p) in index:
                    param.append(index[id(p)])
                else:
                    param.append(-1)
                    c8[0] = float(p)
            else:
                if comp.m == 1:
                    u = comp.compute_unitary()
                    kind.append(0)
                    mode.append(r[0])
                    param.append(-1)
                    c8[0] = float(np.angle(u[0, 0]))
                    if not np.isclose(abs(u[0, 0]), 1.0):
                        raise ValueError("1-mode components must be phase shifters")
                elif comp.m == 2:
                    if any(id(p) in index for p in comp._params.values()):
                        raise ValueError("only phase shifters may carry trainable parameters")
                    u = comp.compute_unitary()
                    kind.append(1)
                    mode.append(r[0])
                    param.append(-1)
                    c8[:] = np.stack([u.ravel().real, u.ravel().imag], axis=1).ravel()
                else:
                    raise ValueError("mesh programs support 1- and 2-mode components only")
            const.append(c8)
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.mode = np.ascontiguousarray(mode, dtype=np.int32)
        self.param = np.ascontiguousarray(param, dtype=np.int32)
        self.const = np.ascontiguousarray(np.array(const).reshape(-1), dtype=np.float64)
        self.nops = len(kind)

    @property
    def nparams(self) -> int:
        return len(self.params)

    def unitary(self, phases: np.ndarray) -> np.ndarray:
        """Host evaluation of the program (for tests and checks)."""
        U = np.eye(self.m, dtype=complex)
        for k, q, p, c in zip(self.kind, self.mode, self.param, self.const.reshape(-1, 8)):
            if k == 0:
                ph = phases[p] if p >= 0 else c[0]
                U[q] *= np.exp(1j * ph)
            else:
                B = (c[0::2] + 1j * c[1::2]).reshape(2, 2)
                U[q:q + 2] = B @ U[q:q + 2]
        return U


def _flatten(circuit: ACircuit, offset: int = 0):
    """(modes, elementary component) pairs in circuit order."""
    if isinstance(circuit, Circuit):
        for r, c in circuit:
            if isinstance(c, Circuit):
                yield from _flatten(c, offset + r[0])
            else:
                yield tuple(x + offset for x in r), c
    else:
        yield tuple(range(offset, offset + circuit.m)), circuit


class MeshPlan:
    """A mesh program, an input state and target outputs uploaded once
    (pcvl_mesh_plan_create); each call enqueues the adjoint kernel for a batch
    of phase vectors, with no allocation or synchronisation in the library."""

    def __init__(self, program: MeshProgram, in_occ, targets):
        self._lib = _native.load()
        self.program = program
        self.in_occ = np.ascontiguousarray(list(in_occ), dtype=np.int32)
        self.targets = np.ascontiguousarray(np.asarray([list(t) for t in targets], dtype=np.int32))
        self.ntargets = len(self.targets)
        self._plan = ctypes.c_void_p()
        i32 = ctypes.POINTER(ctypes.c_int32)
        _native.check(self._lib.pcvl_mesh_plan_create(
            program.kind.ctypes.data_as(i32), program.mode.ctypes.data_as(i32), program.param.ctypes.data_as(i32),
            program.const.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), program.nops, program.m,
            program.nparams, self.in_occ.ctypes.data_as(i32), self.targets.ctypes.data_as(i32), self.ntargets,
            ctypes.byref(self._plan)), "pcvl_mesh_plan_create")
        self.weights = None

    def set_weights(self, weights=None):
        """Per-target weights w_s of the loss sum_s w_s P(s) (None: all 1)."""
        if weights is None:
            self.weights = None
            _native.check(self._lib.pcvl_mesh_plan_set_weights(self._plan, None), "pcvl_mesh_plan_set_weights")
            return
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
        if w.size != self.ntargets:
            raise ValueError(f"expected {self.ntargets} target weights, got {w.size}")
        _native.check(self._lib.pcvl_mesh_plan_set_weights(self._plan, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                      "pcvl_mesh_plan_set_weights")
        self.weights = w

    def __call__(self, phases, want_probs: bool = False, stream=None):
        import torch
        if isinstance(phases, np.ndarray):
            phases = torch.from_numpy(np.ascontiguousarray(phases, dtype=np.float64)).cuda()
        phases = phases.to(torch.float64).contiguous()
        B, Pn = phases.shape
        if Pn != self.program.nparams:
            raise ValueError(f"expected {self.program.nparams} phases per set, got {Pn}")
        loss = torch.empty(B, dtype=torch.float64, device=phases.device)
        grad = torch.empty((B, Pn), dtype=torch.float64, device=phases.device)
        probs = torch.empty((B, self.ntargets), dtype=torch.float64, device=phases.device) if want_probs else None
        _native.check(self._lib.pcvl_mesh_plan_execute(
            self._plan, _native.ptr(phases), B, _native.ptr(loss), _native.ptr(grad),
            _native.ptr(probs) if probs is not None else None, _native.stream_handle(stream)),
            "pcvl_mesh_plan_execute")
        return (loss, grad, probs) if want_probs else (loss, grad)

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _native.check(self._lib.pcvl_mesh_plan_destroy(self._plan), "pcvl_mesh_plan_destroy")
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MAX_PLAN_TARGETS = 512   # targets of one kernel plan; more are split over several plans


def mesh_adjoint(program: MeshProgram, phases, in_occ, targets, want_probs: bool = False, weights=None):
    """Run the kernel.  phases: (B, P) float64 CUDA tensor or numpy.
    Returns (loss (B,), grad (B, P)[, probs (B, T)]) as CUDA tensors, with
    loss = sum_s w_s P(s) over the targets (weights None: all w_s = 1), e.g.
    the expectation of a diagonal observable over every output state.  A
    plan takes up to 512 targets (run in blocks of 64 inside the kernel);
    more, or a set whose tables do not fit the kernel's LDS (-4), are split
    over several plans, whose losses and gradients add (the gradient is
    linear in the targets).  The program keeps one MeshPlan per (input
    state, targets)."""
    import torch
    in_occ = tuple(int(v) for v in in_occ)
    targets = np.ascontiguousarray(np.asarray([list(t) for t in targets], dtype=np.int32))
    w = None if weights is None else np.asarray(weights, dtype=np.float64).reshape(-1)
    if w is not None and w.size != len(targets):
        raise ValueError(f"expected {len(targets)} target weights, got {w.size}")
    plans = program.__dict__.setdefault("_plans", {})
    loss = grad = None
    probs = []
    t0, size = 0, MAX_PLAN_TARGETS
    while t0 < max(1, len(targets)):
        part = targets[t0:t0 + size]
        key = (in_occ, part.shape, part.tobytes())
        plan = plans.get(key)
        if plan is None:
            try:
                plan = MeshPlan(program, in_occ, part)
            except _native.NativeError as e:
                if "failed (-4)" not in str(e) or size == 1:
                    raise
                size //= 2   # the part's tables do not fit the kernel's LDS: smaller parts
                continue
            plans[key] = plan
        wp = None if w is None else w[t0:t0 + size]
        t0 += size
        if not (wp is None and plan.weights is None) and not (wp is not None and plan.weights is not None
                                                              and np.array_equal(wp, plan.weights)):
            plan.set_weights(wp)
        res = plan(phases, want_probs)
        loss = res[0] if loss is None else loss + res[0]
        grad = res[1] if grad is None else grad + res[1]
        if want_probs:
            probs.append(res[2])
    if want_probs:
        return loss, grad, torch.cat(probs, dim=1)
    return loss, grad


def _autograd_function():
    import torch

    class TargetProbability(torch.autograd.Function):
        """loss[b] = sum_{s in targets} w_s P_b(s); backward from the adjoint kernel."""

        @staticmethod
        def forward(ctx, phases, program, in_occ, targets, weights):
            loss, grad = mesh_adjoint(program, phases.detach(), in_occ, targets, weights=weights)
            ctx.save_for_backward(grad)
            return loss

        @staticmethod
        def backward(ctx, grad_loss):
            (grad,) = ctx.saved_tensors
            return grad_loss[:, None] * grad, None, None, None, None

    return TargetProbability


_TP = None


def target_probability(phases, program: MeshProgram, in_occ, targets, weights=None):
    """Differentiable sum_s w_s P(s) over the targets (weights None: all 1)
    for a batch of phase vectors."""
    global _TP
    if _TP is None:
        _TP = _autograd_function()
    return _TP.apply(phases, program, in_occ, targets, weights)
