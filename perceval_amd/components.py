"""Linear-optics circuit algebra: just what the backends need to obtain the
m x m unitary of a circuit (ABackend.set_circuit -> circuit.compute_unitary,
perceval/backends/_abstract_backends.py:45-54).

Conventions follow the reference:
* BS (perceval/components/unitary_components.py:48-183): default convention
  Rx, template [[1, i], [i, 1]]; Ry [[1, -1], [1, 1]]; H [[1, 1], [1, -1]];
  entries scaled by cos(theta/2) / sin(theta/2) and the phases
  u00 *= e^{i(phi_tl+phi_tr)}, u01 *= e^{i(phi_tr+phi_bl)},
  u10 *= e^{i(phi_tl+phi_br)}, u11 *= e^{i(phi_bl+phi_br)}.
* PS (:262-271): [[e^{i phi}]].
* PERM (:464-480): U[perm[i], i] = 1.
* Circuit (perceval/components/linear_circuit.py:487-503): components are
  embedded in an m x m identity and multiplied on the left, U = cU @ U.
* GenericInterferometer rectangle (perceval/components/generic_interferometer.py:150-165)
  and the MZI "phase first" cell (perceval/components/core_catalog/mzi.py:69-83).
Only numeric parameters are supported (no symbolic computation).
"""
from __future__ import annotations

import math
from enum import IntEnum

import numpy as np


class Parameter:
    """Named numeric circuit parameter (perceval/utils/parameter.py, numeric subset)."""

    def __init__(self, name: str, value: float | None = None, min_v=None, max_v=None):
        self.name = name
        self._value = None if value is None else float(value)
        self.min = min_v
        self.max = max_v

    @property
    def defined(self) -> bool:
        return self._value is not None

    def set_value(self, v: float):
        self._value = float(v)

    def reset(self):
        self._value = None

    def __float__(self):
        if self._value is None:
            raise ValueError(f"parameter {self.name} has no value")
        return self._value

    def __repr__(self):
        return f"Parameter({self.name}={self._value})"


P = Parameter


def _param(value, name, owner_params: dict):
    if isinstance(value, Parameter):
        owner_params[value.name] = value
        return value
    if isinstance(value, str):
        p = Parameter(value)
        owner_params[p.name] = p
        return p
    p = Parameter(name, float(value))
    return p


class ACircuit:
    def __init__(self, m: int, name: str | None = None):
        self._m = m
        self.name = name
        self._params: dict = {}

    @property
    def m(self) -> int:
        return self._m

    @property
    def requires_polarization(self) -> bool:
        return False

    def get_parameters(self, all_params: bool = False) -> list:
        ps = list(self._params.values())
        return ps if all_params else [p for p in ps if not _is_fixed(p)]

    def compute_unitary(self, use_symbolic: bool = False, assign: dict | None = None) -> np.ndarray:
        if use_symbolic:
            raise NotImplementedError("symbolic unitaries are not supported by this simulation core")
        if assign:
            for k, v in assign.items():
                self._params[k].set_value(v)
        return self._compute_unitary()

    def _compute_unitary(self) -> np.ndarray:
        raise NotImplementedError

    @property
    def U(self) -> np.ndarray:
        return self.compute_unitary()

    def __floordiv__(self, other):
        c = Circuit(self.m)
        c.add(0, self)
        if isinstance(other, tuple):
            pos, comp = other
            c.add(pos, comp)
        else:
            c.add(0, other)
        return c


def _is_fixed(p: Parameter) -> bool:
    return getattr(p, "_fixed", False)


def _fixed(name, value) -> Parameter:
    p = Parameter(name, value)
    p._fixed = True
    return p


class BSConvention(IntEnum):
    Rx = 0
    Ry = 1
    H = 2


class BS(ACircuit):
    def __init__(self, theta=math.pi / 2, phi_tl=0, phi_bl=0, phi_tr=0, phi_br=0,
                 convention: BSConvention = BSConvention.Rx):
        super().__init__(2, "BS")
        self._convention = convention
        self._theta = self._p(theta, "theta")
        self._phi_tl = self._p(phi_tl, "phi_tl")
        self._phi_bl = self._p(phi_bl, "phi_bl")
        self._phi_tr = self._p(phi_tr, "phi_tr")
        self._phi_br = self._p(phi_br, "phi_br")

    def _p(self, v, name):
        if isinstance(v, (Parameter, str)):
            return _param(v, name, self._params)
        return _fixed(name, v)

    @staticmethod
    def H(theta=math.pi / 2, phi_tl=0, phi_bl=0, phi_tr=0, phi_br=0):
        return BS(theta, phi_tl, phi_bl, phi_tr, phi_br, convention=BSConvention.H)

    @staticmethod
    def Rx(theta=math.pi / 2, phi_tl=0, phi_bl=0, phi_tr=0, phi_br=0):
        return BS(theta, phi_tl, phi_bl, phi_tr, phi_br, convention=BSConvention.Rx)

    @staticmethod
    def Ry(theta=math.pi / 2, phi_tl=0, phi_bl=0, phi_tr=0, phi_br=0):
        return BS(theta, phi_tl, phi_bl, phi_tr, phi_br, convention=BSConvention.Ry)

    @staticmethod
    def r_to_theta(r: float) -> float:
        return 2 * math.acos(math.sqrt(r))

    @staticmethod
    def theta_to_r(theta: float) -> float:
        return math.cos(float(theta) / 2) ** 2

    def _compute_unitary(self):
        th = float(self._theta)
        c, s = math.cos(th / 2), math.sin(th / 2)
        tl, bl, tr, br = float(self._phi_tl), float(self._phi_bl), float(self._phi_tr), float(self._phi_br)
        if self._convention == BSConvention.Rx:
            t = np.array([[1, 1j], [1j, 1]], dtype=complex)
        elif self._convention == BSConvention.Ry:
            t = np.array([[1, -1], [1, 1]], dtype=complex)
        else:
            t = np.array([[1, 1], [1, -1]], dtype=complex)
        t[0, 0] *= np.exp(1j * (tl + tr)) * c
        t[0, 1] *= np.exp(1j * (tr + bl)) * s
        t[1, 1] *= np.exp(1j * (bl + br)) * c
        t[1, 0] *= np.exp(1j * (tl + br)) * s
        return t


class PS(ACircuit):
    def __init__(self, phi):
        super().__init__(1, "PS")
        self._phi = _param(phi, "phi", self._params) if isinstance(phi, (Parameter, str)) else _fixed("phi", phi)

    @property
    def phi(self) -> Parameter:
        return self._phi

    def _compute_unitary(self):
        ph = float(self._phi)
        return np.array([[math.cos(ph) + 1j * math.sin(ph)]], dtype=complex)


class Unitary(ACircuit):
    def __init__(self, U, name: str | None = None):
        U = np.asarray(U, dtype=complex)
        assert U.ndim == 2 and U.shape[0] == U.shape[1], "Unitary needs a square matrix"
        super().__init__(U.shape[0], name or "U")
        self._u = U.copy()

    def _compute_unitary(self):
        return self._u.copy()


class PERM(Unitary):
    def __init__(self, perm: list):
        assert sorted(perm) == list(range(len(perm))), f"{perm} is not a permutation"
        n = len(perm)
        u = np.zeros((n, n), dtype=complex)
        for i, v in enumerate(perm):
            u[v, i] = 1
        super().__init__(u, "PERM")


class Barrier(ACircuit):
    def __init__(self, m: int, visible: bool = True):
        super().__init__(m, "BARRIER")

    def _compute_unitary(self):
        return np.eye(self.m, dtype=complex)


class Circuit(ACircuit):
    def __init__(self, m: int, name: str | None = None):
        super().__init__(m, name)
        self._components: list = []

    def add(self, port, component: ACircuit, merge: bool = True):
        if isinstance(port, int):
            r = tuple(range(port, port + component.m))
        else:
            r = tuple(port)
            assert len(r) == component.m, "port range does not match the component size"
        assert r[0] >= 0 and r[-1] < self.m and list(r) == list(range(r[0], r[0] + len(r))), \
            "component must sit on consecutive modes of the circuit"
        if merge and isinstance(component, Circuit):
            for sub_r, sub_c in component._components:
                self._components.append((tuple(x + r[0] for x in sub_r), sub_c))
        else:
            self._components.append((r, component))
        for p in component._params.values():
            self._params[p.name] = p
        return self

    def __floordiv__(self, other):
        if isinstance(other, tuple):
            pos, comp = other
            return self.add(pos, comp)
        return self.add(0, other)

    def __iter__(self):
        return iter(self._components)

    def ncomponents(self) -> int:
        return len(self._components)

    def _compute_unitary(self):
        u = np.eye(self.m, dtype=complex)
        for r, c in self._components:
            cu = c.compute_unitary()
            lo, hi = r[0], r[-1] + 1
            # cU @ u with cU the component embedded in the identity
            u[lo:hi, :] = cu @ u[lo:hi, :]
        return u


class InterferometerShape(IntEnum):
    RECTANGLE = 0
    TRIANGLE = 1


class GenericInterferometer(Circuit):
    """Rectangular (Clements) or triangular mesh of 2-mode cells
    (perceval/components/generic_interferometer.py:150-180)."""

    def __init__(self, m: int, fun_gen, shape: InterferometerShape = InterferometerShape.RECTANGLE,
                 depth: int | None = None, phase_shifter_fun_gen=None, phase_at_output: bool = False):
        super().__init__(m, "Generic interferometer")
        self._depth_per_mode = [0] * m
        if phase_shifter_fun_gen and not phase_at_output:
            for i in range(m):
                self.add(i, phase_shifter_fun_gen(i))
        if shape == InterferometerShape.RECTANGLE:
            max_depth = m if depth is None else depth
            idx = 0
            for i in range(max_depth):
                for j in range(i % 2, m - 1, 2):
                    if depth is not None and (self._depth_per_mode[j] == depth or self._depth_per_mode[j + 1] == depth):
                        continue
                    self.add((j, j + 1), fun_gen(idx), merge=True)
                    self._depth_per_mode[j] += 1
                    self._depth_per_mode[j + 1] += 1
                    idx += 1
        else:
            idx = 0
            for i in range(1, m):
                for j in range(i, 0, -1):
                    self.add((j - 1, j), fun_gen(idx), merge=True)
                    idx += 1
        if phase_shifter_fun_gen and phase_at_output:
            for i in range(m):
                self.add(i, phase_shifter_fun_gen(i))


def mzi_phase_first(i: int | None = None, phi_a=None, phi_b=None, theta_a=math.pi / 2, theta_b=math.pi / 2) -> Circuit:
    """PS(phi_a) on mode 0, BS, PS(phi_b) on mode 0, BS (core_catalog/mzi.py:69-83)."""
    phi_a = phi_a if phi_a is not None else (f"phi_a{i}" if i is not None else "phi_a")
    phi_b = phi_b if phi_b is not None else (f"phi_b{i}" if i is not None else "phi_b")
    return (Circuit(2, name="MZI") // (0, PS(phi=phi_a)) // BS(theta=theta_a) // (0, PS(phi=phi_b))
            // BS(theta=theta_b))


def postprocessed_cz() -> Circuit:
    """core_catalog/postprocessed_cz.py:50-59."""
    theta_13 = BS.r_to_theta(1 / 3)
    return (Circuit(6, name="PostProcessed CZ")
            .add(1, PERM([2, 1, 3, 0]))
            .add(0, Barrier(6, visible=False))
            .add((0, 1), BS.H(theta_13))
            .add((2, 3), BS.H(theta_13))
            .add((4, 5), BS.H(theta_13))
            .add(1, PERM([3, 1, 0, 2])))


def postprocessed_cnot() -> Circuit:
    """core_catalog/postprocessed_cnot.py:47-51."""
    return (Circuit(6, name="PostProcessed CNOT")
            .add((2, 3), BS.H())
            .add(0, postprocessed_cz(), merge=True)
            .add((2, 3), BS.H()))


class _CatalogItem:
    def __init__(self, builder):
        self._builder = builder

    def build_circuit(self, **kwargs) -> Circuit:
        return self._builder(**kwargs)


catalog = {
    "mzi phase first": _CatalogItem(lambda **kw: mzi_phase_first(**kw)),
    "postprocessed cz": _CatalogItem(lambda **kw: postprocessed_cz()),
    "postprocessed cnot": _CatalogItem(lambda **kw: postprocessed_cnot()),
}


def clements_mesh(m: int) -> GenericInterferometer:
    """GenericInterferometer(m, MZI phase first, RECTANGLE): m(m-1)/2 cells and
    m(m-1) phase parameters phi_a{i}, phi_b{i}."""
    return GenericInterferometer(m, lambda i: mzi_phase_first(i=i), shape=InterferometerShape.RECTANGLE)
