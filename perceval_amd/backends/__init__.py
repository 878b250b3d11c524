"""Backend registry and factory (perceval/backends/__init__.py:39-62).

BACKEND_LIST holds the backends this simulation core provides; every
caller that looks a backend up by name reads this same dict.
"""
import warnings

from ._abstract_backends import ABackend, ASamplingBackend, AStrongSimulationBackend
from ._clifford2017 import Clifford2017Backend
from ._naive import NaiveBackend
from ._naive_approx import NaiveApproxBackend
from ._slos import SLOSBackend

BACKEND_LIST = {
    "CliffordClifford2017": Clifford2017Backend,
    "Naive": NaiveBackend,
    "NaiveApprox": NaiveApproxBackend,
    "SLOS": SLOSBackend,
}


class BackendFactory:
    @staticmethod
    def get_backend(backend_name: str = "SLOS", **kwargs) -> ABackend:
        name = backend_name
        if name in BACKEND_LIST:
            return BACKEND_LIST[name](**kwargs)
        warnings.warn(f'Backend "{name}" not found. Falling back on SLOS')
        return BACKEND_LIST['SLOS'](**kwargs)

    @staticmethod
    def list():
        return list(BACKEND_LIST.keys())
