"""MI355X-native simulation core for Perceval's Naive, SLOS and
CliffordClifford2017 backends: HIP kernels behind a C ABI (include/pcvl_hip.h),
loaded with ctypes; torch provides device memory and streams."""
__version__ = "0.1.0"

from .backends import (ABackend, ASamplingBackend, AStrongSimulationBackend, BACKEND_LIST, BackendFactory,
                       Clifford2017Backend, NaiveBackend, SLOSBackend)
from .components import (ACircuit, BS, BSConvention, Barrier, Circuit, GenericInterferometer,
                         InterferometerShape, P, Parameter, PERM, PS, Unitary, catalog, clements_mesh)
from .utils._random import random_seed
from .utils.dist_metrics import tvd_dist
from .utils.globals import global_params
from .utils.mask import FSMask
from .utils.matrix import Matrix
from .utils.states import (BasicState, BSCount, BSDistribution, BSSamples, FockState, StateVector,
                           allstate_iterator, max_photon_state_iterator)
