"""tensorcircuit_amd: the dense statevector path of TensorCircuit
(Circuit gate application, expectation_ps / Pauli-sum energies,
backend.grad / backend.vmap) on hand-written HIP kernels for MI355X.

Everything computes on the HIP device through ``libtcm.so``; there is no CPU
fallback (importing the package fails when the library is not built).
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401  (fails loudly without libtcm.so)
from . import cons, experimental, gates, interfaces, quantum, templates, torchnn, translation
from .backend import backend
from .circuit import Circuit, expectation
from .cons import get_backend, get_contractor, runtime_dtype, set_backend, set_contractor, set_dtype, set_jit
from .engine import PauliSum, Program
from .gates import array_to_tensor, num_to_tensor
from .interfaces import torch_interface
from .torchnn import QuantumNet, TorchLayer

K = backend
