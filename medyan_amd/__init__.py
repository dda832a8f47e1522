"""medyan_amd -- MI355X-native implementation of MEDYAN's mechanical-equilibration
path (force fields, cylinder neighbor lists, Polak-Ribiere minimizer).

The compute path is libmdy_hip.so (hand-written HIP for gfx950 behind the C ABI
in include/mdy.h); this package holds its Python host binding.
"""
from . import _abi  # noqa: F401
from .network import Network, WORKLOADS, generate  # noqa: F401

__all__ = ["Network", "WORKLOADS", "generate", "mechanics"]
