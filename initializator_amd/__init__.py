"""MI355X-native per-timestep Bloch-Torrey magnetization update.

The solver runs in libnmr_hip.so (hand-written gfx950 HIP kernels behind the
C ABI of include/nmr_hip.h).  ``HipSolver`` is the Python host wrapper;
``generators`` makes seeded synthetic inputs; ``layout`` converts between the
compact layout and the reference's padded layout.
"""
from . import generators, layout  # noqa: F401

__all__ = ["generators", "layout", "HipSolver", "load_library"]


def load_library():
    from . import _lib

    return _lib.load()


def __getattr__(name):
    if name == "HipSolver":
        from .solver import HipSolver

        return HipSolver
    raise AttributeError(name)
