"""MI355X-native level-set interface transport (WENO3 + SSP-RK3 advection,
Sussman reinitialisation, water volume) for curvilinear grids, FP64.

The numerical work runs in hand-written HIP kernels of liblsm_hip.so behind a
C ABI (include/lsm_hip.h); this package is the PyTorch-side front end.
"""
from . import inputs  # noqa: F401

__all__ = ["inputs", "api"]


def __getattr__(name):
    # the GPU front end imports torch; load it on first use
    if name == "api":
        import importlib
        return importlib.import_module(__name__ + ".api")
    raise AttributeError(name)
