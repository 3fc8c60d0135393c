"""xfdtd_amd: MI355X (gfx950) field engine behind the XFDTD Simulation API.

Layers:
  include/xfdtd/xfdtd.hpp   C++ Simulation / Object / PML / TFSF3D / NffftFd ...
  include/xfdtd_sim.h       C-ABI over that C++ API
  include/xfdtd_engine.h    C-ABI of the HIP engine (the device boundary)
  xfdtd_amd.scene           scene description shared by the engine and the tests
  xfdtd_amd.sim.Simulation  Python handle over include/xfdtd_sim.h
"""
from . import scene  # noqa: F401
from ._lib import LIB_PATH, lib  # noqa: F401


def Simulation(*args, **kwargs):
    from .sim import Simulation as _S
    return _S(*args, **kwargs)
