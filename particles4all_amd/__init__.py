"""particles4all_amd: an MI355X (gfx950) implementation of the Particles4All
simulation step behind a C ABI (include/p4a.h, lib/libp4a.so).

The drop-in surface is the JavaScript `Sim` in node/sim.js (Node.js N-API
addon); this Python package carries the ctypes plumbing used by the tests
and the benchmark, plus the synthetic workload definitions.
"""
from . import _lib  # noqa: F401
from ._lib import (Sim, Batch, Scene, default_params, lattice_scene, preset_scene,  # noqa: F401
                   lib, BUF, STAGE)

__all__ = ["Sim", "Batch", "Scene", "default_params", "lattice_scene", "preset_scene", "lib"]
