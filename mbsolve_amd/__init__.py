"""mbsolve_amd -- MI355X (gfx950) Maxwell-Bloch FDTD solvers behind the
mbsolve solver API.

``import mbsolve_amd.lib as mb`` gives the mbsolve API (like ``mbsolve.lib``);
``import mbsolve_amd.solverhip`` registers the ``hip-fdtd-*`` solvers (like
``mbsolve.solvercpu`` registers ``cpu-fdtd-*``).  The native libraries are
built in-tree by ``python -m mbsolve_amd.build``.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))

__all__ = ["PKG_DIR"]
