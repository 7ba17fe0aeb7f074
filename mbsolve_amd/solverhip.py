"""Importing this module registers the hip-fdtd-* solvers with the solver
factory, as ``import mbsolve.solvercpu`` does for the reference's CPU
solvers (solver-cpu/swig/solvercpu.i)."""
from ._mbsolve import load_hip_solvers

load_hip_solvers()
