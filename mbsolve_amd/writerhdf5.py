"""Importing this module registers the "hdf5" writer with the writer
factory, as ``import mbsolve.writerhdf5`` does for the reference
(writer-hdf5/swig/writerhdf5.i).  ``read_file(name)`` returns the
attributes and results of a file written by it."""
from ._writerhdf5 import load_writer, read_file  # noqa: F401

load_writer()
