"""`import pbi4 as pb` for scripts written against the reference package (its Tutorials/Script
call sequence and command line): the drop-in IO, DENSITIES and MEASUREMENTS of pbi4_amd under the
reference's module name.  Put this directory on PYTHONPATH to run such scripts unchanged."""
from pbi4_amd.pbi4 import DENSITIES, IO, MEASUREMENTS, NumpyArrayEncoder  # noqa: F401
