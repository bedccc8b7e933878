"""MI355X-native engine for the PandemicThrust daily simulation loop.

The product is the gfx950 library lib/libpt_hip.so behind the C ABI in
include/pandemic/pt_abi.h, plus the C++ host program (host/) that keeps the
reference's PandemicSim flow and file contract.  This package exposes a
ctypes binding of the same ABI for Python callers.
"""
from .api import Interventions, PTError, Params, PandemicBatch, PandemicSim, glibc_initial_infected, glibc_seeds, lib  # noqa: F401
