"""MI355X-native Zheng-Shu-Chew bubbly-flow lattice-Boltzmann step.

The compute lives in ``libzsc.so`` (hand-written HIP for gfx950 behind the
C-ABI of ``include/zsc.h``); this package is its Python host binding.
"""
from .zsc import (Context, Inputs, Params, bytes_per_cell, derive_params, generate_bubbles,
                  init_phi_host, read_bubbles, read_inputs, write_bubbles)

__all__ = ["Context", "Inputs", "Params", "bytes_per_cell", "derive_params", "generate_bubbles",
           "init_phi_host", "read_bubbles", "read_inputs", "write_bubbles"]
