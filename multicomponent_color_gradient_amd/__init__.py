"""MI355X-native D3Q27 colour-gradient lattice Boltzmann solver.

The hot path (one time step: moments, colour gradient, regularised BGK
collision, surface-tension perturbation, Latva-Kokko recolouring, streaming,
jet inlet / convective outlet, radial statistics) runs in hand-written HIP
kernels for gfx950 behind the C ABI of include/mcg.h.  This package is the
Python binding of that ABI plus the reference's case API and output formats.
"""
from ._abi import McgCase, McgError, load_library, library_path  # noqa: F401
from .case import (derived, interior_nodes, jet_case, periodic_bubble_case, perturbed_jet_case,  # noqa: F401
                   reference_jet, rising_bubble_case)
from .solver import Solver, bubbleIn, jetIn  # noqa: F401

__version__ = "0.1.0"
