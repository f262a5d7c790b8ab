"""eigenfns_amd — MI355X-native matrix-free Maxwell eigensolver hot path.

Drop-in for the reference package ``eigenfns`` on its hot path:
  eigenfns_amd.operator   ← eigenfns.operator   (Θ, preconditioners, basis)
  eigenfns_amd.solver     ← eigenfns.solver     (gram/combine/deflate/SVQB, deflated block LOBPCG)
  eigenfns_amd.chebyshev  ← eigenfns.chebyshev  (Lanczos λ_max, lowpass ChebSI, KPM count)
  eigenfns_amd.interior   ← eigenfns.interior   (bandpass filter, RR extraction, filtered SI)
  eigenfns_amd.kpm        ← scripts/exp/exp_kpm_dos.py + exp_kpm_analyze.py (KPM DOS moments)
  eigenfns_amd.structure  ← eigenfns.structure  (rasterizer) + synthetic rod networks
  eigenfns_amd.io         ← eigenfns.io         (block checkpoints, the reference's file format)
All device arithmetic runs in the HIP library libeigenfns_hip.so (C ABI,
include/eigenfns_hip.h); torch supplies device memory and streams.

``eigenfns_amd.install_alias()`` registers the modules under the reference's
import names (``eigenfns``, ``eigenfns.operator``, ...) for unchanged callers.
"""
from __future__ import annotations

import importlib
import sys

__version__ = "0.3.0"

_SUBMODULES = ("operator", "solver", "chebyshev", "interior", "kpm", "structure", "linalg", "io")


def install_alias(name: str = "eigenfns") -> None:
    """Make ``import eigenfns.operator`` (etc.) resolve to this package."""
    pkg = sys.modules[__name__]
    sys.modules.setdefault(name, pkg)
    for sub in _SUBMODULES:
        mod = importlib.import_module(f"{__name__}.{sub}")
        sys.modules.setdefault(f"{name}.{sub}", mod)


def __getattr__(attr):
    if attr in _SUBMODULES:
        return importlib.import_module(f"{__name__}.{attr}")
    raise AttributeError(attr)
