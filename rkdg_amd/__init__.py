"""rkdg_amd — MI355X-native explicit RKDG right-hand side (FE_Evolution::Mult and its
limiter / RK stage path) behind a C-ABI (include/rkdg_capi.h).

Importing the package is cheap; the HIP library is loaded on first use by
``rkdg_amd.capi`` and fails loudly if it is missing.
"""
from . import mesh, scheme  # noqa: F401

__all__ = ["mesh", "scheme"]
