"""CPU oracle for normal-equation CP-ALS — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy + scipy LAPACK, the algorithm of ITensorCPD.jl's
dense-tensor CP-ALS path (``decompose`` -> ``als_optimize`` -> ``optimize`` with
``KRPFreeNormal`` / ``KRPNormal``, ``ldiv_solve!!``, ``row_norm`` and the
``FitCheck`` / ``NoCheck`` convergence checks).  Each function cites the
reference file:line it follows.

It is the checker, never the product:
  * only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
    leg may import it;
  * the GPU engine (``itensorcpd.jl_amd``) never imports or calls it.

Parity status of each function (what pins it) is listed in the docstring of
``oracle.als``; the functions with nothing that pins them are marked there as
"parity unpinned".
"""
from .als import *  # noqa: F401,F403
