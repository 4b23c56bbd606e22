"""CPU oracle for the Tractor render-and-fit hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy (FP64 unless the reference itself computes in
FP32), the reference algorithms that the HIP library replaces.  It exists to
check the HIP path; it is never the product path:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* ``tractor_amd`` never imports it and has no CPU fallback that routes here.

Every function cites the reference file:line it follows.  What pins each
function (reference known answers, closed forms, library routines, brute
force) is listed in ``tests/test_oracle_*.py``; functions without such a check
are marked "parity unpinned" in their docstrings.
"""
from .tractor_oracle import *  # noqa: F401,F403
