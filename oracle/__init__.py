"""CPU oracle for the diffractive transfer-matrix path.

TEST INFRASTRUCTURE ONLY.  This package is a complex128 NumPy restatement of
the reference algorithm (IJSComplexMatter/dtmm, ``transfer_field`` ->
``transfer_4x4`` / ``transfer_2x2`` -> ``propagate_*_effective_1``).  It is
imported only by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, where it acts as the checker (or as the
timed CPU baseline).  The product package ``dtmm_amd`` never imports it and
has no CPU fallback.

Every function cites the reference file:line it follows.  The pinning of each
function (golden vectors, closed forms, invariants) lives in
``tests/test_oracle_*.py``.  Functions without such a check are marked
"parity unpinned" in their docstring.
"""

from .dtmm_oracle import *  # noqa: F401,F403
