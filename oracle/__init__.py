"""FP64 CPU oracle for the 2D pair-HMM forward/backward path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker
(or, for ``cpu_baseline``, as the timed CPU restatement).

The oracle restates the reference algorithm in plain numpy (log space,
``NEG_INF = -1e30`` sentinels, the same emission log-floors) so that the
HIP kernels can be compared against it.  Each function cites the reference
file:line it follows (paths relative to the reference's ``python/``
directory).  ``oracle/c/fb2d_oracle.c`` is the same antidiagonal recursion in
C (used as the timed CPU baseline and cross-checked against the numpy
version).

How each part is pinned (see ``tests/test_oracle_*.py``):

* ``tkf``       -- closed forms (row sums, TKF92(ext=0) == TKF91, the
                   L'Hopital branch continuous across |1-kappa| = 1e-4 against
                   mpmath at 50 digits).
* ``mixdom``    -- row sums of chi, and chi == the null-eliminated exploded
                   machine built independently by ``mixdom.exploded_chi``
                   (brute-force power series of the null block).
* ``emissions`` -- against direct per-cell evaluation of the emission formula
                   (and C=1 per-class == per-domain).
* ``dp2d``      -- brute-force enumeration of every alignment path on tiny
                   pairs (``brute``), forward logP == backward logP, the exact
                   count identities (sum over M/D destinations == Lx, over
                   M/I destinations == Ly, start row sums to 1, flow
                   conservation), and finite differences
                   d logP / d log T[k,k'] == n[k,k'].
"""
