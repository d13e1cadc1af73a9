"""CPU oracle for the B-spline interpolation path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results or as the
timed CPU baseline.  The product (``bsplineinterpolation_amd``) never does.
"""
from .oracle import Oracle, OracleSpline, RefSpline, ref_available, build_oracle  # noqa: F401
