"""CPU oracle for the gtrap hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg may import this package, and only as the
checker or the timed CPU baseline; the product (``gtrap_amd``) never imports
it.  Every function cites the reference file:line it restates.

Pinning status of each function is listed in the module docstrings and in the
test files that pin them (tests/test_oracle_*.py).
"""
