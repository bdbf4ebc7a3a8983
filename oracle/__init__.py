"""CPU oracle for CENTRE's entropy hot path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, the smoke check and the
CPU-baseline leg of ``bench.py``.  The shipped library never calls it.
"""
from .oracle import *  # noqa: F401,F403
