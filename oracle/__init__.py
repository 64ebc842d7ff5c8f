"""CPU oracle of the stepper -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
(as the checker or the timed CPU baseline), never by the product path.
"""
from .oracle import Oracle, load, orc32  # noqa: F401
