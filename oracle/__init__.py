"""CPU oracle for the per-substep slurry-plus-media step (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker (or as the timed CPU reference path). The product
(``milldynamics3_amd``) never imports it.
"""
from .oracle import OracleSim, default_params, flatten_params, oracle_lib  # noqa: F401
