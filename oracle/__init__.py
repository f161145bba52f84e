"""CPU oracle for the pressure-projection path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker / the timed CPU
baseline.  The shipped library never calls into it.
"""

from .oracle import *  # noqa: F401,F403
