"""CPU oracle for the template-matching path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker (or as the timed CPU baseline). The product
package ``cistem_downstream_bah_amd`` never imports it.

How each oracle function is pinned (golden vectors, closed forms, brute force) is listed in the
header of ``tm_oracle.cpp``; every function has at least one such check in ``tests/``.
"""
from .oracle import *  # noqa: F401,F403
