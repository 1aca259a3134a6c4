"""CPU oracle for the template-matching path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker (or as the timed CPU baseline). The product
package ``cistem_downstream_bah_amd`` never imports it.

Parity status of each oracle function (how it is pinned) is listed in ``oracle/PINNING.md``.
"""
from .oracle import *  # noqa: F401,F403
