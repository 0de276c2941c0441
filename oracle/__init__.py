"""CPU oracle for the pyofss fibre stepping path.

TEST INFRASTRUCTURE ONLY. This package is a NumPy restatement of the
reference's arithmetic (KarinaSaraeva/pyofssChanged, ``pyofss/``), written so
that the HIP engine in ``pyofsschanged_amd`` can be checked against it. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline. The product package never imports anything from here.

Every function cites the reference file:line it restates. Parity pins for
each part are listed in ``oracle/pyofss_np.py``'s header; the parts that no
pin covers are marked "parity unpinned" there.
"""

from .pyofss_np import *  # noqa: F401,F403
