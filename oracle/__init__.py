"""CPU oracle for the closed-shell CCSD iteration.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker / CPU baseline.
The product package ``mpqc_amd`` never imports it.

Parity status of each function is listed in ``oracle/ccsd.py`` and
``oracle/diis.py`` headers.
"""
