"""CPU oracle for the IPC Newton step -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker or the CPU baseline.
The product (``fastipc_amd``) never imports it.
"""
