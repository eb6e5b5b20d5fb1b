"""CPU oracle for the TFHE PBS path -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package, and only as the checker.  The product (``luthe_amd`` and
``libluthe.so``) never imports, links or calls it.
"""
