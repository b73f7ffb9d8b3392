"""TEST INFRASTRUCTURE ONLY -- the CPU oracle of the explicit SMA step.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
library (``alya_sma_amd``) never links or calls it.
"""
