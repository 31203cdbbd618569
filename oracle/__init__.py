"""CPU oracle for the halo2-axiom-gpu hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline. The product path (``halo2-gpu_amd``) never imports it.

* ``pyoracle`` - pure-Python big-integer restatement for tiny inputs (naive DFT,
  affine double-and-add, closed forms). It pins the C oracle.
* ``coracle``  - ctypes binding of ``libhalo2_oracle.so`` (plain C restatement of
  the reference CPU algorithms, multithreaded) for sizes up to 2^24.
"""
