"""Test infrastructure only: the CPU oracle for the PROFESS energy/potential path.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline). The product (``profess_amd``) never imports it and fails loudly
when its HIP library is missing instead of falling back to anything here.

Contents
  profess_oracle  numpy restatement of CalculatePotentialPlus/CalculateEnergy,
                  the per-geometry tables and the NTN optimizer (file:line
                  citations into the reference on every function).
  refbuild        loader for ``oracle/_ref/libprofess_ref.so``: the reference's
                  own self-contained Fortran files (RKSUITE, DCSRCH/DCSTEP,
                  cubic splines) compiled from their sources by ``oracle/ref``.
"""
