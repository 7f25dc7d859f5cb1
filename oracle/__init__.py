"""CPU oracle for the HEALPix spherical-harmonic-transform hot path.

TEST INFRASTRUCTURE ONLY.  Nothing under ``oracle/`` is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product path (``hellblast_amd``) never calls into it.

Modules
-------
geometry   HEALPix RING tables (ring_tables_healpix.cu:12-51, ring_tables.py:17-53)
legendre   normalised, exponent-scaled lambda_lm recursion (legendre.cu:2-35,
           sht_raw.py:26-28) and the literal unnormalised recursion
sht        FP64 analysis/synthesis/alm2cl/adjoints in the sht_fallback +
           isht convention (sht_fallback.py:36-65, isht.py:5-26, alm_ops.py:19-32)
spin       spin-weighted (s=1,2) transforms in the healpy E/B convention
literal    literal NumPy restatements of sht_raw/sht_fallback/isht for small lmax
"""
