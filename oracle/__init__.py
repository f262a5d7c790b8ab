"""CPU oracle for the transverse-H Maxwell eigensolver path (NumPy/SciPy, FP64).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``eigenfns_amd``)
imports, links or runs anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker or as the timed CPU baseline, never as the thing
measured or shipped.

Each module restates one part of the reference algorithm and cites the
reference file:line it follows (paths are relative to the reference checkout):

  operator_np   eigenfns/operator.py:44-184      make_basis, Θ, preconditioners
  linalg_np     eigenfns/solver.py:35-126        gram, combine, deflate, SVQB
  lobpcg_np     eigenfns/solver.py:136-382       deflated block LOBPCG
  chebyshev_np  eigenfns/chebyshev.py:28-64,213-267  Lanczos λ_max, KPM count
  interior_np   eigenfns/interior.py:37-88,312-427  bandpass coefficients, filter, RR
  kpm_np        scripts/exp/exp_kpm_dos.py:125-163, exp_kpm_analyze.py:22-70
  structure_np  eigenfns/structure.py:29-151     pen-like rasterizer
  audit_np      scripts/exp/exp_i2_v2.py:129-152 deflated-probe completeness audit (I2)
  dense_np      tests/test_operator.py:16-24     dense Θ matrix by probing
  rng_np        (no reference counterpart)       counter-based RNG shared with the HIP kernels

Pinning (what fixes each function independently of its own formula) is listed
in each module header; every function is pinned by at least one test in
``tests/test_oracle_*.py``.
"""
