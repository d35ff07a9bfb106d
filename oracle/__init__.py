"""CPU oracle for the orbital-free DFT hot path: TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 numpy restatement of the reference algorithm
(DFTpy, shaoxc/dftpy).  It exists to check the HIP product path; nothing in
``dftpy_amd`` imports it, and only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may call it -- always as the checker or the
timed CPU baseline, never as the thing measured or shipped.

Each function cites the reference file:line it restates (paths relative to
the reference's ``src/dftpy``).  How each part is pinned:

* grid (``grid.py``) -- closed forms (B = 2*pi*A^-T, QE lattices), tests.
* FFT conventions (``field.py:373-390,739-764``) -- analytic Gaussian pair,
  brute-force DFT on tiny 2/3/5 grids, round trip, Parseval.
* TF / LDA -- closed forms, PZ continuity at rs=1, V = dE/drho by finite
  differences.
* vW -- uniform density gives zero; analytic Gaussian sqrt(rho).
* WT (+ Lindhard kernel) -- golden 2.916818700014412 on Al_fde_rho.pp
  (``examples/test/test_functional.py:27``), Lindhard limits.
* Hartree -- Parseval form and a Gaussian closed form.
* local PP -- golden 1.3497046 (``examples/test/test_read_pp.py:12,31``),
  PME vs exact structure factor.
* Ewald -- goldens of ``examples/test/test_ewald.py:33-35``.
* optimizer + line search -- GaAs_random TN total energy -43.3046
  (``examples/test/test_density_opt.py:46``); dcsrch step sequence against
  scipy's own Moré-Thuente port on scalar functions.
* TD Crank-Nicolson -- dipole 0.011423 after 10 CN steps
  (``examples/test/test_propagator.py:81``), norm conservation.
* LMGP -- golden 2.9145872498710492 at atol 1e-3 only
  (``examples/test/test_functional.py:34``).
"""
