"""CPU oracle for the apyib_amd hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy float64/complex128, the reference
algorithms of apyib that the GPU path replaces.  Every function cites the
reference file:line it follows.  It exists to check the GPU path: only the
tests (tests/), ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU
baseline.  Nothing in the ``apyib_amd`` package imports it; the product path
raises if its HIP library is missing.

Modules
  so        spin-orbital CISD / CID / MP2 and DIIS     (ci_wfn.py, mp2_wfn.py, utils.py)
  spatial   spin-adapted CISD / CID / MP2               (ci_wfn.py, mp2_wfn.py)
  transforms AO->MO, MO->SO, overlaps, phase            (utils.py)
  brute     determinant-basis CI by Slater-Condon rules (independent pin)
  davidson  lowest-root Davidson on the CISD operator   (pin for the GPU Davidson)

Parity status of each function (how it is pinned) is listed in each module
header.
"""
