"""CPU oracle for the ADRT oblique inverses — TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy, the reference algorithms of dsrim/adrt-pinv that the HIP engine
replaces: the ADRT levels and their transpose (adrtcore.py), the per-frequency butterflies and the Chebyshev
row filter (fastspec.py, fasthybrid.py, cpp/oblique.cpp), the row engine and its host orchestration
(fastops.py), the spectral bounds (fastcpp.py) and CGLS (pinv.py).  Every function cites the reference
file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this
package, and only as the checker (or the timed CPU baseline).  The product path (the package next to this
directory) never imports it and fails loudly when the HIP library is missing.

Parity pins (what each module is checked against, see ``tests/test_oracle.py``):
  radon    - digital-line closed form, transpose identity, brute-force dense matrix, golden fixtures.
  rows     - closed form A_k(s,t) = exp(i w_k delta_s(t)), adjointness, eigenvector test of the filter,
             dense eigh of the level Gram, golden fixtures of RowEngine.run.
  bounds   - lambda_max(M_0) = N^2, cert >= true row sum, golden fixtures of the C++ power iteration.
  inverses - L R x = x invariant, G1 spectrum / dense closed form, packing = per-quadrant operators,
             golden end-to-end fixtures of FastLp / FastHybrid at fixed setup scalars.
  lsq      - dense pseudoinverse for n <= 4.
  layout   - a stand-in package layout that is a known permutation of the core's is recovered exactly (the adrt
             package is not installed, so its own layout is unpinned).
"""
