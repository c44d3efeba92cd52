"""CPU oracle for the periodic-box pair-counting path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy FP64, what the reference MeasureIA code
computes for ``MeasureIABox.measure_xi_w`` and ``measure_xi_multipoles``.  It is
the checker for the HIP path and the CPU baseline of the benchmark.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product package ``measure_ia_amd`` never does,
and never falls back to it.

Modules and the reference code each one follows (paths relative to the
reference repository root):

  prep.py   input preparation, responsivity, jackknife region ids, bin edges
            (src/measureia/measure_w_box_jk.py:97-155,
             src/measureia/measure_IA_base.py:155-167, 404-452)
  brute.py  O(N_p*N_s) pair counters, loop over positions, vectorised over
            shapes, with the reference's own op order and np.log10 binning
            (src/measureia/measure_w_box_jk.py:157-211,
             src/measureia/measure_m_box_jk.py:164-223,
             src/measureia/measure_w_box.py:144-184,
             src/measureia/measure_m_box.py:338-383)
  post.py   analytic randoms, estimators, w projection, multipoles,
            jackknife combination
            (src/measureia/measure_IA_base.py:229-351, 454-668,
             src/measureia/measure_w_box_jk.py:213-291)
  tree.py   KDTree + per-shape numpy + multiprocessing counter, the
            reference's fast CPU path, used as the timed CPU baseline
            (src/measureia/measure_w_box_jk.py:293-541, 543-857,
             src/measureia/measure_m_box_jk.py:306-568)

How each part is pinned (see tests/test_oracle_*.py):
  * post.py: against every relation the reference's shipped output fixtures
    hold (tests/data/processed/TNG300/*.hdf5 -> tests/golden/*.npz):
    RR closed forms (incl. the N_p-1 multipole quirk), xi = DD/RR - 1,
    xi_g+ = S+D/RR, w, multipoles, jackknife mean/std/cov -- bit-exact.
  * prep.jackknife_region_indices: the reference's known-answer test
    (tests/test_w_jk.py:16-23) plus boundary cases against the literal
    triple loop.
  * brute.py: brute force on hand-placed points with known (r_p, pi),
    e+ = +e, 0, -e geometry; independent per-pair Python-scalar loop on tiny
    inputs; uniform-random DD against the analytic RR; weight scaling
    (tests/test_weights.py:5-84); jackknife leave-one-region-out identity.
  * tree.py: equals brute.py on seeded inputs (DD exact, S+D 1e-12).
"""
