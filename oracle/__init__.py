"""CPU oracle for the TFCE_mediation permutation hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  The product package
``tfce_mediation_amd`` never imports it, and its GPU path fails loudly instead
of falling back here.

Contents
--------
* ``tfce``     ctypes front end of ``tfce_oracle.c`` (C restatement of
               ``lib/fast_tfce.hpp:11-95``) and of the reference build in
               ``oracle/_ref`` when it exists.
* ``ols``      FP64 numpy restatement of ``cynumstats.pyx:28-74`` (tval_int,
               calc_beta_se, resid_covars, calcF) and ``pyfunc.py:130-162``
               (calc_sobelz).
* ``fwe``      restatement of the FWER writers (``calculate_fweP_vertex.py:37-75``,
               ``calculate_fweP_voxel.py:23-57``, ``tm_func.py:403-573``,
               ``pyfunc.py:107-126``).
* ``graphs``   adjacency builders restated from ``pyfunc.py:37-76`` and
               ``tm_func.py:502-562`` plus brute-force BFS components.
* ``imgreg``   the element-wise regressor and VIF image
               (``pyfunc.py:264-271, 1024-1114``).

Pins (each checked by a test that shares no code with the function)
-------------------------------------------------------------------
* ``tfce.tfce``          the reference header itself built in ``oracle/_ref``
                         (bit for bit), brute-force BFS partitions and closed
                         forms (``tests/test_oracle_tfce.py``).
* ``ols.tval_int`` / ``perm_t`` / ``calc_beta_se`` / ``resid_covars``
                         QR, ``scipy.stats.linregress``, lstsq, invariances.
* ``ols.calcF``          F = t^2 of ``linregress`` at k = 2; F from the R^2 of
                         ``np.linalg.lstsq``; the grand-mean TSS offset.
* ``ols.calc_sobelz``    closed forms of the three z variants, symmetric path b.
* ``fwe.find_nearest*`` / ``fwep_map``   hand-made nulls with literal answers.
* ``fwe.apply_mfwer``    one surface: standardised-log nearest lookup; three
                         surfaces with log-mask-size weights: per-permutation
                         winners worked out by hand (``tests/test_oracle_fwe.py``).
* ``graphs.create_adjac_vertex`` / ``create_adjac_voxel``   hand-derived lists
                         (octahedron; 3x3x3 cube, 4x4x4 cube less one voxel,
                         6- and 26-connectivity) and a coordinate brute force
                         (``tests/test_adjacency.py``).
* ``mlp``                central finite differences (``tests/test_oracle_mlp.py``).
* ``imgreg.image_regression``   ``ols.tval_int`` when every element has the same
                         regressor; Frisch-Waugh-Lovell with QR least squares
                         (``tests/test_oracle_imgreg.py``).
* ``imgreg.image_reg_VIF``      1 / (1 - r^2) of ``scipy.stats.linregress``.
"""
