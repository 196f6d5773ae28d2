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
"""
