"""CPU oracle for the non-Cartesian reconstruction path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``hastycucompute_amd``) imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker or as the timed CPU baseline.

Contents (every function cites the reference file:line whose behaviour it
restates; paths are relative to the HastyCuCompute source tree):

* ``ndft``      -- FP64 direct non-uniform DFT, type 1 and type 2, with the
                   sign / centred-mode conventions of ``nufft.cppm:42-55`` and
                   the coordinate pairing of ``nufft.cppm:193-199``.
* ``kb``        -- Kaiser-Bessel gridding NUFFT (spread / FFT / deapodise)
                   restated in numpy FP64 plus an OpenMP C spreader
                   (``kbspread.c``) used for the timed CPU baseline.
* ``toeplitz``  -- PSF kernel (``nufft.cppm:386-441``), ``toeplitz_multiply``
                   (``nufft.cppm:448-479``) and the elementwise semantics of
                   ``toeplitz_multiplication`` (``toeplitz_load_3D.cu:18-79``,
                   ``toeplitz.cppm:365-501``).
* ``solvers``   -- SENSE normal operator (``normal.cppm:65-88``), CG
                   (``cg.cppm:122-184``) and power iteration (``alg.cppm:13-30``).

How each function is pinned (closed forms, invariants, brute force) is listed
in ``tests/test_oracle.py``; the NUFFT arithmetic itself lives in an external
library (cufinufft, unpinned ``master`` tag) that no reference test checks, so
the NDFT definition is the anchor.
"""
