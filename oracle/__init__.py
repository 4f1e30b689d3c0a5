"""CPU oracle for the registration hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy (and plain C for the bitwise prescreen), the
formulas of the reference `jd_curl.three_d` hot path. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import
it, and only as the checker: the product (`seq_vp_3d_amd`) never imports,
links or calls anything here.

Modules and what pins each (see the docstring of every function for the
reference file:line it follows):

* `q1`        -- Q1 geometry, soft JD/curl energy, localized barrier, stable
                 differences, gradient, GGN diagonal, T62 node JD.
                 Pinned by closed forms (u=0, affine maps), central finite
                 differences (gradient) and brute-force unit-vector HVPs
                 (GGN diagonal) in tests/test_oracle_q1.py.
* `poisson`   -- DST-I Poisson inverse. Pinned by the independent operator
                 action (M M^-1 x = x) and a dense Kronecker solve.
* `jd_root`   -- T66 global first JD root. Pinned by np.roots on random
                 cubics and by the closed-form det along a line.
* `prescreen` -- T75 Bernstein prescreen (C, exact IEEE order, no FMA).
                 Pinned by exact rational arithmetic (fractions) on random and
                 near-degenerate cells: the bound must contain the exact
                 Bernstein minimum.
* `spline`    -- not-a-knot cubic tensor spline. Pinned by exact cubic
                 reproduction and node interpolation.
* `synthetic` -- the synthetic pair generator (blobs + sinusoidal modes).
* `certificate` -- the whole-cell interval certificate (edge hull, prescreen,
                 Bernstein fallback, map records). Pinned by exact rationals,
                 closed forms and the T05 folding cell.
* `preopt`    -- the FP32 pre-optimiser's objective, Adam/cap step and pyramid.
                 Pinned by finite differences and SciPy's own routines.
* `jd_affine` -- the T62 node JD measure and its T69 face/interior extension.
                 Pinned by quadratic bowls, L-BFGS-B and the Bernstein hull.
* `jd_step`   -- JD witnesses, the point enclosure and the historical-witness
                 step prediction. Pinned by exact rationals, known cubic roots
                 and brute-force scans of det(A + sV).
"""
