"""MI355X-native registration hot path of seq_vp_3d (jd_curl.three_d).

Device work runs in libseqvp_hip.so (hand-written gfx950 HIP kernels behind a C ABI, see
include/seqvp.h); the Python modules mirror the reference's operator interface:

* q1.FusedQ1 / q1.FusedJDRoot   -- Q1 energies, gradient, GGN diagonal, T62 node JD, T66 root
* poisson.DevicePoisson         -- DST-I Poisson metric inverse (FP64 / FP32 MFMA)
* spline.DeviceSpline           -- cubic B-spline sampler of the moving image
* prescreen                     -- T75 Bernstein prescreen (bitwise with the CPU kernel)
"""
from .problem import Grid3D, Q1Problem, Weights3D, make_grid_3d  # noqa: F401

__version__ = "0.1.0"
