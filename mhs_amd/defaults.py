"""Every threshold of the assembled path, in one module that imports nothing.

Values are the reference's (src/mhs/defaults.py:23-148 and
src/mhs/fullspace/defaults.py:20-37); the GPU kernels receive them through the
launch descriptor, so the host and device cannot disagree.
"""

# host ceiling (defaults.py:23-26)
MATRIX_MAX_BYTES = 8 * 1024**3
MATRIX_RAM_FRACTION = 0.5

# full-space hybrid producer (fullspace/defaults.py:20-23)
D_STAR = 10.0
D_STAR_DISTANT = 40.0
FAR_GAUSS_N = 12
FAR_GAUSS_N_DISTANT = 8

# edge primitives (fullspace/defaults.py:26-33)
SERIES_U_OVER_RHO = 2.0
SMALL_U_OVER_RHO = 0.5
SERIES_TERMS = 32

# image Q-family budget law (defaults.py:60-66)
IMAGE_Q_BUDGET_C = 10.0
IMAGE_Q_GAUSS_MIN = 8
IMAGE_Q_GAUSS_MAX = 192

# continuous layout weld (defaults.py:91)
WELD_TOL_OVER_H = 1.0e-9

# fold (defaults.py:108-148)
FOLD_SOURCE_DEPTH_EPS = 3.0
FOLD_BAND_DEPTH_EPS = 10.0
FOLD_REACH_EPS = 60.0
FOLD_N_RADIAL = 6
FOLD_N_ANGULAR = 4
FOLD_CREASE_MARGIN_EPS = 5.0
FOLD_FINE_SPACING_EPS = 0.25
FOLD_MEDIUM_SPACING_EPS = 2.0
FOLD_N_COARSE = 24
FOLD_FINE_SUBPANEL_EPS = 6.0
FOLD_MEDIUM_SUBPANEL_EPS = 16.0
FOLD_R1_MARGIN_DIAM = 8.0

# float32 out= storage tolerance (defaults.py:40-43)
OUT_FLOAT32_PARITY = 1.0e-6

# block drivers (parallel.py): the reference's in-flight budget of its worker
# pool (defaults.py:82), and the result size of one device block here
PARALLEL_INFLIGHT_BYTES = 512 << 20
DEVICE_BLOCK_BYTES = 2 << 30
