"""numpy views of the C-ABI records declared in include/localizer_hip.h.

``CANDIDATE_DTYPE`` mirrors ``LzrCandidate`` (the reference ``Particle``,
source/Storage.h:90-101, plus frame / window / model) and ``FIT_DTYPE`` mirrors
``LzrFit`` (the fields of ``LocalizedPosition_2DGauss`` and
``LocalizedPosition_Ellipsoidal2DGauss``, source/Storage.h:117-186).
"""
import numpy as np

MODEL_SYMMETRIC = 0
MODEL_ELLIPSOIDAL = 1
MODEL_ELLIPSOIDAL_SYMMETRIC_PSF = 2
MODEL_FIXED_WIDTH = 3

MAX_HALF_WIDTH = 15
MAX_PARAMS = 7

ACCEPTED = 0
REJECT_EDGE = 1
REJECT_AMPLITUDE = 2
REJECT_POSITION = 3
REJECT_WIDTH = 4
REJECT_EIGEN = 5
REJECT_ANISOTROPIC = 6
REJECT_PSF_WIDTH = 7
REJECT_INVALID = 8
REJECT_DISTANCE = 9

LM_SUCCESS = 0
LM_CONTINUE = -2
LM_FAILURE = -1
LM_ENOPROG = 27
LM_ETOLF = 29
LM_ETOLX = 30
LM_ETOLG = 31

OPT_MAX_ITERATIONS = 1
OPT_EPSABS = 2
OPT_EPSREL = 3
OPT_DIAG_UPDATE_ROWS = 4
OPT_WAVES_PER_CU = 5
OPT_MAX_HALF_WIDTH = 6
OPT_OCCUPANCY = 7
OPT_ROUNDS = 8
OPT_PARK_THRESHOLD = 9
OPT_STRAGGLER_TRIALS = 10
OPT_DEFAULT_HALF_WIDTH = 11
OPT_STRAGGLER_GROUP = 12
OPT_NET_MAX_INTENSITY = 13

CANDIDATE_DTYPE = np.dtype(
    [
        ("x", "<f8"),
        ("y", "<f8"),
        ("intensity", "<f8"),
        ("background", "<f8"),
        ("frame", "<i4"),
        ("half_width", "<i2"),
        ("model", "<i2"),
    ],
    align=True,
)

FIT_DTYPE = np.dtype(
    [
        ("params", "<f8", (MAX_PARAMS,)),
        ("deviations", "<f8", (MAX_PARAMS,)),
        ("integral", "<f8"),
        ("integral_deviation", "<f8"),
        ("sd1", "<f8"),
        ("sd2", "<f8"),
        ("theta", "<f8"),
        ("sd1_deviation", "<f8"),
        ("sd2_deviation", "<f8"),
        ("chi", "<f8"),
        ("status", "<i4"),
        ("lm_status", "<i4"),
        ("iterations", "<i2"),
        ("trials", "<i2"),
        ("accepted_steps", "<i2"),
        ("model", "<i2"),
    ],
    align=True,
)

assert CANDIDATE_DTYPE.itemsize == 40
assert FIT_DTYPE.itemsize == 192

# LzrStart: an LM start point (width > 0), the reference start (0) or skip (< 0).
START_DTYPE = np.dtype(
    [("amplitude", "<f8"), ("width", "<f8"), ("x", "<f8"), ("y", "<f8"), ("background", "<f8")]
)
assert START_DTYPE.itemsize == 40

# Learned initializer (include/localizer_hip.h, LZR_NET_*).
NET_HALF_WIDTH = 5
NET_OUTPUTS = 5
NET_OFF = {
    "conv1_w": (0, (32, 3, 3)),
    "conv1_b": (288, (32,)),
    "conv2_w": (320, (32, 32, 3, 3)),
    "conv2_b": (9536, (32,)),
    "fc1_w": (9568, (64, 1568)),
    "fc1_b": (109920, (64,)),
    "fc2_w": (109984, (5, 64)),
    "fc2_b": (110304, (5,)),
}
NET_NPARAMS = 110309

# Truth records written by the synthetic generator (localizer_amd/csrc/lzr_synth.h).
EMITTER_DTYPE = np.dtype(
    [
        ("x", "<f8"),
        ("y", "<f8"),
        ("photons", "<f8"),
        ("sigma_x", "<f8"),
        ("sigma_y", "<f8"),
        ("amplitude", "<f8"),
        ("background", "<f8"),
        ("frame", "<i4"),
        ("model", "<i2"),
        ("half_width", "<i2"),
    ],
    align=True,
)


def n_params(model):
    """Number of fitted parameters of a model code."""
    m = int(model)
    return 5 if m == MODEL_SYMMETRIC else (4 if m == MODEL_FIXED_WIDTH else 7)


def default_half_width(initial_psf_width):
    """The reference window: cutoff_radius = ceil(4 * initialPSFWidth)
    (source/Localization.h:68)."""
    return int(np.ceil(initial_psf_width * 4.0))

# Localization bitmap (include/localizer_hip.h, LzrPosition)
POSITION_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("integral", "<f8"), ("x_deviation", "<f8"),
                           ("background", "<f8")])
assert POSITION_DTYPE.itemsize == 40
DEVIATION_CONSTANT = 0
DEVIATION_FIT_UNCERTAINTY = 1
DEVIATION_GAUSSIAN_MASK = 2
WEIGHING_SAME = 0
WEIGHING_INTEGRAL = 1
