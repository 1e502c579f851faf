"""Constants of the reference's public surface (constants.py:5-39)."""
import math

ARCMIN_PER_DEGREE = 60
INVERSE_LENSE = -1
FORWARD_LENSE = +1
NSIDE_DEFAULT = 256
THETA_PIX_DEFAULT = 2.5
PIX_WIDTH_DEFAULT = math.radians(THETA_PIX_DEFAULT / ARCMIN_PER_DEGREE)
DEFAULT_NSTEPS = 10  # lense_flow.py:289 default n


class Parametrization:
    T = 0
    QU = 1
    EB = 2


class Basis:
    MAP = 0
    FOURIER = 1
