"""Velocity sets and default parameters (oracle side, FP64).

Follows direbm/constants.py:10-52 (D2Q7: rest + unit hexagon, W = [1/2, 1/12 x 6], cs^2 = 1/4,
DX = 1, ALPHA = 4, KAPPA_HARD = 4) and direbm/lattices.py:44-61 (D3Q13 icosahedral: rest + 12
normalised (0, +-1, +-phi) cyclic vertices, W = [2/5, 1/20 x 12], cs^2 = 1/5).
"""

from __future__ import annotations

import numpy as np

_S = np.sqrt(3.0) / 2.0

# D2Q7: direction order follows direbm/constants.py:16-27.
C2 = np.array(
    [[0.0, 0.0], [1.0, 0.0], [0.5, _S], [-0.5, _S], [-1.0, 0.0], [-0.5, -_S], [0.5, -_S]],
    dtype=np.float64,
)
W2 = np.array([0.5] + [1.0 / 12.0] * 6, dtype=np.float64)
CS2_2D = 0.25

# D3Q13 icosahedral: vertex order follows direbm/lattices.py:46-55.
_PHI = (1.0 + np.sqrt(5.0)) / 2.0
_ICO = np.array(
    [
        [0, 1, _PHI], [0, 1, -_PHI], [0, -1, _PHI], [0, -1, -_PHI],
        [1, _PHI, 0], [1, -_PHI, 0], [-1, _PHI, 0], [-1, -_PHI, 0],
        [_PHI, 0, 1], [_PHI, 0, -1], [-_PHI, 0, 1], [-_PHI, 0, -1],
    ],
    dtype=np.float64,
)
_ICO = _ICO / np.linalg.norm(_ICO[0])
C3 = np.vstack([np.zeros(3), _ICO])
W3 = np.array([0.4] + [0.05] * 12, dtype=np.float64)
CS2_3D = 0.2

DX = 1.0
ALPHA = 4.0
KAPPA_HARD = 4
KAPPA_SOFT = 5
DEFAULT_TAU = 0.6

# Axial (di, dj) index shifts of the hex LBM, direbm/lbm.py:22 (same as direbm/warp/lbm.py:19-20).
HEX_OFFSETS = [(0, 0), (1, 0), (0, 1), (-1, 1), (-1, 0), (0, -1), (1, -1)]

# Gap-fill emptiness threshold of the GPU path (direbm/warp/propagation.py:243).
EPS_EMPTY = 1e-6
