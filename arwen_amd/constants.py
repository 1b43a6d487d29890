"""WRF physical constants used by the dycore step.

Values follow WRF ``share/module_model_constants.F`` as carried by the
reference (gpuwm/core/constants.py:24-40).  Derived constants that WRF folds
in float32 arithmetic are spelled as float32 chains so the device ``#define``
values agree bit for bit with the reference preamble
(gpuwm/core/kernels/__init__.py:52-56, which writes ``float(v)!r + "f"``).
"""
import numpy as np

G = 9.81
RD = 287.0
RV = 461.6
CP = 7.0 * RD / 2.0        # 1004.5
CV = CP - RD               # 717.5
P0 = 1.0e5
T0 = 300.0
GAMMA = CP / CV            # 1.4
RCP = RD / CP
RCV = RD / CV
# r_v / r_d as a float32 divide of float32 operands (WRF folds it that way).
RVOVRD = float(np.float32(np.float32(RV) / np.float32(RD)))

#: Constants handed to the HIP kernels (compile-time defines in csrc).
DEVICE_DEFINES = {"G": G, "RD": RD, "CP": CP, "CV": CV, "P0": P0,
                  "GAMMA": GAMMA, "RVOVRD": RVOVRD}
