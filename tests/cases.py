"""Frames shared by the tests.  Coordinates come from the reference's own test
locations (src/engine/test_cases.h, src/tools/verify.cpp:11789-11881) and are stored
as data in tests/golden/locations.json."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LOC = json.load(open(os.path.join(GOLDEN, "locations.json")))


def pow10(k):
    return "1" + "0" * k


def shift_decimal(x, shift):
    """x + shift for decimal strings, exact (Python's decimal module at the length of x)."""
    from decimal import Decimal, getcontext

    getcontext().prec = len(x) + 50
    return format(Decimal(x) + Decimal(shift), "f")


def golden(name, H, W):
    return np.fromfile(os.path.join(GOLDEN, name), np.float32).reshape(H, W)


# name -> (cre, cim, scale, W, H, mxit, c_method)
FRAMES = {
    # verify.cpp:11838-11843 (golden minibrot875.f32)
    "minibrot875": (LOC["deep1_x"], LOC["deep1_y"], pow10(875), 24, 18, 300000, 0),
    # verify.cpp:11851-11857 (golden minibrot875_extref.f32)
    "extref875": (shift_decimal(LOC["deep1_x"], "2.2e-875"), LOC["deep1_y"], pow10(875), 24, 18, 300000, 0),
    # verify.cpp:11805: Misiurewicz c = i at 1e1000, whole frame exterior
    "exterior1000": ("0", "1", pow10(1000), 120, 80, 10000, 0),
    # the same with the exterior distance estimate
    "exterior1000_ede": ("0", "1", pow10(1000), 64, 48, 10000, 2),
    # verify.cpp:11816: deep1 needle at 1e876 (reduced frame)
    "deep876": (LOC["deep1_x"], LOC["deep1_y"], pow10(876), 48, 32, 300000, 0),
    # verify.cpp:11827: sub-pixel nucleus at 1e737, all exterior
    "subpixel737": (LOC["deep1_x"], LOC["deep1_y"], pow10(737), 48, 36, 60000, 0),
    "subpixel737_ede": (LOC["deep1_x"], LOC["deep1_y"], pow10(737), 48, 36, 60000, 2),
    # the resolved nucleus frame with the exterior distance estimate
    "minibrot875_ede": (LOC["deep1_x"], LOC["deep1_y"], pow10(875), 24, 18, 300000, 2),
    # odd sizes and a strip of a taller frame (strip API, mandel_perturbation.cpp:1399-1424)
    "minibrot875_odd": (LOC["deep1_x"], LOC["deep1_y"], pow10(875), 37, 23, 300000, 0),
}
