"""Output files written at DEM's output points from the device state are
byte for byte those written from the oracle's state."""
import filecmp
import glob
import os

import pytest

from parmec_amd import parmec as P
from parmec_amd import pmx

pytestmark = pytest.mark.gpu


def scene():
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    k = 0
    for i in range(6):
        for j in range(5):
            P.SPHERE((0.1 + 0.21 * i, 0.1 + 0.21 * j, 0.3 + 0.05 * ((i + j) % 3)), 0.1, mat, 1)
            k += 1
    P.OBSTACLE([(-2.0, -2.0, 0.0, 3.0, -2.0, 0.0, -2.0, 3.0, 0.0)], 3)
    P.SPRING(0, (0.1, 0.1, 0.3), 7, (0.31, 0.31, 0.35), [-1.0, -1e3, 1.0, 1e3], dashpot=0.1)
    P.GRANULAR(0, 0, 1e5, 0.2, 0.3)
    P.GRAVITY(0.0, 0.0, -9.81)
    P.OUTPUT(format=["DUMP", "VTK"])
    P.OUTPUT(entities=["NUMBER", "LINVEL", "FORCE"], subset=[3, 4], mode=["RB", "SPH"])


def test_output_files_match_oracle(tmp_path):
    from oracle import orc
    print("case output files, 30 spheres, 2 calls", flush=True)
    dirs = []
    for name, make in (("gpu", lambda: pmx.Engine()), ("orc", lambda: orc.Oracle(0))):
        d = os.path.join(str(tmp_path), name)
        os.makedirs(d)
        P.RESET()
        scene()
        P._use_engine(make())
        P.DEM(300 * 1e-4 - 0.5e-4, 1e-4, 25e-4, prefix=os.path.join(d, "out"))
        P.DEM(200 * 1e-4 - 0.5e-4, 1e-4, 25e-4, prefix=os.path.join(d, "out"))
        P.RESET()
        dirs.append(d)
    a = sorted(os.path.basename(f) for f in glob.glob(os.path.join(dirs[0], "*")))
    b = sorted(os.path.basename(f) for f in glob.glob(os.path.join(dirs[1], "*")))
    assert a == b and len(a) > 40
    match, mismatch, errors = filecmp.cmpfiles(dirs[0], dirs[1], a, shallow=False)
    assert not mismatch and not errors, mismatch[:5]
