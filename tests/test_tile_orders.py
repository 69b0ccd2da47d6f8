"""CPU test: the tile orders of the Ozaki product kernels (csrc/ozaki_tiles.hpp, host
and device code) visit every tile of their grid exactly once, over grids of 1 .. 300
tile rows, lower-triangular (64 x 64 tiles) and lower-trapezoidal (128 x 64 tiles with
as many, twice as many and fewer tile columns).  A tile visited twice or never would
leave part of C unwritten or written twice; tests/cpp/ozaki_tiles.cpp is the check."""
import os
import shutil
import subprocess

import pytest

from _tb import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "build", "ozaki_tiles")

pytestmark = pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                                reason="needs make and hipcc")


def test_ozaki_tile_orders_cover_every_tile_once():
    subprocess.run(["make", "-C", CPP, "tiles"], check=True, stdout=subprocess.DEVNULL)
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.strip().endswith("done 0"), p.stdout[-2000:]
