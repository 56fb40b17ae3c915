import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hddm-solver-postprocessors_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def kat():
    import json
    import hddm_lb
    d = json.load(open(os.path.join(GOLDEN, "interpolate_array_kat.json")))
    st = hddm_lb.read_plt(os.path.join(GOLDEN, "surplus.plt"), d["nagents"])
    d["state"] = st
    return d


def tiny_grid(dim, max_level, seed, target=0, p=0.5, start=1):
    import hddm_lb
    ii, jj, lv, ix = hddm_lb.generate_grid(dim, min(start, max_level - 1), max_level, p, seed, target)
    return ii, jj, lv, ix
