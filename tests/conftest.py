import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "dda-fft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libddafft_amd.so)")


@pytest.fixture(scope="session")
def orc32():
    import oracle

    return oracle.Oracle("f32")


@pytest.fixture(scope="session")
def orc64():
    import oracle

    return oracle.Oracle("f64")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def bits(a):
    """View of an array as raw integers, for bit-exact comparisons."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype in (np.complex64, np.float32) else np.uint64)


def directions(theta_deg, plane_deg=0.0):
    """Scattering directions of one plane, computed as main.cpp:354-387 does."""
    th = np.asarray(theta_deg, np.float32).astype(np.float64) * np.pi / 180.0
    d = np.stack([np.sin(th), 0 * th, np.cos(th)], 1).astype(np.float32)
    if plane_deg:
        c = np.float32(np.cos(np.float32(plane_deg) * (np.pi / 180.0)))
        s = np.float32(np.sin(np.float32(plane_deg) * (np.pi / 180.0)))
        x = c * d[:, 0] + s * d[:, 1]
        y = -s * d[:, 0] + c * d[:, 1]
        d = np.stack([x, y, d[:, 2]], 1).astype(np.float32)
    return d
