import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    return torch.device("cuda:0")


def _reference_study_input(N, batch, K_values):
    """The inputs of the reference's precision study (test_precision_accuracy.cu:205-215):
    std::mt19937(42), uniform_real_distribution<float>(-0.5, 0.5), fresh interleaved FP16
    data [batch][N][K][2] for each K in turn from the one generator.  libstdc++ draws one
    32-bit word per float: r = float(u) / 2^32 (1 -> the float below 1), x = 1 * r - 0.5
    in float, then FP16 nearest even."""
    import numpy as np
    bg = np.random.MT19937(0)
    bg._legacy_seeding(42)
    out = {}
    for K in K_values:
        n = 2 * N * K * batch
        u = bg.random_raw(n).astype(np.uint32)
        r = u.astype(np.float32) / np.float32(2.0 ** 32)
        r = np.where(r >= 1, np.nextafter(np.float32(1), np.float32(0)), r).astype(np.float32)
        out[K] = (r + np.float32(-0.5)).astype(np.float32).astype(np.float16)
    return out


@pytest.fixture(scope="session")
def reference_study_input():
    return _reference_study_input
