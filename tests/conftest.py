import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    from patchtable_amd import _ffi
    return _ffi.load()


def speed1(extra=None):
    """PatchTableParams() + set_speed(1) [+ extra switches]; the leading '-x 0' pair is the
    token set_from_switches(string) drops, as in the reference."""
    from patchtable_amd import PatchTableParams
    p = PatchTableParams()
    p.set_speed(1)
    if extra:
        p.set_from_switches("-x 0 " + extra)
    return p


@pytest.fixture(scope="session")
def pair_cfg1():
    from patchtable_amd import synth_pair
    return synth_pair(1, 300, 400)


def seeded(shape, seed):
    return np.random.default_rng(seed).random(shape, dtype=np.float32)
