"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on CPU."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

RAY_KEYS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("vk-rt-cuda-raytracing_amd")


@pytest.fixture(scope="session")
def scenes():
    return importlib.import_module("vk-rt-cuda-raytracing_amd.scenes")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def to_dev(rays: np.ndarray, dev):
    """float32 (8, n) SoA rays -> dict of device tensors."""
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(rays[i])).to(dev) for i, k in enumerate(RAY_KEYS)}


def hit_buffers(n, dev):
    import torch
    return dict(t=torch.empty(n, device=dev), prim_id=torch.empty(n, dtype=torch.int32, device=dev),
                u=torch.empty(n, device=dev), v=torch.empty(n, device=dev))
