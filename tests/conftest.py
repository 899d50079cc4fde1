import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("web-gs-renderer_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    return torch


@pytest.fixture(scope="session")
def ctx(torch_cuda):
    from gsr import lib
    c = lib.Context(0)
    yield c
    c.close()


# ---- device array helpers (uint32/uint64 travel as int32/int64 views) ----
def to_dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64)).cuda()
    return torch.from_numpy(a).cuda()


def to_host(t, dtype):
    a = t.detach().cpu().numpy()
    return a.view(dtype)


def empty_dev(torch, n, dtype):
    tmap = {np.uint32: torch.int32, np.uint64: torch.int64, np.float32: torch.float32, np.int32: torch.int32}
    return torch.zeros(max(int(n), 1), dtype=tmap[dtype], device="cuda")
