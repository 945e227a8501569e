"""Shared fixtures. GPU tests are marked ``gpu`` and call libhv.so through its C ABI."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhv.so kernels)")


@pytest.fixture(scope="session")
def ctx():
    from helio_amd.passes import Context

    c = Context(0)
    yield c
    c.close()


def canonical_ids(ids: np.ndarray, args: np.ndarray) -> np.ndarray:
    """Sort ids inside each bucket range (the order inside a bucket is unspecified)."""
    out = ids.copy()
    first = np.asarray(args["first_instance"] if args.dtype.names else args[:, 4], np.int64)
    cnt = np.asarray(args["instance_count"] if args.dtype.names else args[:, 1], np.int64)
    for f, c in zip(first, cnt):
        if c > 1:
            out[f:f + c] = np.sort(out[f:f + c])
    return out


def args_as_u32(args: np.ndarray) -> np.ndarray:
    """Structured DRAW_INDEXED_INDIRECT array -> (n, 5) u32 view."""
    return np.ascontiguousarray(args).view(np.uint32).reshape(-1, 5)
