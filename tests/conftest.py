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


def check_cull_invariants(ids: np.ndarray, args: np.ndarray, stats: np.ndarray) -> None:
    """The compaction contract of one view, independent of any reference:
    bucket counts sum to the ids, first_instance is their exclusive prefix,
    stats[0] = stats[1] + stats[2] + stats[3], and the ids are the drawn
    candidates that the Hi-Z test kept (stats[3] - stats[4])."""
    a = args_as_u32(args) if args.dtype.names else np.asarray(args)
    cnt = a[:, 1].astype(np.int64)
    assert cnt.sum() == len(ids)
    assert np.array_equal(a[:, 4].astype(np.int64), np.concatenate([[0], np.cumsum(cnt)[:-1]]))
    assert int(stats[0]) == int(stats[1]) + int(stats[2]) + int(stats[3])
    assert len(ids) == int(stats[3]) - int(stats[4])
