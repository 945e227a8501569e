"""Frozen fixtures (tests/golden/*.npz): the oracle must keep reproducing them
(CPU), and the HIP path must equal them (GPU)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from tests.conftest import args_as_u32, canonical_ids

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def _cam(z):
    return np.frombuffer(z["camera"].tobytes(), L.CAMERA)


def test_oracle_reproduces_config0_golden():
    z = _load("config0_small.npz")
    meta = (z["mesh_ids"] | ((z["flags"] & 1) << 31)).astype(np.uint32)
    ml = np.ascontiguousarray(z["mesh_lods"]).view(L.MESH_LOD).reshape(-1)
    ov = oracle.view_from_camera(_cam(z), mesh_lods=ml, mesh_count=256)
    ids, args, stats = oracle.cull_view(ov, z["bounds"], meta, 0, len(meta), threads=3)
    assert np.array_equal(ids, z["ids"]) and np.array_equal(args, z["args"]) and np.array_equal(stats, z["stats"])
    assert len(ids) > 100


def test_oracle_reproduces_hiz_golden():
    z = _load("hiz_small.npz")
    assert np.array_equal(oracle.hiz_build(z["depth"], 0), z["max_levels"])
    assert np.array_equal(oracle.hiz_build(z["depth"], 1), z["min_levels"])


@pytest.mark.gpu
def test_gpu_matches_config0_golden(ctx):
    from helio_amd.passes import InstanceCullPass, Scene

    z = _load("config0_small.npz")
    scene = Scene(ctx, len(z["bounds"]))
    scene.upload_soa(z["bounds"], z["mesh_ids"], z["flags"])
    scene.set_mesh_lods(np.ascontiguousarray(z["mesh_lods"]).view(L.MESH_LOD).reshape(-1), 256, 4)
    p = InstanceCullPass(ctx)
    p.prepare(scene, _cam(z))
    p.execute()
    ids, args, stats = p.read()
    assert np.array_equal(stats, z["stats"])
    assert np.array_equal(args_as_u32(args), z["args"])
    assert np.array_equal(canonical_ids(ids, args), z["ids"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode,key", [(L.HIZ_MAX, "max_levels"), (L.HIZ_MIN, "min_levels")])
def test_gpu_matches_hiz_golden(ctx, mode, key):
    from helio_amd.passes import HiZBuildPass

    z = _load("hiz_small.npz")
    d = z["depth"]
    hiz = HiZBuildPass(ctx, d.shape[1], d.shape[0], mode)
    hiz.upload_depth(d)
    hiz.execute()
    assert np.array_equal(hiz.read_packed(), z[key])
