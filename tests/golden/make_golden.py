"""Writes the small golden fixtures under tests/golden/ from the seeded
generators and the CPU oracle. Re-run only when the canonical semantics change
on purpose; the fixtures then freeze them for both the oracle and the HIP path.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from helio_amd import synth  # noqa: E402


def config0_small():
    sc = synth.instance_scene(4096, config=0, index=77, cube=150.0)
    cam = synth.perspective_view(synth.rng(0, 77), position=(75, 75, -5), direction=(0.1, 0.05, 1)).camera
    ov = oracle.view_from_camera(cam, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
    ids, args, stats = oracle.cull_view(ov, sc.bounds, sc.meta(), 0, sc.n, threads=1)
    np.savez_compressed(HERE / "config0_small.npz", bounds=sc.bounds, mesh_ids=sc.mesh_ids, flags=sc.flags,
                        mesh_lods=sc.mesh_lods.view(np.uint32).reshape(-1, 3), camera=cam.view(np.uint8),
                        ids=ids, args=args, stats=stats)


def hiz_small():
    g = np.random.default_rng(78)
    depth = g.random((45, 77), dtype=np.float32)
    np.savez_compressed(HERE / "hiz_small.npz", depth=depth, max_levels=oracle.hiz_build(depth, 0),
                        min_levels=oracle.hiz_build(depth, 1))


if __name__ == "__main__":
    config0_small()
    hiz_small()
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))
