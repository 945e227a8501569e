"""Layout KATs of the GPU-shared structs (reference layout tests, libhelio)."""
from __future__ import annotations

import ctypes as C

import pytest

from helio_amd import _lib
from helio_amd import layouts as L

# Sizes asserted by the reference: crates/libhelio/src/meshlet.rs:131-166,
# crates/helio-pass-virtual-geometry/src/lib.rs:295-298,
# crates/helio-pass-hiz/tests/hiz_tests.rs (HiZUniforms 16 B), and the struct
# definitions in crates/libhelio/src/{camera,instance,draw}.rs.
SIZES = [
    (L.CAMERA, 368), (L.INSTANCE, 144), (L.AABB, 32), (L.DRAW_CALL, 20), (L.DRAW_INDEXED_INDIRECT, 20),
    (L.MESHLET, 64), (L.VG_OBJECT, 128), (L.VG_DRAW, 16), (L.VG_WORK_ITEM, 8), (L.INSTANCE_CULL, 16),
    (L.ID_CULL_UNIFORMS, 112), (L.OCCL_CULL_PARAMS, 56), (L.VG_CULL_UNIFORMS, 48), (L.HIZ_UNIFORMS, 16),
    (L.MESH_LOD, 12),
]


@pytest.mark.parametrize("dt,size", SIZES, ids=[str(s[1]) for s in SIZES])
def test_sizes(dt, size):
    assert dt.itemsize == size


def test_offsets():
    assert L.CAMERA.fields["view_proj"][1] == 128
    assert L.CAMERA.fields["position_near"][1] == 256
    assert L.CAMERA.fields["forward_far"][1] == 272
    assert L.CAMERA.fields["prev_view_proj"][1] == 304
    assert L.INSTANCE.fields["bounds"][1] == 112
    assert L.INSTANCE.fields["mesh_id"][1] == 128
    assert L.INSTANCE.fields["flags"][1] == 136
    assert L.MESHLET.fields["cone_cutoff"][1] == 28
    assert L.MESHLET.fields["lod_error"][1] == 44
    assert L.MESHLET.fields["first_index"][1] == 48
    assert L.MESHLET.fields["instance_index"][1] == 60
    assert L.VG_OBJECT.fields["local_bounds"][1] == 16
    assert L.VG_OBJECT.fields["lod_errors"][1] == 32
    assert L.VG_OBJECT.fields["lod_first_meshlets"][1] == 64
    assert L.VG_OBJECT.fields["lod_meshlet_counts"][1] == 96
    assert L.DRAW_INDEXED_INDIRECT.fields["base_vertex"][1] == 12


def test_publication_budget_bytes():
    """VirtualGeometryBudget::publication_bytes: 36 B per slot + 11 counters
    (crates/helio-pass-virtual-geometry/src/lib.rs:96-98, test :301-307)."""
    slot = L.DRAW_INDEXED_INDIRECT.itemsize + L.VG_DRAW.itemsize
    assert slot == 36
    assert 262_144 * slot + 11 * 4 == 9 * 1024 * 1024 + 44


def test_ctypes_mirrors():
    assert C.sizeof(_lib.CullConfig) == 4 + 28 + 12
    assert C.sizeof(_lib.ViewRequest) == 8 + 8 + 4 + 4 + 8
