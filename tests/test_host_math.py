"""Host-side math of libhv.so (runs on the CPU, no GPU needed) against the
oracle and the reference's unit-test vectors."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd.passes import build_vg_work_items, extract_frustum_planes, instance_cull_data, mip_levels

lib = oracle.lib()


def test_planes_bit_exact_with_oracle():
    g = np.random.default_rng(11)
    for _ in range(2000):
        vp = (g.normal(size=16) * g.choice([1e-3, 1.0, 1e3])).astype(np.float32)
        got = extract_frustum_planes(vp).reshape(24)
        ref = np.zeros(24, np.float32)
        lib.or_extract_frustum_planes(vp.ctypes.data, ref.ctypes.data)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("w,h", [(1, 1), (1920, 1080), (4096, 4096), (0, 7), (3, 5), (65535, 2)])
def test_mip_levels_matches_oracle(w, h):
    assert mip_levels(w, h) == lib.or_mip_levels(w, h)


def test_instance_cull_data_bit_exact_with_oracle():
    g = np.random.default_rng(12)
    n = 3000
    inst = np.zeros(n, L.INSTANCE)
    m = g.normal(size=(n, 16)).astype(np.float32)
    m[:, [3, 7, 11]] = np.where(g.random((n, 3)) < 0.8, 0.0, m[:, [3, 7, 11]] * 1e-5)
    m[:, 15] = np.where(g.random(n) < 0.9, 1.0, m[:, 15])
    # A quarter are exact rotations x uniform scale (cone enabled).
    k = n // 4
    q = g.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    rot = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y), np.zeros(k),
                    2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x), np.zeros(k),
                    2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y), np.zeros(k)], axis=1)
    m[:k, :12] = (rot * g.uniform(0.1, 10, (k, 1))).astype(np.float32)
    m[5, 0] = np.nan
    m[6, :] = 0
    inst["model"] = m
    got = instance_cull_data(inst)
    for i in range(n):
        mx, mn, cone, valid = C.c_float(), C.c_float(), C.c_uint32(), C.c_uint32()
        lib.or_instance_cull_data(inst[i:i + 1].ctypes.data, C.byref(mx), C.byref(mn), C.byref(cone), C.byref(valid))
        row = got[i]
        assert np.float32(row["max_scale"]).view(np.uint32) == np.float32(mx.value).view(np.uint32), i
        assert np.float32(row["min_scale"]).view(np.uint32) == np.float32(mn.value).view(np.uint32), i
        assert (row["cone_cull_enabled"], row["valid_transform"]) == (cone.value, valid.value), i
    assert got["cone_cull_enabled"][:k].sum() > 0.9 * k
    assert got["valid_transform"][5] == 0 and got["valid_transform"][6] == 0


def test_instance_cull_data_reference_kats():
    """crates/helio-pass-virtual-geometry/src/lib.rs:315-372 through the product."""
    def icd(model):
        inst = np.zeros(1, L.INSTANCE)
        inst["model"] = np.asarray(model, np.float32)
        return instance_cull_data(inst)[0]
    r = icd([2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 4, 5, 6, 1])
    assert (r["valid_transform"], r["cone_cull_enabled"], r["max_scale"], r["min_scale"]) == (1, 1, 2.0, 2.0)
    r = icd([4, 0, 0, 0, 0, 2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])
    assert (r["valid_transform"], r["cone_cull_enabled"], r["max_scale"], r["min_scale"]) == (1, 0, 4.0, 1.0)
    assert icd([0] * 16)["valid_transform"] == 0
    assert icd([1, 0, 0, 0, 0.5, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])["cone_cull_enabled"] == 0
    assert icd([-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])["cone_cull_enabled"] == 0


def test_work_items_reference_kat():
    """work_items_cover_exact_fixed_meshlet_spans
    (crates/helio/src/scene/virtual_geometry/rebuild.rs:234-259)."""
    objs = np.zeros(5, L.VG_OBJECT)
    objs["max_meshlet_count"] = [0, 1, 64, 65, 130]
    wi = build_vg_work_items(objs)
    got = [(int(a), int(b)) for a, b in zip(wi["object_index"], wi["local_meshlet_base"])]
    # Object indices here are the positions 0..4 standing for the reference's 3, 5, 7, 11, 13.
    assert got == [(1, 0), (2, 0), (3, 0), (3, 64), (4, 0), (4, 64), (4, 128)]


def test_vg_counter_layout_decoding():
    """VirtualGeometryDebugStats::from_counters (lib.rs:390-403): slot 0 attempted,
    1 rejected, 2..9 LOD histogram, 10 max LOD."""
    c = np.array([123, 4, 0, 2, 5, 0, 0, 1, 0, 0, 6], np.uint32)
    assert c[0] == 123 and c[1] == 4
    hist = c[2:10]
    assert hist.sum() == 8
    nz = np.nonzero(hist)[0]
    assert (nz[0], nz[-1]) == (1, 5)
    assert min(int(c[10]), 7) == 6
