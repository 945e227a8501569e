"""The CPU oracle of the instance/draw rebuild, pinned before the GPU is compared with it.

The reference has no tests for rebuild_instance_buffers_{persistent,optimized}
(crates/helio/src/scene/objects/rebuild.rs:91-361) or
rebuild_shadow_partition_buffers (:374-425), so the oracle is pinned by:
- two hand-worked cases whose every output is written out below, one per mode,
  including a draw group that spans a (class, graph_hash) change (grouping
  compares only (mesh_id, material_id), rebuild.rs:283-286) and a missing
  material reading (0, 0);
- invariants on seeded random sets: the slots form a permutation that carries
  every instance and aabb byte; keys are non-decreasing and ties keep the
  dense order (Rust's sort_by_key is stable); groups tile the slots with
  uniform (mesh, material) and differ from their neighbours; ranges tile the
  groups with uniform keys; the shadow lists are the stable movability split
  with first_instance = the new slot; visibility is object_is_visible
  (crates/helio/src/scene/helpers.rs:87-89);
- a pure-Python restatement with sorted() on small sets.
Plausible mistakes each fail one of them: an unstable sort, the key fields in
another order, groups split on a class change, ranges counted over objects in
optimized mode, the shadow lists in sorted instead of dense order.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth


def _set(mesh, mat, slot, mov, groups, materials):
    n = len(mesh)
    inst = np.zeros(n, L.INSTANCE)
    inst["mesh_id"] = mesh
    inst["material_id"] = mat
    inst["bounds"][:, 0] = np.arange(n)  # tags the object in the copied bytes
    ab = np.zeros(n, L.AABB)
    ab["min"][:, 0] = np.arange(n) + 100
    dr = np.zeros(n, L.DRAW_CALL)
    dr["index_count"] = 10 + np.arange(n)
    dr["first_index"] = 1000 + np.arange(n)
    dr["vertex_offset"] = -5 - np.arange(n)
    dr["first_instance"] = np.arange(n)
    dr["instance_count"] = 1
    ob = np.zeros(n, L.OBJECT_INFO)
    ob["material_slot"] = slot
    ob["movability"] = mov
    ob["groups"] = groups
    mt = np.zeros(len(materials), L.MATERIAL_KEY)
    for i, (c, h) in enumerate(materials):
        mt[i]["material_class"], mt[i]["graph_hash"] = c, h
    return inst, ab, dr, ob, mt


MATS = [(1, 7), (0, 9), (1, 7), (0, 3)]


def test_kat_persistent():
    # objects 0..5: material slots 0, 1, 5 (missing), 2, 3, 1
    inst, ab, dr, ob, mt = _set([2, 1, 1, 2, 0, 1], [0, 1, 1, 0, 0, 1], [0, 1, 5, 2, 3, 1], [2, 0, 1, 2, 0, 2],
                                [0, 2, 1, 3, 0, 4], MATS)
    r = oracle.scene_rebuild(L.REBUILD_PERSISTENT, inst, ab, dr, ob, mt, group_hidden=2)
    # (class, hash): o0 (1,7) o1 (0,9) o2 (0,0) o3 (1,7) o4 (0,3) o5 (0,9) -> order 2, 4, 1, 5, 0, 3
    order = [2, 4, 1, 5, 0, 3]
    assert list(r["counts"]) == [6, 4, 3, 3]
    assert list(r["gpu_slots"]) == [4, 2, 0, 5, 1, 3]
    assert list(r["instances"]["bounds"][:, 0]) == order
    assert list(r["aabbs"]["min"][:, 0]) == [o + 100 for o in order]
    assert [tuple(x) for x in r["draw_calls"]] == [(10 + o, 1000 + o, 0xFFFFFFFF - 4 - o, s, 1)
                                                  for s, o in enumerate(order)]
    assert [tuple(x) for x in r["indirect"]] == [(10 + o, 1, 1000 + o, 0xFFFFFFFF - 4 - o, s)
                                                for s, o in enumerate(order)]
    assert [(x["material_class"], x["graph_hash"], x["start"], x["count"]) for x in r["ranges"]] == [
        (0, 0, 0, 1), (0, 3, 1, 1), (0, 9, 2, 2), (1, 7, 4, 2)]
    # movable: o0, o3, o5 (dense order); static: o1, o2, o4
    assert [tuple(x) for x in r["shadow_movable"]] == [(10, 1, 1000, 0xFFFFFFFB, 4), (13, 1, 1003, 0xFFFFFFF8, 5),
                                                      (15, 1, 1005, 0xFFFFFFF6, 3)]
    assert [x[4] for x in r["shadow_static"]] == [2, 0, 1]
    # groups 0, 2, 1, 3, 0, 4 with hidden = 2 -> visible 1, 0, 1, 0, 1, 1 (sorted order)
    assert list(r["visibility"]) == [1, 1, 0, 1, 1, 0]


def test_kat_optimized_group_spans_a_class_change():
    inst, ab, dr, ob, mt = _set([2, 1, 1, 2, 0, 1], [0, 1, 1, 0, 0, 1], [0, 1, 5, 2, 0, 1], [2, 0, 1, 2, 0, 2],
                                [0, 2, 1, 3, 0, 4], MATS)
    r = oracle.scene_rebuild(L.REBUILD_OPTIMIZED, inst, ab, dr, ob, mt, group_hidden=2)
    # keys: o2 (0,0,1,1) o1 (0,9,1,1) o5 (0,9,1,1) o4 (1,7,0,0) o0 (1,7,2,0) o3 (1,7,2,0)
    order = [2, 1, 5, 4, 0, 3]
    assert list(r["instances"]["bounds"][:, 0]) == order
    assert list(r["gpu_slots"]) == [4, 1, 0, 5, 3, 2]
    # groups: {o2, o1, o5} (one (mesh, material) run across (0,0) -> (0,9)), {o4}, {o0, o3}
    assert list(r["counts"]) == [3, 2, 3, 3]
    assert [tuple(x) for x in r["draw_calls"]] == [(12, 1002, 0xFFFFFFF9, 0, 3), (14, 1004, 0xFFFFFFF7, 3, 1),
                                                  (10, 1000, 0xFFFFFFFB, 4, 2)]
    assert [tuple(x) for x in r["indirect"]] == [(12, 3, 1002, 0xFFFFFFF9, 0), (14, 1, 1004, 0xFFFFFFF7, 3),
                                                (10, 2, 1000, 0xFFFFFFFB, 4)]
    # ranges over draw groups keyed by each group's first object: (0,0) then (1,7) x 2
    assert [(x["material_class"], x["graph_hash"], x["start"], x["count"]) for x in r["ranges"]] == [
        (0, 0, 0, 1), (1, 7, 1, 2)]
    assert [x[4] for x in r["shadow_movable"]] == [4, 5, 2]
    assert [x[4] for x in r["shadow_static"]] == [1, 0, 3]
    assert list(r["visibility"]) == [1, 0, 1, 1, 1, 0]


def test_zero_objects_leave_the_shadow_counts():
    e = np.zeros(0, L.INSTANCE)
    r = oracle.scene_rebuild(L.REBUILD_OPTIMIZED, e, np.zeros(0, L.AABB), np.zeros(0, L.DRAW_CALL),
                             np.zeros(0, L.OBJECT_INFO), np.zeros(0, L.MATERIAL_KEY), counts_init=(9, 9, 7, 5))
    assert list(r["counts"]) == [0, 0, 7, 5]


def _keys(s: synth.ObjectSet, idx, mode):
    ob, mt = s.objects[idx], s.materials
    slot = ob["material_slot"].astype(np.int64)
    ok = slot < len(mt)
    cls = np.where(ok, mt["material_class"][np.minimum(slot, max(len(mt) - 1, 0))] if len(mt) else 0, 0)
    gh = np.where(ok, mt["graph_hash"][np.minimum(slot, max(len(mt) - 1, 0))] if len(mt) else 0, 0)
    cols = [cls.astype(object), gh.astype(np.uint64).astype(object)]
    if mode == L.REBUILD_OPTIMIZED:
        cols += [s.instances["mesh_id"][idx].astype(object), s.instances["material_id"][idx].astype(object)]
    return list(zip(*cols))


def check_invariants(s: synth.ObjectSet, r: dict, mode: int, group_hidden: int):
    n = len(s.instances)
    slots = r["gpu_slots"].astype(np.int64)
    assert np.array_equal(np.sort(slots), np.arange(n))
    order = np.empty(n, np.int64)
    order[slots] = np.arange(n)
    assert r["instances"].tobytes() == s.instances[order].tobytes()
    assert r["aabbs"].tobytes() == s.aabbs[order].tobytes()
    keys = _keys(s, order, mode)
    for a in range(n - 1):  # non-decreasing, ties in dense order
        assert keys[a] < keys[a + 1] or (keys[a] == keys[a + 1] and order[a] < order[a + 1])
    dc = r["draw_calls"].view(np.uint32).reshape(-1, 5)
    G = int(r["counts"][0])
    assert len(dc) == G
    assert int(dc[:, 4].sum()) == n
    assert np.array_equal(dc[:, 3], np.concatenate([[0], np.cumsum(dc[:-1, 4])]).astype(np.uint32))
    mm = list(zip(s.instances["mesh_id"][order].tolist(), s.instances["material_id"][order].tolist()))
    heads = dc[:, 3].astype(np.int64)
    tmpl = s.draws.view(np.uint32).reshape(-1, 5)
    for g in range(G):
        h, c = heads[g], int(dc[g, 4])
        assert np.array_equal(dc[g, :3], tmpl[order[h], :3])
        if mode == L.REBUILD_PERSISTENT:
            assert c == 1
        else:
            assert len(set(mm[h:h + c])) == 1
            if g:
                assert mm[heads[g - 1]] != mm[h]
    ind = r["indirect"].view(np.uint32).reshape(-1, 5)
    assert np.array_equal(ind[:, [0, 1, 2, 3, 4]], dc[:, [0, 4, 1, 2, 3]])
    rg = r["ranges"]
    assert int(r["counts"][1]) == len(rg)
    assert int(rg["count"].sum()) == G
    assert np.array_equal(rg["start"], np.concatenate([[0], np.cumsum(rg["count"][:-1])]).astype(np.uint32))
    gkeys = [k[:2] for k in _keys(s, order[heads], L.REBUILD_PERSISTENT)]
    for j, x in enumerate(rg):
        want = (int(x["material_class"]), int(x["graph_hash"]))
        assert all(tuple(int(v) for v in gkeys[g]) == want for g in range(x["start"], x["start"] + x["count"]))
        if j:
            assert (int(rg[j - 1]["material_class"]), int(rg[j - 1]["graph_hash"])) != want
    mov = s.objects["movability"] == L.MOVABILITY_MOVABLE
    for lst, sel in ((r["shadow_movable"], mov), (r["shadow_static"], ~mov)):
        e = lst.view(np.uint32).reshape(-1, 5)
        d = np.flatnonzero(sel)
        assert len(e) == len(d)
        assert np.array_equal(e[:, 0], tmpl[d, 0]) and np.array_equal(e[:, 2], tmpl[d, 1])
        assert np.array_equal(e[:, 3], tmpl[d, 2]) and (e[:, 1] == 1).all()
        assert np.array_equal(e[:, 4], slots[d].astype(np.uint32))
    gr = s.objects["groups"][order]
    want_vis = ((gr == 0) | ((gr & np.uint64(group_hidden)) == 0)).astype(np.uint32)
    assert np.array_equal(r["visibility"], want_vis)


@pytest.mark.parametrize("mode", [L.REBUILD_PERSISTENT, L.REBUILD_OPTIMIZED])
@pytest.mark.parametrize("kw", [dict(), dict(material_count=0), dict(material_count=3000, hash_pool=2),
                                dict(mesh_count=1, material_count=1, class_count=1, hash_pool=1, p_missing=0.0,
                                     p_foreign_material_id=0.0),
                                dict(wide_ids=True)])
def test_invariants(mode, kw):
    s = synth.scene_objects(3000, seed=len(kw), **kw)
    r = oracle.scene_rebuild(mode, s.instances, s.aabbs, s.draws, s.objects, s.materials, group_hidden=0b1010)
    check_invariants(s, r, mode, 0b1010)


def _python_rebuild(s: synth.ObjectSet, mode: int):
    n, mt = len(s.instances), s.materials
    def pair(i):
        sl = int(s.objects["material_slot"][i])
        return (int(mt["material_class"][sl]), int(mt["graph_hash"][sl])) if sl < len(mt) else (0, 0)
    def key(i):
        k = pair(i)
        if mode == L.REBUILD_OPTIMIZED:
            k += (int(s.instances["mesh_id"][i]), int(s.instances["material_id"][i]))
        return k
    order = sorted(range(n), key=key)
    groups = []
    for slot, o in enumerate(order):
        mm = (int(s.instances["mesh_id"][o]), int(s.instances["material_id"][o]))
        if mode == L.REBUILD_OPTIMIZED and groups and groups[-1][2] == mm:
            groups[-1][1] += 1
        else:
            groups.append([slot, 1, mm, pair(o)])
    ranges = []
    for gi, g in enumerate(groups):
        if ranges and ranges[-1][0] == g[3]:
            ranges[-1][2] += 1
        else:
            ranges.append([g[3], gi, 1])
    return order, [(g[0], g[1]) for g in groups], [(r[0][0], r[0][1], r[1], r[2]) for r in ranges]


@pytest.mark.parametrize("mode", [L.REBUILD_PERSISTENT, L.REBUILD_OPTIMIZED])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_python_restatement_small(mode, seed):
    s = synth.scene_objects(400, mesh_count=6, material_count=9, class_count=3, hash_pool=3, seed=seed,
                            p_foreign_material_id=0.3)
    r = oracle.scene_rebuild(mode, s.instances, s.aabbs, s.draws, s.objects, s.materials)
    order, groups, ranges = _python_rebuild(s, mode)
    slots = np.empty(len(order), np.int64)
    slots[order] = np.arange(len(order))
    assert np.array_equal(r["gpu_slots"], slots)
    dc = r["draw_calls"].view(np.uint32).reshape(-1, 5)
    assert [(int(a), int(b)) for a, b in dc[:, 3:5]] == groups
    assert [(int(x["material_class"]), int(x["graph_hash"]), int(x["start"]), int(x["count"]))
            for x in r["ranges"]] == ranges


try:
    from hypothesis import HealthCheck, given, settings
    from hypothesis import strategies as st
except ImportError:  # pragma: no cover
    given = None


if given is not None:
    @settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])
    @given(n=st.integers(0, 300), meshes=st.integers(1, 8), materials=st.integers(0, 12),
           classes=st.integers(1, 4), pool=st.integers(1, 4), foreign=st.floats(0.0, 0.6),
           seed=st.integers(0, 10_000), mode=st.sampled_from([L.REBUILD_PERSISTENT, L.REBUILD_OPTIMIZED]))
    def test_oracle_equals_python_restatement_random(n, meshes, materials, classes, pool, foreign, seed, mode):
        s = synth.scene_objects(n, mesh_count=meshes, material_count=materials, class_count=classes,
                                hash_pool=pool, p_foreign_material_id=foreign, seed=seed)
        r = oracle.scene_rebuild(mode, s.instances, s.aabbs, s.draws, s.objects, s.materials, group_hidden=0b101)
        if n == 0:
            assert list(r["counts"][:2]) == [0, 0]
            return
        order, groups, ranges = _python_rebuild(s, mode)
        slots = np.empty(n, np.int64)
        slots[order] = np.arange(n)
        assert np.array_equal(r["gpu_slots"], slots)
        dc = r["draw_calls"].view(np.uint32).reshape(-1, 5)
        assert [(int(a), int(b)) for a, b in dc[:, 3:5]] == groups
        assert [(int(x["material_class"]), int(x["graph_hash"]), int(x["start"]), int(x["count"]))
                for x in r["ranges"]] == ranges
        check_invariants(s, r, mode, 0b101)
