"""A caller may record a whole occlusion frame into its own graph: every
execute only enqueues work, so the Hi-Z build of a depth buffer and the
per-instance occlusion cull that reads it are captured on the caller's stream
(here through the HIP runtime the library itself uses: stream capture,
instantiate, launch) and replayed. Before each replay the cull's outputs are
overwritten (and every other replay follows an eager frame of the same
passes); each replay's outputs equal the eager frame's, which equal the
oracle's."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd._lib import load
from helio_amd.passes import HiZBuildPass, InstanceCullPass, Scene, Stream
from tests.conftest import args_as_u32, canonical_ids, check_cull_invariants
from tests.test_gpu_inst_cull import _half_wall_depth

pytestmark = pytest.mark.gpu


def _same(got, ref):
    ids, args, stats = got
    rids, rargs, rstats = ref
    check_cull_invariants(ids, args, stats)
    assert np.array_equal(stats, rstats), (stats, rstats)
    assert np.array_equal(args_as_u32(args), rargs)
    assert np.array_equal(canonical_ids(ids, args), rids)


def test_occlusion_frame_recorded_in_a_caller_graph(ctx):
    sc = synth.instance_scene(300_000, config=1, index=21)
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    view = synth.perspective_view(synth.rng(1, 77), reversed_z=True)
    depth = _half_wall_depth(view, True, dist=60.0)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
    hiz.upload_depth(depth)
    p = InstanceCullPass(ctx, occlusion=True, reversed_z=True)
    p.prepare(scene, view.camera, 0, sc.n, hiz)

    flags = L.CULL_OCCLUSION | L.CULL_REVERSED_Z
    ov = oracle.view_from_camera(view.camera, flags=flags, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
    rids, rargs, rstats = oracle.cull_view(ov, sc.bounds, sc.meta(), 0, sc.n, oracle.pyramid(depth, 1), None,
                                           threads=8)
    assert rstats[4] > 1000, rstats  # the wall occludes
    ref = (rids, rargs, rstats)

    # The eager frame (it also sizes everything the passes use).
    hiz.execute()
    p.execute()
    _same(p.read(), ref)

    hip = C.CDLL("libamdhip64.so.7")  # the runtime libhv.so is linked against
    vp, P = C.c_void_p, C.POINTER
    for fn, args in (("hipStreamBeginCapture", [vp, C.c_int]), ("hipStreamEndCapture", [vp, P(vp)]),
                     ("hipGraphInstantiate", [P(vp), vp, vp, vp, C.c_size_t]), ("hipGraphLaunch", [vp, vp]),
                     ("hipStreamSynchronize", [vp]), ("hipGraphExecDestroy", [vp]), ("hipGraphDestroy", [vp])):
        getattr(hip, fn).argtypes, getattr(hip, fn).restype = args, C.c_int

    def ok(e):
        assert e == 0, e

    st = Stream(ctx)
    graph, exe = vp(), vp()
    print("capturing the pyramid build and the occlusion cull on a library stream", flush=True)
    ok(hip.hipStreamBeginCapture(st.handle, 1))  # thread-local capture
    hiz.execute(stream=st.handle)
    p.execute(stream=st.handle)
    ok(hip.hipStreamEndCapture(st.handle, C.byref(graph)))
    ok(hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0))
    o = p.outputs()
    lib = load()

    def poison():
        # Every output word the cull writes, overwritten outside the graph, so
        # a replay must write them all again.
        for ptr_, words in ((o.ids, o.id_capacity), (o.args, o.bucket_count * 5), (o.stats, 9)):
            assert lib.hv_memset_d32(ctx.handle, ptr_, 0xFFFFFFFF, words, None) == 0
        ctx.synchronize()

    for i in range(5):
        if i % 2:  # an eager frame of the same passes between replays
            hiz.execute()
            p.execute()
        ctx.synchronize()
        poison()
        ok(hip.hipGraphLaunch(exe, st.handle))
        ok(hip.hipStreamSynchronize(st.handle))
        print(f"replay {i}", flush=True)
        _same(p.read(), ref)
    ok(hip.hipGraphExecDestroy(exe))
    ok(hip.hipGraphDestroy(graph))
    st.close()
    p.close()
    hiz.close()
    scene.close()
