"""A script that leaves its context, scenes, passes, a batch and a server open
exits cleanly: the wrappers close them at interpreter exit, newest first."""
from __future__ import annotations

import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np
    from helio_amd import synth
    from helio_amd.passes import BatchCull, Context, CullServer, InstanceCullPass, Scene
    ctx = Context(0)
    scenes = []
    for i, n in enumerate((20_000, 50_000)):
        sc = synth.instance_scene(n, config=4, index=i)
        s = Scene(ctx, sc.n)
        s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
        s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
        scenes.append((s, n))
    g = synth.rng(4, 7)
    cams = [synth.perspective_view(g).camera for _ in range(8)]
    p = InstanceCullPass(ctx)
    p.prepare(scenes[0][0], cams[0], 0, scenes[0][1])
    p.execute()
    b = BatchCull(ctx, 4)
    b.prepare([(s, cams[k], 0, n) for k, (s, n) in enumerate(scenes)])
    b.execute(graph=True)
    srv = CullServer(ctx, max_batch_views=8, max_inflight=2, max_batch_candidates=1 << 20, copy_ids=False)
    reqs = [CullServer.request(scenes[k % 2][0], cams[k], 0, scenes[k % 2][1], want_ids=False) for k in range(8)]
    lat, wall = srv.run_trace([r[0] for r in reqs], np.arange(8) * 1e-4)
    print("ran", len(lat), flush=True)
""")


def test_exit_with_objects_open():
    print("child script: context, 2 scenes, pass, batch, server left open", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=str(ROOT))], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "ran 8" in r.stdout
