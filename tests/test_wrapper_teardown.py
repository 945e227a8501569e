"""Teardown of the Python wrappers: objects still alive at interpreter exit
are closed newest first (contexts last) before module teardown."""
from __future__ import annotations

from helio_amd import passes


class _Fake:
    def __init__(self, log, name):
        self.log, self.name = log, name
        passes._track(self)

    def close(self):
        self.log.append(self.name)


def test_live_objects_close_newest_first():
    saved = list(passes._live)
    passes._live.clear()
    try:
        log = []
        objs = [_Fake(log, n) for n in ("ctx", "scene", "batch", "server")]
        dropped = _Fake(log, "dropped")
        del dropped  # already gone: nothing to close
        passes._close_live()
        assert log == ["server", "batch", "scene", "ctx"]
        assert passes._live == []
        del objs
    finally:
        passes._live[:] = saved
