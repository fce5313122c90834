"""roctx ranges: balanced push/pop around the traced host steps, and no-ops
without the roctx library (HVR_ROCTX=0)."""

from types import SimpleNamespace

import numpy as np

from pyvr_amd.hip_renderer import trace


class Recorder:
    def __init__(self):
        self.events = []

    def roctxRangePushA(self, name):
        self.events.append(("push", name.decode()))
        return 0

    def roctxRangePop(self):
        self.events.append(("pop", None))
        return 0

    def roctxMarkA(self, msg):
        self.events.append(("mark", msg.decode()))


def test_spans_nest_and_pop_on_exceptions(monkeypatch):
    rec = Recorder()
    monkeypatch.setattr(trace, "_lib", rec)
    monkeypatch.setattr(trace, "_tried", True)
    with trace.span("a"):
        with trace.span("b"):
            trace.mark("m")
    try:
        with trace.span("c"):
            raise ValueError
    except ValueError:
        pass
    assert rec.events == [("push", "a"), ("push", "b"), ("mark", "m"), ("pop", None), ("pop", None),
                          ("push", "c"), ("pop", None)]


def test_disabled_by_environment(monkeypatch):
    monkeypatch.setattr(trace, "_lib", None)
    monkeypatch.setattr(trace, "_tried", False)
    monkeypatch.setenv("HVR_ROCTX", "0")
    assert not trace.enabled()
    with trace.span("x"):
        trace.mark("y")


def test_serving_loop_is_traced(monkeypatch):
    """One poll of the serving loop over the fake backend records an
    admission range per launch and a collection range per finished batch."""
    import test_serve_logic as t

    rec = Recorder()
    monkeypatch.setattr(trace, "_lib", rec)
    monkeypatch.setattr(trace, "_tried", True)
    srv, be = t.make(3, max_batch_requests=2)
    for k in "ABC":
        srv.submit(t.req(k))
    srv.poll()
    be.finish(0)
    srv.poll()
    names = [n for e, n in rec.events if e == "push"]
    assert names.count("hvr.serve.admit") == 2 and names.count("hvr.serve.collect") == 1
    depth = 0
    for e, _ in rec.events:
        depth += {"push": 1, "pop": -1}.get(e, 0)
        assert depth >= 0
    assert depth == 0
