"""Tuning-table logic of the launch tuner (no GPU): keys, exact and
nearest-source lookup, and what apply() sets on a context."""
import json

from flexrt_imaging_reion_amd import autotune


class FakeRT:
    def __init__(self):
        self.tuning, self.graph = {}, None

    def set_tuning(self, **kw):
        self.tuning.update(kw)

    def set_step_graph(self, on=True):
        self.graph = bool(on)


CFG = {"n": 128, "nfreq": 5, "itercount": 5}


def _table():
    e = lambda g, q: {"settings": {"graph": g, "deposit_queue": q}, "ms_per_step": 1.0}
    return {autotune.table_key("gfx950", CFG, 1, 3935): e(1, 0),
            autotune.table_key("gfx950", CFG, 1, 100): e(0, 1),
            autotune.table_key("gfx950", CFG, 0, 3935): e(0, 0)}


def test_key_names_the_shape():
    assert autotune.table_key("gfx950", CFG, 1, 3935) == "gfx950|n=128|nfreq=5|iter=5|prec=1|src=3935"


def test_lookup_exact_then_nearest_source_count():
    t = _table()
    k, e = autotune.lookup(t, "gfx950", CFG, 1, 3935)
    assert k.endswith("src=3935") and e["settings"]["graph"] == 1
    k, e = autotune.lookup(t, "gfx950", CFG, 1, 2000)  # log-nearest: 3935
    assert k.endswith("src=3935")
    k, e = autotune.lookup(t, "gfx950", CFG, 1, 150)
    assert k.endswith("src=100")
    assert autotune.lookup(t, "gfx942", CFG, 1, 3935) == (None, None)
    assert autotune.lookup(t, "gfx950", dict(CFG, n=64), 1, 3935) == (None, None)


def test_apply_sets_every_key_with_defaults_filled_in():
    rt = FakeRT()
    key, s = autotune.apply(rt, CFG, 1, 3935, table=_table(), arch="gfx950")
    assert key is not None and rt.graph is True
    assert set(rt.tuning) == set(autotune.DEFAULTS) - {"graph"}
    assert rt.tuning["deposit_queue"] == 0 and rt.tuning["fp64_chunk"] == autotune.DEFAULTS["fp64_chunk"]
    rt2 = FakeRT()
    assert autotune.apply(rt2, dict(CFG, nfreq=3), 1, 3935, table=_table(), arch="gfx950") == (None, None)
    assert rt2.graph is None and not rt2.tuning


def test_committed_table_is_well_formed():
    t = autotune.load_table()
    for k, v in t.items():
        assert k.count("|") == 5, k
        assert set(v["settings"]) <= set(autotune.DEFAULTS), k
        for name, val in v["settings"].items():
            assert val in autotune.SPACE[name] or val == autotune.DEFAULTS[name], (k, name, val)
        assert v["ms_per_step"] <= v["default_ms_per_step"] * 1.001, k
        json.dumps(v)
