"""Launch tuner of the RT step: times frt_step under the launch choices the
library offers (frt_set_tuning keys and the captured step graph) on a
workload, keeps the fastest, and records the choice in a table that
``apply`` reads back.

Search: coordinate descent from the library defaults -- each parameter in
turn takes every value of its range with the others held, the fastest is
kept, and the sweep repeats until a full sweep changes nothing (at most
`sweeps` times).  Every candidate runs the same steps from the workload's
initial state, and its final n_HI must be bit-identical to the defaults'
(the tuning changes only launch shapes and work order, never results); a
candidate that differs is an error, not a choice.

Table (JSON, committed beside this module): key "<arch>|n=<n>|nfreq=<F>|
iter=<I>|prec=<p>|src=<nsrc>" -> {"settings": {...}, "ms_per_step": t,
"default_ms_per_step": t0, "steps": K, "workload": name, "trials": [...]}.
``apply`` looks up the exact key, else the entry of the same arch, grid,
bins, passes and precision with the nearest source count.

    python -m flexrt_imaging_reion_amd.autotune --workload batch128 --precision 1 --out table.json
"""
import argparse
import json
import os
import time

import numpy as np

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")

DEFAULTS = {"graph": 0, "deposit_queue": 0, "march_queue": 0, "queue_chunk": 256, "blocks_per_cu": 8,
            "grid_cap": 1 << 20, "fp64_chunk": 4096}
SPACE = {"graph": [0, 1], "deposit_queue": [0, 1], "march_queue": [0, 1], "queue_chunk": [256, 512, 1024],
         "blocks_per_cu": [4, 8, 16], "grid_cap": [16384, 65536, 1 << 20], "fp64_chunk": [2048, 4096]}


def device_arch():
    import ctypes

    from ._lib import load

    buf = ctypes.create_string_buffer(64)
    if load().frt_device_arch(buf, 64) != 0:
        raise RuntimeError("frt_device_arch failed")
    return buf.value.decode()


def table_key(arch, cfg, precision, nsrc):
    return "%s|n=%d|nfreq=%d|iter=%d|prec=%d|src=%d" % (arch, int(cfg["n"]), int(cfg["nfreq"]),
                                                        int(cfg["itercount"]), int(precision), int(nsrc))


def load_table(path=TABLE_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(table, arch, cfg, precision, nsrc):
    """The entry for this workload shape: exact key, else the nearest source
    count among entries of the same arch, grid, bins, passes and precision."""
    key = table_key(arch, cfg, precision, nsrc)
    if key in table:
        return key, table[key]
    prefix = key.rsplit("|src=", 1)[0] + "|src="
    best = None
    for k, v in table.items():
        if k.startswith(prefix):
            d = abs(np.log(max(int(k[len(prefix):]), 1)) - np.log(max(nsrc, 1)))
            if best is None or d < best[0]:
                best = (d, k, v)
    return (best[1], best[2]) if best else (None, None)


def set_launch(rt, settings):
    s = dict(DEFAULTS, **settings)
    rt.set_tuning(**{k: v for k, v in s.items() if k != "graph"})
    rt.set_step_graph(bool(s["graph"]))


def apply(rt, cfg, precision, nsrc, table=None, arch=None):
    """Applies the table's choice for this workload shape to `rt`; returns
    (key, settings) or (None, None) when the table has no entry."""
    table = load_table() if table is None else table
    arch = arch or device_arch()
    key, entry = lookup(table, arch, cfg, precision, nsrc)
    if entry is None:
        return None, None
    set_launch(rt, entry["settings"])
    return key, entry["settings"]


def time_setting(rt, wl, settings, steps, warm=0):
    """ms per step of `steps` steps after `warm` untimed steps from the
    initial state (the whole sequence runs twice; the second is timed), and
    the final n_HI.  warm = 0 times the workload's own run from its initial
    state, graph captures included; warm > 0 times a later stretch of a long
    run, where a captured graph is replayed without recapture."""
    from . import inputs

    set_launch(rt, settings)
    for _ in range(2):
        inputs.setup(rt, wl)
        for _ in range(warm):
            rt.step()
        rt.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            rt.step()
        rt.synchronize()
        dt = time.perf_counter() - t0
    return 1e3 * dt / steps, rt.download_field("nH1")


def tune(workload, precision, steps=None, sweeps=2, warm=0, log=print):
    from . import inputs
    from .rt import FlexRT

    wl = inputs.make_workload(workload)
    steps = int(steps or wl["steps"])
    space = dict(SPACE)
    if precision != 0:
        space.pop("fp64_chunk")
    trials = []
    with FlexRT(dict(wl["cfg"], precision=precision)) as rt:
        best = dict(DEFAULTS)
        t_best, ref = time_setting(rt, wl, best, steps, warm)
        t_default = t_best
        trials.append({"settings": dict(best), "ms_per_step": round(t_best, 4)})
        log("defaults: %.3f ms/step" % t_best)
        for _ in range(sweeps):
            changed = False
            for name, values in space.items():
                for v in values:
                    if v == best[name]:
                        continue
                    cand = dict(best, **{name: v})
                    if not cand["deposit_queue"] and not cand["march_queue"] and name in ("queue_chunk",
                                                                                        "blocks_per_cu"):
                        continue  # no queue kernel runs: nothing to time
                    t, x = time_setting(rt, wl, cand, steps, warm)
                    if not np.array_equal(x.view(np.uint32), ref.view(np.uint32)):
                        raise RuntimeError("launch setting %r changed the results" % cand)
                    trials.append({"settings": cand, "ms_per_step": round(t, 4)})
                    log("%s=%s: %.3f ms/step" % (name, v, t))
                    if t < t_best * 0.995:  # keep a change only for a clear gain
                        best, t_best, changed = cand, t, True
            if not changed:
                break
    key = table_key(device_arch(), wl["cfg"], precision, len(wl["src_lum"]))
    return key, {"settings": best, "ms_per_step": round(t_best, 4), "default_ms_per_step": round(t_default, 4),
                 "steps": steps, "warm_steps": warm, "workload": workload, "trials": trials}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batch128")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--warm", type=int, default=0,
                    help="untimed steps before the timed ones (0: time the run from the initial state)")
    ap.add_argument("--out", default=TABLE_PATH, help="table to update (JSON)")
    a = ap.parse_args()
    key, entry = tune(a.workload, a.precision, a.steps, a.sweeps, a.warm,
                      log=lambda m: print("autotune: " + m, flush=True))
    table = load_table(a.out)
    table[key] = entry
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({key: {k: v for k, v in entry.items() if k != "trials"}}))


if __name__ == "__main__":
    main()
