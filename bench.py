"""Benchmark of the per-timestep RT core on one MI355X.

Workload (the metric's configuration): 128^3 cells, 40 cMpc/h box at z=7,
lognormal density (sigma_ln 1, P(k)~k^-2, seed 2), 4096 sources drawn with
probability ~(1+delta)^2 and Ndot log-uniform in [1e51, 1e56] /s, 5 frequency
bins, clumping C=3, 20 steps of 0.5 Myr.  One bench "step" is one RT timestep
(march, 5 C2Ray passes, merge, split, thermal update, casting).

The chosen (mixed) form is timed over --steps steps from the initial state
after --warmup untimed steps, without instrumentation (`value`); the same
steps are then run again from the initial state with HIP events around every
launch, which give the per-kernel times of `roofline` (that pass's wall time
is reported as ms_per_step_with_launch_events).  The FP64 form is timed the
same way beside it
and its final state is the reference for the mixed-form errors (the FP64 GPU
path is bit-identical to the CPU checker, see tests/test_gpu_parity.py).
The CPU checker runs a bounded sample of the same workload on one host core.
With --restart-files DIR, the mixed run's restart files (ray, gas, treion in
the reference's binary formats) are also written and read back, and the CPU
checker's ray-file writer is timed on a sample of the same rays.  With
--driver-ops, the driver's per-step device ops (cosmic expansion, hydro-step
reload) are timed on the final state.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-RT-updates/s at 128^3, 4096 src: mixed vs FP64 GPU at stated x_HI err"
KERNELS = ["march", "scan_segments", "seg_keys", "sort_segments", "pack_avg",
           "deposit_mixed", "deposit_fp64", "chem", "count_alive", "pdf", "merge_flag",
           "scan_merge", "sort_merge", "merge_groups", "split_flag",
           "scan_split", "out_keys", "sort_rays", "compact", "cast", "zero_ray_dt", "add_ray_dt",
           "thermal"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batch128")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--restart-files", default=None, metavar="DIR",
                    help="after the mixed run, time writing and reading its restart files "
                         "(ray, gas, treion) in DIR (removed afterwards), and the CPU checker's "
                         "ray-file writer on a sample of the same rays")
    ap.add_argument("--driver-ops", action="store_true",
                    help="after the mixed run, time the driver's per-step device ops "
                         "(update_cosmo_expansion, update_hydro_step) on its state, and the CPU "
                         "checker's on the same grid")
    ap.add_argument("--adjoint", action="store_true",
                    help="time one find_avg_unu forward with its adjoint tape and the adjoint "
                         "(frt_adj_find_avg_unu) on the differentiation workload (64^3, 16 sources)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="run with the library's default launches instead of the tuning table's "
                         "choice for this workload (flexrt_imaging_reion_amd/tuning_table.json)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the runs, copy MIB MiB through the library's plain copy kernel "
                         "(reference dispatch for HBM counter calibration)")
    return ap.parse_args()


def kernel_bytes(name, wl, st, mixed):
    """Algorithmic HBM bytes of one launch of `name` in a step with stats `st`.

    deposit: per ray 4 (seg count); per moved ray 4F (photons in);
      per segment 4 (cell) + 4 (length) + 4 (sorted slot) + 8 (avg_nh1, avg_nhe1)
      + 2 acc (H and He deposits) in the first passes, (F+1) acc (photon
      record) in the last pass, which adds 4F + 5 per moved ray (photons out,
      cell, alive).
    chem: per cell 8 (run bounds) + 36 (nH, nHe, clump, f_H1, f_He1, recomb x2,
      avg x2 or their iteration-0 sources) + 8 (avg out); per segment 2 acc in
      the first passes, (F+1) acc in the last, which adds 12F per cell (per-bin
      fields) and 100 per cell (finalize + thermal: ~25 fields).
    march: per ray 56 read (cell, direction, position, time, r, dr) + 5 (flags)
      + 28 written back; per segment 8 (cell, length).
    The per-pass figures are averaged over the itercount passes of the step.
    """
    F = int(wl["cfg"]["nfreq"])
    I = int(wl["cfg"]["itercount"])
    acc = 4.0 if mixed else 8.0
    n3 = float(wl["cfg"]["n"]) ** 3
    nr, nm, ns = float(st["rays_start"]), float(st["rays_moved"]), float(st["segments"])
    if name.startswith("deposit"):
        first = nr * 4 + nm * 4 * F + ns * (20 + 2 * acc)
        last = nr * 4 + nm * 4 * F + ns * (20 + (F + 1) * acc) + nm * (4 * F + 5)
        return ((I - 1) * first + last) / I
    if name == "chem":
        first = n3 * 52 + ns * 2 * acc
        last = n3 * 52 + ns * (F + 1) * acc + n3 * (12 * F + 100)
        return ((I - 1) * first + last) / I
    if name == "march":
        return nr * 89 + ns * 8
    return None


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def run_form(wl, precision, steps, warmup, profile, files_dir=None, driver_ops=False):
    """Warm up, then time `steps` steps of the workload from its initial state
    without instrumentation (the throughput), then, if `profile`, run the same
    steps again from the initial state with every launch bracketed by HIP
    events on the library's stream (per-kernel times for the roofline)."""
    import torch

    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd.rt import FlexRT

    cfg = dict(wl["cfg"])
    cfg["precision"] = precision
    free0, total = torch.cuda.mem_get_info()
    g = FlexRT(cfg)
    inputs.setup(g, wl)
    run_form.launch = None
    if getattr(run_form, "use_tuning", True):
        from flexrt_imaging_reion_amd import autotune

        key, settings = autotune.apply(g, wl["cfg"], precision, len(wl["src_lum"]))
        run_form.launch = {"table_key": key, "settings": settings} if key else None
        log("precision %d: launch settings %s" % (precision, settings if key else "defaults (no table entry)"))
    free1, _ = torch.cuda.mem_get_info()
    run_form.device_memory_gb = round((free0 - free1) / 1e9, 2)
    log("precision %d: context ready (%.1f GB of device memory), %d warmup steps"
        % (precision, (free0 - free1) / 1e9, warmup))
    for _ in range(warmup):
        g.step()

    def timed(prof):
        inputs.setup(g, wl)  # back to the initial state: time steps 1..K of the workload
        g.synchronize()
        g.set_profiling(prof)
        g.reset_kernel_times()
        st = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.append(g.step())
        g.synchronize()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, st

    wall, stats = timed(False)
    log("precision %d: %d steps in %.3f s, %d rays at the end" % (precision, steps, wall, stats[-1]["rays_end"]))
    state = {f: g.download_field(f) for f in ("nH1", "nH", "temp", "gamma_H1_tot")}
    ktimes, wall_prof = {}, None
    if profile:
        wall_prof, _ = timed(True)
        log("precision %d: profiled pass %.3f s" % (precision, wall_prof))
        ktimes = {k: g.kernel_time(k) for k in KERNELS}
        g.set_profiling(False)
    files = restart_files(g, wl, files_dir) if files_dir else None
    if driver_ops:
        files = dict(files or {}, driver_ops=time_driver_ops(g, wl))
    g.close()
    return wall, stats, ktimes, state, wall_prof, files


def time_driver_ops(g, wl, reps=5, sample_rays=2_000_000):
    """update_cosmo_expansion and update_hydro_step on the context's state
    (HIP-event kernel times, mean of `reps` calls; algorithmic bytes: per cell
    4 fields read and 12 fields + the packed opacity pair written = 72 B, per
    ray one 16-byte position read and written = 32 B), and the CPU checker's
    update_cosmo_expansion on the same cells and a sample of the same rays."""
    from oracle.pyoracle import Oracle

    n3 = int(wl["cfg"]["n"]) ** 3
    nrays = g.ray_count()
    a0 = float(wl["cfg"]["a_rt"])
    g.synchronize()
    g.set_profiling(True)
    g.reset_kernel_times()
    for r in range(reps):
        g.update_cosmo_expansion(np.float32(a0 * 1.001 ** (r + 1)), np.float32(a0 * 1.001 ** r))
    g.synchronize()
    kc, nc = g.kernel_time("cosmo_cells")
    kr, nr = g.kernel_time("cosmo_rays")
    g.reset_kernel_times()
    avg_rho0 = np.float32(g.download_field("rho").astype(np.float64).mean())
    g.update_hydro_step(np.float32(wl["cfg"]["zz"]), np.float32(a0), avg_rho0, np.float32(3.15e13))
    g.synchronize()
    kh, _ = g.kernel_time("cosmo_cells")
    g.reset_kernel_times()
    avg_rho = np.float32(g.download_field("rho").astype(np.float64).mean())
    for _ in range(reps):
        g.otf_averages(avg_rho)
    g.synchronize()
    ko, no = g.kernel_time("otf_cells")
    kor, _ = g.kernel_time("otf_rays")
    kof, _ = g.kernel_time("otf_final")
    g.set_profiling(False)
    t0 = time.perf_counter()
    otf_gpu = g.otf_averages(avg_rho)
    otf_wall = time.perf_counter() - t0
    cell_ms, ray_ms = kc / max(nc, 1), kr / max(nr, 1)
    out = {"cells": n3, "rays": nrays,
           "cosmo_cells_ms": round(cell_ms, 4), "cosmo_cells_GBps": round(72 * n3 / (cell_ms * 1e-3) / 1e9, 1),
           "cosmo_rays_ms": round(ray_ms, 4), "cosmo_rays_GBps": round(32 * nrays / (ray_ms * 1e-3) / 1e9, 1),
           "hydro_step_cells_ms": round(kh, 4),
           # write_otf: 21 cell fields of 4 B read; per ray its photon vector(s)
           "otf_cells_ms": round(ko / max(no, 1), 4),
           "otf_cells_GBps": round(84 * n3 / (ko / max(no, 1) * 1e-3) / 1e9, 1),
           "otf_rays_ms": round(kor / max(no, 1), 4),
           "otf_final_ms": round(kof / max(no, 1), 4),
           "otf_call_wall_ms": round(1e3 * otf_wall, 3)}
    rays = g.rays()
    m = min(len(rays["cell"]), sample_rays)
    nf = int(wl["cfg"]["nfreq"])
    part = {k: np.ascontiguousarray(v[:m] if k != "nphot" else v[: m * nf]) for k, v in rays.items()}
    o = Oracle(wl["cfg"])
    try:
        from flexrt_imaging_reion_amd._abi import FIELDS

        for k in FIELDS:
            o.upload_field(k, g.download_field(k))
        o.set_sources(wl["src_cells"], wl["src_lum"])
        t0 = time.perf_counter()
        otf_cpu = o.otf_averages(avg_rho)
        to = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.update_cosmo_expansion(np.float32(a0 * 1.001), np.float32(a0))
        tc = time.perf_counter() - t0
        o.upload_rays(part)
        t0 = time.perf_counter()
        o.update_cosmo_expansion(np.float32(a0 * 1.001), np.float32(a0))
        tr = time.perf_counter() - t0 - tc
    finally:
        o.close()
    out["cpu_baseline"] = {"kind": "port", "cores": 1, "cells_ms": round(1e3 * tc, 2),
                           "rays_sample": m, "rays_ms_per_million": round(1e3 * tr / m * 1e6, 2),
                           "otf_cells_ms": round(1e3 * to, 2)}
    out["otf_max_rel_vs_checker"] = max(
        abs(float(x) - float(y)) / abs(float(y)) for k in ("fH1", "fH2", "nH", "GH1", "T", "MFP", "Heat", "Cool")
        for x, y in zip(np.atleast_1d(otf_gpu[k])[:2], np.atleast_1d(otf_cpu[k])[:2]) if y)
    log("driver ops: cells %.3f ms, rays %.3f ms (GPU); cells %.1f ms (CPU)" % (cell_ms, ray_ms, 1e3 * tc))
    return out



def time_adjoint(precision=0, steps=5, reps=5):
    """Forward (update_rays + find_avg_unu) without and with the tape, and the
    adjoint, on the adjoint64 workload after `steps` steps: wall ms, best of
    `reps` (the state is restored from a device checkpoint before each)."""
    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd.rt import FlexRT

    wl = inputs.make_workload("adjoint64", steps=steps)
    g = FlexRT(dict(wl["cfg"], precision=precision))
    inputs.setup(g, wl)
    for _ in range(steps):
        g.step()
    ck = g.checkpoint()
    n3 = int(wl["cfg"]["n"]) ** 3
    a = np.random.default_rng(0).standard_normal(n3)

    def fwd():
        g.restore(ck)
        g.update_rays()
        g.find_avg_unu()
        g.synchronize()

    out = {}
    for tape in (False, True):
        g.adj_enable(tape)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fwd()
            ts.append(time.perf_counter() - t0)
        out["forward_with_tape_ms" if tape else "forward_ms"] = round(1e3 * min(ts), 3)
    ts = []
    for _ in range(reps):
        fwd()
        t0 = time.perf_counter()
        g.adj_find_avg_unu(a)
        ts.append(time.perf_counter() - t0)
    out["adjoint_ms"] = round(1e3 * min(ts), 3)
    out["workload"] = "adjoint64: 64^3, 16 sources, C=2, state after %d steps, %d rays" % (steps, g.ray_count())
    ck.free()
    g.close()
    log("adjoint: %s" % out)
    return out


def restart_files(g, wl, d, sample=2_000_000):
    """Write and read back the context's restart files in the reference's
    formats (C ABI frt_write_ray_file & co.: pack on the GPU, one copy to the
    host, one file write), best of 3, then the CPU checker's restatement of
    write_rays_binary (one stream write per field) on the first `sample` rays
    of the same list, whose bytes must equal the library file's prefix."""
    from oracle.pyoracle import Oracle

    os.makedirs(d, exist_ok=True)
    paths = {k: os.path.join(d, k) for k in ("ray", "gas", "treion", "ray_cpu")}
    out = {}

    def best(fn):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    try:
        for name, write, read in [("ray", g.write_ray_file, g.read_ray_file),
                                  ("gas", g.write_gas_file, g.read_gas_file),
                                  ("treion", g.write_treion_file, g.read_treion_file)]:
            tw = best(lambda: write(paths[name]))
            size = os.path.getsize(paths[name])
            tr = best(lambda: read(paths[name]))
            out[name] = {"bytes": size, "write_s": round(tw, 4), "read_s": round(tr, 4),
                         "write_GBps": round(size / tw / 1e9, 3), "read_GBps": round(size / tr / 1e9, 3)}
            log("restart %s file: %d bytes, write %.3f s, read %.3f s" % (name, size, tw, tr))
        rays = g.rays()
        n = len(rays["cell"])
        m = min(n, sample)
        part = {k: np.ascontiguousarray(v[:m] if k != "nphot" else v[: m * int(wl["cfg"]["nfreq"])])
                for k, v in rays.items()}
        o = Oracle(wl["cfg"])
        try:
            o.upload_rays(part)
            t0 = time.perf_counter()
            o.write_ray_file(paths["ray_cpu"])
            tc = time.perf_counter() - t0
        finally:
            o.close()
        rec = 57 + 4 * int(wl["cfg"]["nfreq"])
        with open(paths["ray"], "rb") as f:
            head = f.read(m * rec)
        with open(paths["ray_cpu"], "rb") as f:
            cpu = f.read()
        out["ray"]["rays"] = n
        out["cpu_baseline"] = {"kind": "port", "rays": m, "write_s": round(tc, 4),
                                  "rays_per_s": round(m / tc, 1), "GBps": round(len(cpu) / tc / 1e9, 3),
                                  "same_bytes_as_library": cpu[:-8] == head and cpu[-8:] == (-1).to_bytes(
                                      8, "little", signed=True),
                                  "cores": 1}
        log("cpu checker ray file: %d rays in %.3f s" % (m, tc))
    finally:
        for p in paths.values():
            if os.path.exists(p):
                os.remove(p)
    return out


def pmc_traffic(kernel, workload, mixed):
    """HBM bytes per launch of `kernel` measured by rocprofv3 PMC for this
    workload and form (profiles/rt_step_<workload>_<form>_traffic.txt: the
    FETCH_SIZE and WRITE_SIZE counters scaled by a plain-copy calibration)."""
    path = os.path.join(REPO, "profiles", "rt_step_%s_%s_traffic.txt" % (workload, "mixed" if mixed else "fp64"))
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(kernel + ":"):
            kv = dict(t.split("=") for t in line.split(":", 1)[1].split())
            return float(kv["fetch_bytes"]) + float(kv["write_bytes"]), os.path.relpath(path, REPO)
    return None, None


# Algorithmic bytes of the step's terms (per unit), as the scope table of the
# hot path fixes them: a moved ray's record read and written (72 B each way),
# a stored segment written once by the march (8 B) and read with its cell's two
# opacities in every C2Ray pass (b_seg), per cell and pass the Gamma and photon
# outputs, chemistry inputs and opacity outputs (b_ci), and per cell the
# finalize and thermal update (b_cf).
B_RAY = 144.0
B_SEG_MARCH = 8.0
B_SEG = {True: 16.0, False: 24.0}
B_CI = {True: 104.0, False: 208.0}
B_CF = {True: 108.0, False: 216.0}


def step_bytes(wl, st, mixed):
    """B_step = N_ray 144 + N_seg (8 + I b_seg) + N_cell (I b_ci + b_cf), with
    N_ray the rays moved and N_seg the segments stored in that step."""
    I = float(wl["cfg"]["itercount"])
    n3 = float(wl["cfg"]["n"]) ** 3
    return (float(st["rays_moved"]) * B_RAY + float(st["segments"]) * (B_SEG_MARCH + I * B_SEG[mixed]) +
            n3 * (I * B_CI[mixed] + B_CF[mixed]))


def spec_bytes(name, wl, st, mixed):
    """Bytes of one launch of `name` by the same per-unit figures: a deposit
    pass moves N_seg b_seg, a cell pass N_cell b_ci (the last adds b_cf), the
    march N_ray 144 + N_seg 8; None for the other kernels."""
    I = float(wl["cfg"]["itercount"])
    n3 = float(wl["cfg"]["n"]) ** 3
    if name.startswith("deposit"):
        return float(st["segments"]) * B_SEG[mixed]
    if name == "chem":
        return n3 * B_CI[mixed] + n3 * B_CF[mixed] / I
    if name == "march":
        return float(st["rays_moved"]) * B_RAY + float(st["segments"]) * B_SEG_MARCH
    return None


def step_roofline(wl, stats, wall, mixed):
    """Whole-step figure: the steps' summed B_step over their wall time."""
    tot = sum(step_bytes(wl, st, mixed) for st in stats)
    gbs = tot / wall / 1e9
    return {"bytes_per_step": round(tot / len(stats)), "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}


def roofline(wl, stats, ktimes, mixed):
    """Roofline figure of the kernel with the largest share of GPU time:
    achieved = its algorithmic bytes per launch (the per-unit figures of
    step_bytes times the units one launch handles) over its average launch
    time.  The finer per-kernel byte model (kernel_bytes: slot indices, path
    reads, per-ray photon loads, partial writes) is reported beside it."""
    best = None
    for name, (ms, launches) in ktimes.items():
        if launches == 0:
            continue
        if best is None or ms > best[1]:
            best = (name, ms, launches)
    name, ms, launches = best
    total_bytes = 0.0
    n_launch = 0
    for st in stats:
        b = kernel_bytes(name, wl, st, mixed)
        if b is None:
            break
        per_step = wl["cfg"]["itercount"] if (name.startswith("deposit") or name == "chem") else 1
        total_bytes += b * per_step
        n_launch += per_step
    if n_launch == 0 or n_launch != launches:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": name, "kernel_ms_total": ms, "launches": launches}
    avg_ms = ms / launches
    model_bytes = total_bytes / n_launch
    spec = [spec_bytes(name, wl, st, mixed) for st in stats]
    avg_bytes = sum(spec) / len(spec) if spec and spec[0] is not None else model_bytes
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    traffic, source = pmc_traffic(name, wl["name"], mixed)
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5),
           "traffic": round(traffic) if traffic is not None else None, "kernel": name,
           "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": round(avg_bytes),
           "launches": launches, "model_bytes_per_launch": round(model_bytes),
           "model_frac": round(model_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
    if source:
        out["traffic_source"] = source
        out["traffic_GBps"] = round(traffic / (avg_ms * 1e-3) / 1e9, 2)
    return out


CPU_SAMPLE_STEPS = (5, 10, 15, 20)


def fp64_snapshots(wl, steps, sample, make=None):
    """Run the FP64 form (bit-identical to the CPU checker; `make` may supply
    another backend) and keep its full state before each sampled step: every
    field, the ray list, the time and the merge threshold carried into the
    step; also each step's stats."""
    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd._abi import FIELDS
    from flexrt_imaging_reion_amd.rt import FlexRT

    snaps, stats = {}, {}
    g = make() if make else FlexRT(dict(wl["cfg"], precision=0))
    try:
        inputs.setup(g, wl)
        merge = 0.0
        for s in range(1, steps + 1):
            if s in sample:
                snaps[s] = {"fields": {k: g.download_field(k) for k in FIELDS}, "rays": g.rays(),
                            "t": g.get_time(), "nphot_merge": merge}
            stats[s] = g.step()
            merge = stats[s]["nphot_merge"]
    finally:
        g.close()
    return snaps, stats


def resume(wl, snap, s, omp):
    """A CPU checker context holding the state before step s."""
    from flexrt_imaging_reion_amd import inputs
    from oracle.pyoracle import Oracle

    o = Oracle(wl["cfg"], omp=omp)
    inputs.setup(o, wl)
    for k, v in snap["fields"].items():
        o.upload_field(k, v)
    o.upload_rays(snap["rays"])
    o.set_time(snap["t"])
    o.set_step(s - 1)
    o.get_nphot_min_max()
    o.set_nphot_merge(snap["nphot_merge"])
    return o


def final_state(wl, precision, steps, dither=0):
    """Fields and per-step stats after `steps` untimed steps (opacity dither
    seed `dither`, 0 = off)."""
    from flexrt_imaging_reion_amd import inputs
    from flexrt_imaging_reion_amd.rt import FlexRT

    with FlexRT(dict(wl["cfg"], precision=precision)) as g:
        inputs.setup(g, wl)
        g.set_opacity_dither(dither)
        st = [g.step() for _ in range(steps)]
        return {f: g.download_field(f) for f in ("nH1", "nH", "temp", "gamma_H1_tot")}, st


COUNT_KEYS = ("rays_moved", "segments", "rays_dumped", "rays_terminated", "rays_merged_away", "rays_split",
              "rays_end")


def errors_against(state_a, st_a, state_f, st_f):
    """Errors of a run (state_a, st_a) against the FP64 run (state_f, st_f)."""
    xa = state_a["nH1"] / state_a["nH"]
    xf = state_f["nH1"] / state_f["nH"]
    rel = lambda u, v: float(np.max(np.abs(u.astype(np.float64) - v) / np.maximum(np.abs(v), 1e-300)))
    gmax = float(np.abs(state_f["gamma_H1_tot"]).max()) or 1.0
    rx = np.abs(xa.astype(np.float64) - xf) / np.abs(xf.astype(np.float64))
    first_diff = next((k + 1 for k, (sa, sf) in enumerate(zip(st_a, st_f))
                       if any(sa[q] != sf[q] for q in COUNT_KEYS)), None)
    return {
        "x_HI_max_rel": rel(xa, xf),
        "x_HI_rel_p99_9": float(np.percentile(rx, 99.9)),
        "x_HI_rel_median": float(np.median(rx)),
        "cells_x_HI_rel_over_1e-3": int((rx > 1e-3).sum()),
        "first_step_with_count_difference": first_diff,
        "ionized_volume_fraction_abs": abs(float((xa < 0.5).mean()) - float((xf < 0.5).mean())),
        "Gamma_HI_max_abs_over_max": float(np.abs(state_a["gamma_H1_tot"].astype(np.float64) -
                                                  state_f["gamma_H1_tot"]).max()) / gmax,
        "T_max_rel": rel(state_a["temp"], state_f["temp"]),
        "ray_counts_exact": first_diff is None,
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, steps, sample=CPU_SAMPLE_STEPS):
    """The CPU checker's OpenMP build (the reference's parallel loops on all
    host threads, atomic per-cell sums) timed on single steps of the same
    workload, each started from the FP64 GPU state before it (identical to
    the checker's own state there); the 20-step wall time is the piecewise-
    linear interpolation of the sampled step times through (0, 0), summed."""
    from oracle.pyoracle import set_threads

    sample = tuple(s for s in sample if s <= steps)
    snaps, gpu_stats = fp64_snapshots(wl, steps, sample)
    threads = set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    times, rays = {}, {}
    for s in sample:
        o = resume(wl, snaps.pop(s), s, omp=True)
        try:
            t0 = time.perf_counter()
            st = o.step()
            times[s] = time.perf_counter() - t0
            rays[s] = (int(st["rays_end"]), int(gpu_stats[s]["rays_end"]))
        finally:
            o.close()
        log("cpu checker (OpenMP, %d threads) step %d: %.2f s, %d rays (GPU FP64: %d)"
            % (threads, s, times[s], rays[s][0], rays[s][1]))
    xs = [0] + list(sample)
    ys = [0.0] + [times[s] for s in sample]
    est = float(sum(np.interp(k, xs, ys) for k in range(1, steps + 1)))
    n3 = float(wl["cfg"]["n"]) ** 3
    return {"value": round(n3 * steps / est, 1), "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "steps %s of the %d-step workload, each from the FP64 GPU state before it, CPU checker "
                      "OpenMP build (oracle/frt_oracle.cc, g++ -O3 -fopenmp, %d threads); %d-step wall "
                      "interpolated from the sampled steps: %.1f s" % (list(sample), steps, threads, steps, est),
            "step_seconds": {str(s): round(times[s], 3) for s in sample},
            "rays_end_cpu_vs_gpu": {str(s): rays[s] for s in sample},
            "host": {"nproc": os.cpu_count(), "model": cpu_model()}}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the RT core is measured at one GPU's share "
              "(--gpus 1); multi-GPU sharding is a design note, not a measured path.")
        sys.exit(2)
    from flexrt_imaging_reion_amd import inputs

    run_form.use_tuning = not a.no_tuning
    wl = inputs.make_workload(a.workload, steps=max(a.steps, 1))
    log("workload %s built: %d source cells" % (a.workload, len(wl["src_cells"])))
    steps = a.steps
    n3 = float(wl["cfg"]["n"]) ** 3
    wall_m, st_m, kt_m, state_m, wp_m, files = run_form(wl, 1, steps, a.warmup, True, a.restart_files,
                                                        a.driver_ops)
    out = {
        "metric": METRIC,
        "value": round(n3 * steps / wall_m, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * wall_m / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 march/deposit + f64 chemistry",
        "data": "synthetic",
        "config": {"workload": "%s: %d^3 cells, %.0f cMpc/h, z=%.1f, %d source cells, 5 bins, "
                               "%d steps x %.1f Myr, C=%s, mixed precision"
                               % (a.workload, wl["cfg"]["n"], wl["params"]["box_cmpc_h"], wl["params"]["z"],
                                  len(wl["src_cells"]), steps, wl["params"]["dt_myr"], wl["params"]["clump"]),
                   "parallelism": "single GPU"},
        "roofline": roofline(wl, st_m, kt_m, True),
        "step_roofline": step_roofline(wl, st_m, wall_m, True),
        "ms_per_step_with_launch_events": round(1e3 * wp_m / steps, 3),
        # HIP-event time of each launch name (scans and sorts under their
        # step-phase names), ms per step, from the pass with launch events
        "kernel_ms_per_step": {k: round(ms / steps, 4) for k, (ms, n) in sorted(kt_m.items(), key=lambda kv: -kv[1][0])
                               if n} if kt_m else None,
        "rays_end": int(st_m[-1]["rays_end"]),
        "segments_per_step": int(np.mean([s["segments"] for s in st_m])),
        "ray_segments_per_s": round(float(np.sum([s["segments"] for s in st_m])) / wall_m, 1),
        "device_memory_GB": getattr(run_form, "device_memory_gb", None),
        "launch": run_form.launch,
    }
    if not a.no_fp64:
        wall_f, st_f, kt_f, state_f, wp_f, _ = run_form(wl, 0, steps, a.warmup, True)
        out["fp64"] = {"value": round(n3 * steps / wall_f, 1), "ms_per_step": round(1e3 * wall_f / steps, 3),
                       "launch": run_form.launch,
                       "roofline": roofline(wl, st_f, kt_f, False),
                       "step_roofline": step_roofline(wl, st_f, wall_f, False)}
        out["mixed_over_fp64"] = round(wall_f / wall_m, 3)
        out["errors_vs_fp64"] = errors_against(state_m, st_m, state_f, st_f)
        # the same measures for the FP64 form with every stored opacity moved by
        # one float ulp after each pass: the spread rounding at the reference's
        # float storage causes by itself
        state_d, st_d = final_state(wl, 0, steps, dither=1)
        out["fp64_rounding_floor"] = dict(errors_against(state_d, st_d, state_f, st_f),
                                          perturbation="one float ulp on every stored opacity after each pass")
        log("fp64 rounding floor: %s" % out["fp64_rounding_floor"])
    if a.adjoint:
        out["adjoint"] = time_adjoint()
    if files and "driver_ops" in files:
        out["driver_ops"] = files.pop("driver_ops")
    if files:
        out["restart_files"] = files
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, steps)
    if a.calibrate_copy:
        from flexrt_imaging_reion_amd.rt import FlexRT

        with FlexRT(inputs.reference_defaults(8)) as c:
            c.copy_kernel(a.calibrate_copy << 20, 3)
        out["copy_calibration_bytes"] = a.calibrate_copy << 20
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
