"""The drop-in binding (library phases under the reference's main.cc, the
host's own update_thermal in between; flexrt_imaging_reion_amd.dropin), on
the CPU with the checker in both roles:

  * run in that order, with the host holding only what it pulls after
    find_avg_unu, the result equals the fused step bit for bit, so the pulled
    list covers everything the host's temperature update reads
    (gas_funcs.cc:104-242,353-419);
  * the check has teeth: leaving avg_nh1 (read by update_heat_cool,
    gas_funcs.cc:157) out of the pulled list changes the temperatures;
  * the C++ shims of the integration notes compile against
    include/flexrt_gpu.h, and their kPulled / kPushed lists are the Python
    ones.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from flexrt_imaging_reion_amd import dropin, inputs
from flexrt_imaging_reion_amd._abi import FIELDS
from oracle.pyoracle import Oracle
from parity_util import STATE_FIELDS, bits_equal, compare_rays

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(pulled, steps=5):
    wl = inputs.make_workload("lognormal64", n=16, n_src=16, steps=steps)
    cfg = dict(wl["cfg"])
    fused, lib, host = Oracle(cfg), Oracle(cfg), Oracle(cfg)
    for rt in (fused, lib, host):
        inputs.setup(rt, wl)
    dt = host.derived()["dt"]
    saved = dropin.PULLED
    dropin.PULLED = pulled
    try:
        for s in range(steps):
            fused.step()
            dropin.step_dropin(lib, host, s, dt)
    finally:
        dropin.PULLED = saved
    return fused, lib, host


def test_dropin_order_matches_fused_step():
    fused, lib, host = _run(dropin.PULLED)
    try:
        assert lib.get_time() == fused.get_time()
        assert compare_rays(lib.rays(), fused.rays()) is None
        bad = [k for k in STATE_FIELDS if not bits_equal(lib.download_field(k), fused.download_field(k))]
        assert not bad, bad
        for k in ("temp", "heat_rate", "cool_rate", "IF_speed", "Inc_Flux"):
            assert bits_equal(host.download_field(k), fused.download_field(k)), k
    finally:
        for o in (fused, lib, host):
            o.close()


def test_dropin_missing_pull_is_detected():
    fused, lib, host = _run([k for k in dropin.PULLED if k != "avg_nh1"])
    try:
        assert not bits_equal(host.download_field("temp"), fused.download_field("temp"))
    finally:
        for o in (fused, lib, host):
            o.close()


# The reference globals the shims touch, declared with the reference's types
# (global_variables.cc, user_inputs.h, global_constants.h), so the snippets
# can be compiled on their own.
_GLOBALS = r"""
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include "flexrt_gpu.h"
#define Nx 16
#define Ny 16
#define Nz 16
#define Nfreq 5
#define Nsrc_max 4096
#define TRUE 1
#define FALSE 0
static const int itercount = 5, path_length = 5;
static const float Lx0 = 40000.f, cl_factor = 0.3f, const_clump = 1.f, temp_0 = 100.f, tstep_factor = 1.f;
static const double cl = 2.998e10;
static const float zinit = 8.f, fH1_0 = 1.f - 1.2e-3f;
static const int Nt = 5000;
#define yr_to_s 3.154e7
extern float tre_dom[3], deltaNL_dom[3];
extern float a, zz, dt, t, avg_rho, Lx, Ly, Lz, gadget_step_time;
extern double nu_phot[Nfreq];
extern float spectrum[Nfreq];
extern double source_lum[Nx * Ny * Nz];
extern long int num_thresh;
extern double nphot_min, nphot_max;
extern int step;
extern float clump[Nx * Ny * Nz];
extern float dx[Nx], dy[Ny], dz[Nz];
extern float rho[Nx][Ny][Nz], f_He3[Nx][Ny][Nz], mfp[Nx][Ny][Nz], mfp_912[Nx][Ny][Nz], fion[Nx][Ny][Nz],
    tcross[Nx][Ny][Nz], treion_out[Nx][Ny][Nz];
extern bool tcross_flag[Nx][Ny][Nz], neutral_flag[Nx][Ny][Nz];
extern const char *otf_output, *vIF_file, *small_gas_file;
extern int input_grid;
extern const char* ray_file_init;
extern long steps_done_before_restart;
float scale_factor(float t);
float cosmic_time(float a);
"""


def _cpp_blocks():
    with open(os.path.join(REPO, "INTEGRATION.md")) as f:
        text = f.read()
    sec = text.split("## 1. Drop-in shims")[1].split("## 2.")[0]
    return re.findall(r"```cpp\n(.*?)```", sec, re.S)


def test_shim_lists_match_python():
    head = _cpp_blocks()[0]

    def ids(name):
        m = re.search(r"static const int %s\[\] = \{(.*?)\};" % name, head, re.S)
        return [s.strip() for s in m.group(1).split(",") if s.strip()]

    # frt_field's enumerators in order, paired with the Python field names
    with open(os.path.join(REPO, "include", "flexrt_gpu.h")) as f:
        hdr = f.read()
    enum = re.search(r"typedef enum \{\s*(FRT_F_NH = 0.*?)FRT_F_COUNT", hdr, re.S).group(1)
    names = re.findall(r"(FRT_F_\w+)", enum)
    assert len(names) == len(FIELDS)
    cname = dict(zip(FIELDS, names))
    assert ids("kPulled") == [cname[k] for k in dropin.PULLED]
    assert ids("kPushed") == [cname[k] for k in dropin.PUSHED]
    assert ids("kPulledAfterDensityOps") == [cname[k] for k in dropin.PULLED_AFTER_DENSITY_OPS]


def test_shims_compile(tmp_path):
    blocks = _cpp_blocks()
    first = blocks[0]
    helpers = first[:first.index("void init_rt(void)")]  # g_frt, frt_or_die, the lists, pull()
    for i, b in enumerate(blocks):
        src = _GLOBALS + (b if i == 0 else helpers + b)
        p = tmp_path / ("shim%d.cc" % i)
        p.write_text(src)
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-unused-function",
                            "-I", os.path.join(REPO, "include"), str(p)], capture_output=True, text=True)
        assert r.returncode == 0, "block %d:\n%s" % (i, r.stderr[-3000:])
