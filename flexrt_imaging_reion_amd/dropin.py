"""Drop-in binding order: the library's phases under the reference's own
driver, with the unchanged host update_thermal in between.

In this mode (the replacement rt_funcs.cc / ray_tracing.cc shims of the
integration notes) main.cc and gas_funcs.cc stay as they are, so the
temperature update (gas_funcs.cc:362-457) runs on the host's grid globals
between find_avg_unu and new_rays (main.cc:117-198).  The host must then hold
every field that update_thermal and its callees read, and the library every
field the host wrote before it reads it again:

  PULLED  written by the library's update_rays/find_avg_unu and read by the
          host afterwards: update_thermal and set_Treion_default read nH1,
          nH1_prev, n_tot, dne_dt (gas_funcs.cc:105-106,354-355,398-399);
          update_heat_cool reads f_H2, avg_nh1, avg_nhe1, gamma_H1_nu,
          gamma_He1_nu, ne, nH2, nHe2 (:149-241); write_otf, write_gas_binary
          and write_treion read the rates, fractions, mfp, ngam_nu and the
          treion flags (io_funcs.cc:338-393,500-657).
  PUSHED  written by the host's update_thermal (temp, heat_rate, cool_rate,
          IF_speed, Inc_Flux, nH1_prev; gas_funcs.cc:125-128,371-412) or
          set_clump (clump, subgrid.cc:799-811), read by the library's next
          find_avg_unu (temp through the recombination coefficients,
          ray_tracing.cc:661-674; clump, :1096), its file writers and its
          vIF cube.
  PULLED_AFTER_DENSITY_OPS  what the host reads after the library rescaled
          the densities (frt_update_cosmo_expansion, frt_update_hydro_step):
          update_thermal reads rho and ddelta_dt (:399) besides the
          densities.

`step_dropin` runs one driver step in that order, `lib` being the library
(rt.FlexRT) and `host` anything with the reference's update_thermal on its
own copy of the globals and the same upload/download interface.
"""
import numpy as np

PULLED = [
    "nH1", "nHe1", "nH2", "nHe2", "ne", "n_tot", "f_H1", "f_He1", "f_H2", "f_He2", "f_He3",
    "dne_dt", "avg_nh1", "avg_nhe1", "gamma_H1_tot", "gamma_He1_tot", "gamma_H1_nu", "gamma_He1_nu",
    "ngam_nu", "recomb_H2", "recomb_He2", "recrate_H2", "recrate_He2", "treion", "treion_flag",
    "mfp", "mfp_912",
]

PUSHED = ["temp", "heat_rate", "cool_rate", "IF_speed", "Inc_Flux", "nH1_prev", "clump"]

PULLED_AFTER_DENSITY_OPS = [
    "rho", "nH", "nHe", "nH1", "nHe1", "nH2", "nHe2", "ne", "n_tot", "avg_nh1", "avg_nhe1",
    "nH1_prev", "ddelta_dt", "rho_prev",
]


def copy_fields(src, dst, names):
    for k in names:
        dst.upload_field(k, src.download_field(k))


def step_dropin(lib, host, steps_done, dt):
    """One pass of main.cc:117-198 with the library's phases and the host's
    update_thermal: t += dt (float, main.cc:123), update_rays, find_avg_unu,
    (pull), merge, split_rays, update_thermal on the host, (push), new_rays.
    steps_done is the driver's count of completed steps; it keys the
    library's angle streams exactly as frt_step would."""
    lib.set_step(steps_done)
    copy_fields(host, lib, PUSHED)
    t = np.float32(np.float32(host.get_time()) + np.float32(dt))
    host.set_time(t)
    lib.set_time(t)
    lib.update_rays()
    lib.find_avg_unu()
    copy_fields(lib, host, PULLED)
    lib.merge_rays()
    lib.split_rays()
    host.update_thermal()
    copy_fields(host, lib, PUSHED)
    lib.new_rays()
    lib.set_step(steps_done + 1)
