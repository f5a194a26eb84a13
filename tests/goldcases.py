"""Golden fixture catalogue shared by the tests: which fixtures exist and how
to rebuild the case directory each one was made from."""
import ast
import os

import numpy as np

from csu_amd import cases, native, renumber

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RAMP = ["ramp_roe", "ramp_ausm_o1", "ramp_ausmdv_lim1", "ramp_ausmpup",
        "ramp_ausmpup2_sub", "ramp_quad_global_dt", "ramp_tp_dtswitch",
        "ramp_lusgs_ausmdv", "ramp_dplur_ausm", "ramp_lusgs_colour"]
PLATE = ["plate_lam", "plate_lamT_ausmdv", "plate_lam_axi_roe", "plate_euler_axi",
         "plate_lam_pup_lim1", "plate_lusgs_lam", "plate_lusgs_lamT", "plate_dplur_lam"]
SA = ["plate_sa_ausmdv", "plate_sa_blazek_wallT", "plate_sa_pup_lim1", "plate_sa_lusgs", "plate_sa_dplur",
      "plate_sacc_ausmdv", "plate_sacc_lusgs_wallT"]
NACA = ["naca_sa_lusgs", "naca_sst_lusgs"]
SST = ["plate_sst_ausm", "plate_sst_ausmdv_wallT", "plate_sst_pup_axi", "plate_sst_pup2_t1", "plate_sst_lusgs",
       "plate_sa_lusgs_colour", "plate_sst_trans", "plate_sst_trans_lusgs_wallT"]
PLATE = PLATE + SA + SST + NACA
ALL = RAMP + PLATE
# tube mode (time-accurate runs to tmax)
TUBE = ["tube_roe_o1", "tube_ausmdv_o2", "tube_ausmdv_o2_lim1", "tube_ausm_quad", "tube_pup_moving",
        "tube_roe_tp"]


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def write(name, g, d, nmax=1):
    mk = ast.literal_eval(str(g["mesh_kwargs"]))
    inp = ast.literal_eval(str(g["inputs"]))
    ren = mk.pop("renumber", False)
    if name.startswith("plate"):
        cases.write_case(d, cases.plate_mesh(**mk), cases.plate_input(threads=1, nmax=nmax, **inp))
    elif name.startswith("naca"):
        cases.write_case(d, cases.naca_mesh(**mk), cases.naca_input(threads=1, nmax=nmax, **inp))
    elif name.startswith("tube"):
        cases.write_case(d, cases.tube_mesh(**mk), cases.tube_input(threads=1, **inp))
    else:
        cases.write_case(d, cases.ramp_mesh(seed=0, **mk), cases.ramp_input(threads=1, nmax=nmax, **inp))
    if ren:
        renumber.renumber_file(d + "mesh.su2", d + "mesh.su2")
    return d


def host_case(name, g, d, nmax=1):
    return native.HostCase(write(name, g, d, nmax))
