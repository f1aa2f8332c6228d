"""Rebuild a Scenario from a golden fixture (tests/golden/*.npz)."""
import glob
import os

import numpy as np

from fall3d_amd import scenario as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_fixtures():
    return sorted(glob.glob(os.path.join(GOLDEN, "run_*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


# the fields of f3d_gravity, in the order the gravity-current fixtures hold them
GRAVITY_KEYS = ("c_flow_rate", "lambda_grav", "k_entrain", "brunt_vaisala", "center_x", "center_y",
                "xorigr", "yorigr", "dx", "dy", "beg_time")


def scenario_from(g):
    nx, ny, nz, nc = (int(g[k]) for k in ("nx", "ny", "nz", "nc"))
    # gravity-current fixtures keep the input vy apart from the wind the
    # meteo stage leaves (radial wind added)
    gravity = (dict(zip(GRAVITY_KEYS, (float(v) for v in g["gravity"])))
               if "gravity" in g else None)
    return S.Scenario(
        name=os.path.basename(str(g.get("name", "golden"))), nx=nx, ny=ny, nz=nz, nc=nc,
        dx=float(g["dx"]), dy=float(g["dy"]), zlayer=g["zlayer"], diam=g["diam"],
        rhop=g["rhop"], sphe=g["sphe"], fc=np.ones(nc) / nc, vx=g["vx_in"],
        vy=g["vy_in"] if "vy_in" in g else g["vy"],
        vz=g["vz"], tempe=g["tempe"], vptemp=g["vptemp"], rho=g["rho_in"], ustar=g["ustar"],
        zi=g["zi"], rl=g["rl"], topg=g["topg"], vdep=g["vdep"], isrc=g["isrc"],
        jsrc=g["jsrc"], ksrc=g["ksrc"], tmrat=g["tmrat"], flowrate=float(g["flowrate"]),
        modkv=int(g["modkv"]), modkh=int(g["modkh"]), modv=int(g["modv"]),
        rkv0=float(g["rkv0"]), rkh0=float(g["rkh0"]), csh=float(g["csh"]),
        prate=g.get("prate"), lonlat=int(g.get("lonlat", 0)),
        dlon=float(g.get("dlon", 0.0)), dlat=float(g.get("dlat", 0.0)),
        hm1_cols=g["hm1"] if int(g.get("lonlat", 0)) else None,
        gravity=gravity, zs=g.get("zs"))
