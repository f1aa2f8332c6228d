"""Error metrics of a run against a reference run of the same scenario.

* relative L1 = sum|x - x_ref| / sum|x_ref| and relative L-infinity =
  max|x - x_ref| / max|x_ref|, of the interior concentration (all classes)
  and of the ground load cdry(:,:,0) (the class sum, endstep.f90:80-106);
* mass residual rho_M = |B - B_ref| / erumass with B = massvol + massdep +
  outmass, the budget terms cmass3d.f90 and cmass2d.f90 produce, and the
  relative difference of each term (None for a term whose reference value is
  below 1e-12 of the injected mass: early in a run no mass has reached the
  boundary, and the reference's outmass is then an FP64 tail far below what
  an FP32 state holds).

The budget's closure against the injected mass, (B - erumass)/erumass, is
reported beside them; it is not a parity figure (the scheme does not conserve
it, in the reference either).
"""
from __future__ import annotations

import numpy as np

__all__ = ["rel_l1", "rel_linf", "mass", "compare"]


def rel_l1(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    den = float(np.sum(np.abs(b)))
    return 0.0 if den == 0 else float(np.sum(np.abs(a - b)) / den)


def rel_linf(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    den = float(np.max(np.abs(b))) if b.size else 0.0
    return 0.0 if den == 0 else float(np.max(np.abs(a - b)) / den)


def mass(b) -> float:
    return b["massvol"] + b["massdep"] + b["outmass"]


def _rel(x, y, scale):
    if abs(y) <= 1e-12 * abs(scale):
        return None if x != y else 0.0
    return abs(x - y) / abs(y)


def compare(load, load_ref, budget, budget_ref, c=None, c_ref=None) -> dict:
    """Metrics of one result against its reference.  load: (nx,ny) ground
    load (class sum); c: interior concentration, any shape (optional)."""
    out = {
        "load_rel_l1": rel_l1(load, load_ref),
        "load_rel_linf": rel_linf(load, load_ref),
        "rho_M": abs(mass(budget) - mass(budget_ref)) / budget_ref["erumass"],
        "massvol_rel": _rel(budget["massvol"], budget_ref["massvol"], budget_ref["erumass"]),
        "massdep_rel": _rel(budget["massdep"], budget_ref["massdep"], budget_ref["erumass"]),
        "outmass_rel": _rel(budget["outmass"], budget_ref["outmass"], budget_ref["erumass"]),
        "closure": (mass(budget) - budget["erumass"]) / budget["erumass"],
    }
    if c is not None:
        out["c_rel_l1"] = rel_l1(c, c_ref)
        out["c_rel_linf"] = rel_linf(c, c_ref)
    return out
