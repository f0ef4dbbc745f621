"""Algorithmic floating-point operations per tissue cell-step of the path,
re-derived from the fit orders of the fit file (FMA = 2 flop; a division, a
log or an exp term = 1).

The fits are the only part whose cost depends on the data: a fit with
numerator length m and denominator length l (elec/pade.cc:17-36) costs m - 1
Horner multiply-adds for the numerator and, when l >= 2, l - 1 more for the
denominator.  The rest of the step is fixed arithmetic:

  non-gate update (elec/TT06NonGates.c:374-563): the 7 fits fv0..fv6 plus
      177 flop of currents, fluxes and updates (with its 13 divisions);
  gate update (elec/TT06GatesSimd.c, v2): the 24 rate fits of a cell type
      (s0 fits for endo, s1 for mid and epi) plus 100 flop (the m gate's
      16-term exponential, 12 Rush-Larsen / rate updates, 19 divisions);
  4 logs (reversal potentials);
  stencil (elec/FGRDiffusionOMP.cc:117-128): 18 multiply-adds and a scale = 37;
  integrator (elec/simulationLoop.cc:221-227): 4.

    python -m cardioid_amd.flops          (prints the breakdown)
"""
from __future__ import annotations

from typing import Dict

from .fitfile import DEFAULT_FIT_FILE, FIT_ORDER, read_fit_file

NONGATE_FITS = FIT_ORDER[:7]
GATE_FITS_S0 = FIT_ORDER[7:29] + ["sMhu0", "sTauR0"]
GATE_FITS_S1 = FIT_ORDER[7:29] + ["sMhu1", "sTauR1"]
NONGATE_OTHER = 177
GATE_OTHER = 100
LOGS = 4
STENCIL = 2 * 18 + 1
INTEGRATOR = 4


def horner_fmas(l: int, m: int) -> int:
    return (m - 1) + (l - 1 if l >= 2 else 0)


def breakdown(path: str = DEFAULT_FIT_FILE, s_switch: int = 0) -> Dict[str, int]:
    fits = {f.name: f for f in read_fit_file(path).fits}
    ng = sum(horner_fmas(fits[n].l, fits[n].m) for n in NONGATE_FITS)
    gate_names = GATE_FITS_S1 if s_switch else GATE_FITS_S0
    g = sum(horner_fmas(fits[n].l, fits[n].m) for n in gate_names)
    out = {"nongate_fit_fmas": ng, "gate_fit_fmas": g,
           "nongate": 2 * ng + NONGATE_OTHER, "gates": 2 * g + GATE_OTHER, "logs": LOGS,
           "stencil": STENCIL, "integrator": INTEGRATOR}
    out["reaction"] = out["nongate"] + out["gates"] + LOGS
    out["total"] = out["reaction"] + STENCIL + INTEGRATOR
    return out


def flop_per_cell_step(path: str = DEFAULT_FIT_FILE) -> int:
    """Endocardial cells (s0 fits), the figure the rooflines use."""
    return breakdown(path, 0)["total"]


if __name__ == "__main__":
    for sw, name in ((0, "endo (s0)"), (1, "mid/epi (s1)")):
        print(name, breakdown(s_switch=sw))
