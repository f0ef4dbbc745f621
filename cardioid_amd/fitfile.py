"""Reader for Cardioid's rational-fit files (``NAME FIT { ... }`` objects).

The TT06Dev reaction loads its 33 rational fits from the file named by the
REACTION object's ``fitFile`` keyword (elec/TT06Dev_Reaction.cc:606-624, read
by ``scanFit`` at :557-582).  Each object carries ``tol``, ``deltaX``, ``x0``,
``x1``, ``l``, ``m`` and ``coef``; the coefficient layout is numerator
``coef[0..m-1]`` (ascending powers), ``coef[m] = 1``, denominator
``coef[m+1..m+l-1]`` (elec/pade.cc:17-36, 72-74).  The ``functions FIT``
object lists the fit names in the order the reaction uses them.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

# Order of the 33 fits inside the reaction (elec/TT06Dev_Reaction.cc:206-207,
# with jhTauSmooth = 1 substituting the smoothed h/j tau fits, :218-219).
FIT_ORDER = [
    "fv0", "fv1", "fv2", "fv3", "fv4", "fv5", "fv6",
    "mMhu", "mTauR", "hMhu", "hTauRMod", "jMhu", "jTauRMod",
    "Xr1Mhu", "Xr1TauR", "Xr2Mhu", "Xr2TauR", "XsMhu", "XsTauR",
    "rMhu", "rTauR", "dMhu", "dTauR", "fMhu", "fTauR", "f2Mhu", "f2TauR",
    "jLMhu", "jLTauR", "sMhu0", "sTauR0", "sMhu1", "sTauR1",
]
GATE_NAMES = ["m", "h", "j", "Xr1", "Xr2", "Xs", "r", "d", "f", "f2", "jL", "s0", "s1"]

DEFAULT_FIT_FILE = os.path.join(os.path.dirname(__file__), "data", "tt06_fit_tol1e-4_mod1.data")

_OBJ_RE = re.compile(r"(\w+)\s+FIT\s*\{(.*?)\}", re.S)


@dataclass
class Fit:
    name: str
    l: int
    m: int
    coef: np.ndarray
    tol: float = 0.0
    deltaX: float = 0.1
    x0: float = -100.0
    x1: float = 50.0


@dataclass
class FitSet:
    fits: List[Fit] = field(default_factory=list)

    def packed(self):
        """(l[33], m[33], coef packed in order) as int32/int32/float64 arrays."""
        l = np.array([f.l for f in self.fits], dtype=np.int32)
        m = np.array([f.m for f in self.fits], dtype=np.int32)
        coef = np.concatenate([f.coef for f in self.fits]).astype(np.float64)
        return l, m, coef


def _parse_body(body: str) -> Dict[str, str]:
    out: Dict[str, str] = {}
    for item in body.split(";"):
        if "=" not in item:
            continue
        k, v = item.split("=", 1)
        out[k.strip()] = v.strip()
    return out


def read_fit_file(path: str = DEFAULT_FIT_FILE) -> FitSet:
    """Parse a fit file; raises ValueError on a malformed or incomplete set."""
    with open(path, "r") as fh:
        text = fh.read()
    objs = {}
    order = None
    for name, body in _OBJ_RE.findall(text):
        kv = _parse_body(body)
        if name == "functions":
            order = kv["functions"].split()
            continue
        l = int(kv["l"])
        m = int(kv["m"])
        coef = np.array([float(x) for x in kv["coef"].split()], dtype=np.float64)
        if coef.size != l + m:
            raise ValueError(f"fit {name}: {coef.size} coefficients, expected l+m={l + m}")
        if coef[m] != 1.0 and l >= 1:
            raise ValueError(f"fit {name}: coef[m] must be 1")
        objs[name] = Fit(name, l, m, coef, float(kv.get("tol", 0.0)),
                         float(kv.get("deltaX", 0.1)), float(kv.get("x0", -100.0)),
                         float(kv.get("x1", 50.0)))
    if order is None:
        order = FIT_ORDER
    if order != FIT_ORDER:
        raise ValueError(f"fit order {order} differs from the reaction's {FIT_ORDER}")
    missing = [n for n in FIT_ORDER if n not in objs]
    if missing:
        raise ValueError(f"fit file lacks {missing}")
    return FitSet([objs[n] for n in FIT_ORDER])
