"""On-disk interchange of an assembled problem (the reference's problem
assembly stays in Julia; this carries its output to the device backend).

A problem is a directory: `manifest.json` plus one `.npy` file per array.
Arrays are stored as the reference holds them (column-major, `fortran_order`
true, complex128 or float64), so a Julia writer emits the raw column-major
bytes after a `.npy` header (`julia/export_problem.jl`); the reader uses
`numpy.load(..., allow_pickle=False)` only.

manifest.json:
    {"format": "qclhip-problem", "version": 1,
     "scalars": {"gs", "target", "Ep", "hwLO", "NLO", "kbt", "alpha_mix",
                 optional "lambda_P", "F", "z0", "C", "alpha_P"},
     "arrays": {"eps", "wE", "wk", "h" (Nk,Nb,Nb), "Tplus", "Tminus",
                optional "kappa", "x", "wx", "eps_r", "ND", "Phi", "chi" (Nz,Nb),
                "H0", "Z", "Minv" (Nb,Nb)}  -> file names,
     "mechanisms": [{"name", "kind" (dense | momentum_independent |
                     separable_ifr | low_rank), "qK", "is_lo", "rank",
                     "arrays": {"a", "b", "c"} -> file names}]}

The fields follow the reference's types: ModelGrids (types.jl:515-526) for
eps, wE, kappa, wk, x, wx; MaterialProfiles (:539-547) for eps_r and N_D;
BasisData (:558-578) for Phi, chi, H0, Z, Minv, T+, T-; the kernel
representations of production_contracts.jl:262-323; h = project_hamiltonians
(basis.jl:288-308) at the exported Hartree energy.
"""
from __future__ import annotations

import json
import os

import numpy as np

FORMAT = "qclhip-problem"
VERSION = 1
KINDS = ("dense", "momentum_independent", "separable_ifr", "low_rank")
REQUIRED = ("eps", "wE", "wk", "h", "Tplus", "Tminus")
SCALARS = ("gs", "target", "Ep", "hwLO", "NLO", "kbt", "alpha_mix")


class InterchangeError(ValueError):
    """A problem directory that does not follow the format."""


def _save(path, name, x):
    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = x.astype(np.complex128)
    elif x.dtype.kind in "fiu":
        x = x.astype(np.float64)
    else:
        raise InterchangeError(f"array {name!r} must be numeric")
    fname = f"{name}.npy"
    np.save(os.path.join(path, fname), np.asfortranarray(x), allow_pickle=False)
    return fname


def save_problem(path, arrays: dict, scalars: dict, mechanisms: list):
    """Write a problem directory. arrays: name -> ndarray (REQUIRED plus any
    optional ones); scalars: name -> number; mechanisms: dicts with name,
    kind, qK, is_lo, rank and arrays a, b, c (None when unused)."""
    os.makedirs(path, exist_ok=True)
    for k in REQUIRED:
        if k not in arrays:
            raise InterchangeError(f"missing required array {k!r}")
    for k in SCALARS:
        if k not in scalars:
            raise InterchangeError(f"missing required scalar {k!r}")
    man = {"format": FORMAT, "version": VERSION, "scalars": {k: float(v) for k, v in scalars.items()},
           "arrays": {}, "mechanisms": []}
    for k, v in arrays.items():
        man["arrays"][k] = _save(path, k, v)
    for i, m in enumerate(mechanisms):
        if m["kind"] not in KINDS:
            raise InterchangeError(f"unknown kernel kind {m['kind']!r}")
        ent = {"name": str(m["name"]), "kind": m["kind"], "qK": float(m.get("qK", 1.0)),
               "is_lo": bool(m.get("is_lo", False)), "rank": int(m.get("rank", 0)), "arrays": {}}
        for k in ("a", "b", "c"):
            if m.get(k) is not None:
                ent["arrays"][k] = _save(path, f"mech{i}_{k}", m[k])
        man["mechanisms"].append(ent)
    with open(os.path.join(path, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1)


def _load(path, fname):
    full = os.path.join(path, fname)
    if not os.path.isfile(full):
        raise InterchangeError(f"missing array file {fname!r}")
    x = np.load(full, allow_pickle=False)
    if x.dtype not in (np.complex128, np.float64):
        raise InterchangeError(f"{fname}: dtype {x.dtype} (complex128 or float64 expected)")
    return x


def load_problem(path) -> dict:
    """Read and validate a problem directory: dict with arrays, scalars,
    mechanisms (arrays loaded) and the sizes NE, Nk, Nb (and Nz)."""
    mpath = os.path.join(path, "manifest.json")
    if not os.path.isfile(mpath):
        raise InterchangeError("missing manifest.json")
    with open(mpath) as f:
        man = json.load(f)
    if man.get("format") != FORMAT or man.get("version") != VERSION:
        raise InterchangeError(f"not a {FORMAT} version {VERSION} directory")
    arrays = {k: _load(path, v) for k, v in man["arrays"].items()}
    for k in REQUIRED:
        if k not in arrays:
            raise InterchangeError(f"missing required array {k!r}")
    scalars = dict(man["scalars"])
    for k in SCALARS:
        if k not in scalars:
            raise InterchangeError(f"missing required scalar {k!r}")
    NE, Nk = arrays["eps"].size, arrays["wk"].size
    h = arrays["h"]
    if h.ndim != 3 or h.shape[0] != Nk or h.shape[1] != h.shape[2]:
        raise InterchangeError(f"h has shape {h.shape}, expected (Nk, Nb, Nb) with Nk = {Nk}")
    Nb = h.shape[1]
    if arrays["wE"].size != NE:
        raise InterchangeError("wE and eps differ in length")
    for k in ("Tplus", "Tminus", "H0", "Z", "Minv"):
        if k in arrays and arrays[k].shape != (Nb, Nb):
            raise InterchangeError(f"{k} has shape {arrays[k].shape}, expected ({Nb}, {Nb})")
    Nz = None
    for k in ("Phi", "chi"):
        if k in arrays:
            if arrays[k].ndim != 2 or arrays[k].shape[1] != Nb:
                raise InterchangeError(f"{k} has shape {arrays[k].shape}, expected (Nz, {Nb})")
            Nz = arrays[k].shape[0]
    mechs = []
    for ent in man["mechanisms"]:
        if ent["kind"] not in KINDS:
            raise InterchangeError(f"unknown kernel kind {ent['kind']!r}")
        m = dict(ent)
        m["arrays"] = {k: _load(path, v) for k, v in ent["arrays"].items()}
        for k in ("a", "b", "c"):
            m[k] = m["arrays"].get(k)
        mechs.append(m)
    return dict(arrays=arrays, scalars=scalars, mechanisms=mechs, NE=NE, Nk=Nk, Nb=Nb, Nz=Nz)


def device_problem(prob: dict, measure: bool = True):
    """An ScbaProblem on the device from a loaded problem (and, when the
    Poisson arrays are present, an OuterProblem)."""
    import torch
    from . import api, outer, scba
    a, s = prob["arrays"], prob["scalars"]
    g = api.Grid(a["eps"], a["wE"], a["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], api.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, api.h_to_device(a["h"]), api.matrix_to_device(a["Tplus"]),
                         api.matrix_to_device(a["Tminus"]), mechs, s["target"], s["Ep"], s["hwLO"], s["NLO"],
                         s["alpha_mix"], measure=measure)
    op = None
    if all(k in a for k in ("eps_r", "wx", "ND", "Phi", "chi", "H0", "Z", "Minv", "kappa")) and \
            all(k in s for k in ("lambda_P", "F", "z0", "C")):
        Nb = prob["Nb"]
        S = a["H0"] - s["F"] * (a["Z"] - s["z0"] * np.eye(Nb))
        kin = s["C"] * a["kappa"] ** 2
        dev = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=g.ctx.device)
        wx = a["wx"]
        op = outer.OuterProblem(dev(a["eps_r"]), dev(wx), dev(a["ND"]), s["lambda_P"], api.basis_to_device(a["chi"]),
                                api.basis_to_device(a["Phi"]), api.matrix_to_device(S), dev(kin),
                                api.matrix_to_device(a["Minv"]), gs=s["gs"], alpha_P=s.get("alpha_P", 0.5),
                                ND2D=float(np.dot(wx, a["ND"])), Lp=float(np.sum(wx)))
    return p, op
