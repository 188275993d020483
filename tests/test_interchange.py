"""The on-disk problem interchange (qclhip/interchange.py): round trip of a
synthetic problem, the byte layout the Julia writer emits (a `.npy` header
with fortran_order true and raw column-major data, written here by hand),
and validation errors."""
import json
import os
import struct

import numpy as np
import pytest

import scba_fixture as fx
from qclhip import interchange as ic


def _fixture_problem():
    prob = fx.make_problem(N=4, NE=24, Nk=3)
    arrays = {k: prob[k] for k in ("eps", "wE", "wk", "h", "Tplus", "Tminus")}
    scalars = dict(gs=2, target=1.0, Ep=prob["Ep"], hwLO=prob["hwLO"], NLO=prob["NLO"], kbt=prob["kbt"],
                   alpha_mix=prob["alpha_mix"])
    return prob, arrays, scalars


def test_round_trip_is_exact(tmp_path):
    prob, arrays, scalars = _fixture_problem()
    ic.save_problem(tmp_path, arrays, scalars, prob["mechanisms"])
    got = ic.load_problem(tmp_path)
    assert (got["NE"], got["Nk"], got["Nb"]) == (24, 3, 4)
    for k, v in arrays.items():
        assert np.array_equal(got["arrays"][k], v), k
    for m, g in zip(prob["mechanisms"], got["mechanisms"]):
        assert (g["name"], g["kind"], g["rank"], g["is_lo"]) == (m["name"], m["kind"], m["rank"], m["is_lo"])
        for k in ("a", "b", "c"):
            if m[k] is None:
                assert g[k] is None
            else:
                assert np.array_equal(g[k], m[k])
    assert got["scalars"]["NLO"] == scalars["NLO"]


def _julia_style_npy(path, A):
    """The bytes QCLHipExport.write_npy produces: header with fortran_order
    True padded to 64 bytes, then the column-major data."""
    descr = "<c16" if np.iscomplexobj(A) else "<f8"
    shape = f"({A.shape[0]},)" if A.ndim == 1 else "(" + ", ".join(str(d) for d in A.shape) + ")"
    header = f"{{'descr': '{descr}', 'fortran_order': True, 'shape': {shape}, }}"
    total = 10 + len(header) + 1
    header += " " * ((64 - total % 64) % 64) + "\n"
    with open(path, "wb") as f:
        f.write(b"\x93NUMPY\x01\x00" + struct.pack("<H", len(header)) + header.encode("latin1"))
        f.write(np.asarray(A).ravel(order="F").tobytes())


def test_reads_the_julia_writer_layout(tmp_path):
    rng = np.random.default_rng(1)
    h = rng.normal(size=(3, 4, 4)) + 1j * rng.normal(size=(3, 4, 4))
    eps = np.linspace(0, 1, 5)
    arrays = dict(eps=eps, wE=np.full(5, 0.25), wk=np.ones(3) / 3, h=h, Tplus=h[0], Tminus=h[1].conj().T)
    man = {"format": "qclhip-problem", "version": 1,
           "scalars": dict(gs=2.0, target=1.0, Ep=0.1, hwLO=0.036, NLO=0.3, kbt=0.025, alpha_mix=0.3),
           "arrays": {}, "mechanisms": []}
    for k, v in arrays.items():
        _julia_style_npy(os.path.join(tmp_path, k + ".npy"), v)
        man["arrays"][k] = k + ".npy"
    with open(os.path.join(tmp_path, "manifest.json"), "w") as f:
        json.dump(man, f)
    got = ic.load_problem(tmp_path)
    for k, v in arrays.items():
        assert np.array_equal(got["arrays"][k], v), k
    assert got["arrays"]["h"][2, 1, 3] == h[2, 1, 3]


def test_validation_errors(tmp_path):
    prob, arrays, scalars = _fixture_problem()
    with pytest.raises(ic.InterchangeError):
        ic.save_problem(tmp_path, {k: v for k, v in arrays.items() if k != "h"}, scalars, [])
    ic.save_problem(tmp_path, arrays, scalars, prob["mechanisms"])
    os.remove(os.path.join(tmp_path, "Tplus.npy"))
    with pytest.raises(ic.InterchangeError):
        ic.load_problem(tmp_path)
    ic.save_problem(tmp_path, dict(arrays, Tplus=np.zeros((3, 3))), scalars, [])
    with pytest.raises(ic.InterchangeError):
        ic.load_problem(tmp_path)
    with open(os.path.join(tmp_path, "manifest.json")) as f:
        man = json.load(f)
    man["version"] = 99
    with open(os.path.join(tmp_path, "manifest.json"), "w") as f:
        json.dump(man, f)
    with pytest.raises(ic.InterchangeError):
        ic.load_problem(tmp_path)
