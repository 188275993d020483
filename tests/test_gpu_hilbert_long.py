"""FFT Hilbert transform on energy axes beyond one workgroup's LDS.

Above L = 16384 (NE > 8192; the reference documents NE = 30721,
docs/src/theory/19_production.md) each column runs as a four-step
transform through HBM scratch. It is checked against the oracle's
transform (fft_parallel.jl:247-291 restated, any power-of-two L), against
Dawson's integral (closed form: sign and scale of the kernel), against the
one-workgroup path on the lengths both take (4096..16384, several column
chunks included), with the 2L round-off witness, the S^R composition and
the adjoint's dot-product identity.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _uniform(NE, lo=-1.3, hi=0.9):
    eps = np.linspace(lo, hi, NE)
    wE = np.full(NE, eps[1] - eps[0])
    wE[[0, -1]] /= 2
    return eps, wE


def _fields(rng, NE, Nk, Nb, n):
    return [np.asfortranarray(rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb)))
            for _ in range(n)]


def _ip(a, b):
    import torch
    return float(torch.sum(torch.conj(a) * b).real)


@pytest.mark.parametrize("NE", [8193, 12000, 30721, 40000, 70000])
def test_long_axis_vs_oracle(hip, orc, NE):
    """Every row length of the four-step split: L2 = 4096 (L = 32768,
    65536), 8192 (L = 131072) and 16384 (L = 262144)."""
    rng = np.random.default_rng(NE)
    Nk, Nb = 3, 1
    eps, wE = _uniform(NE)
    Gam, SL, SG = _fields(rng, NE, Nk, Nb, 3)
    g = hip.Grid(eps, wE, np.ones(Nk), Nb)
    plan = hip.FFTHilbertPlan(g, verify_roundoff=True)
    assert plan.path == "hbm"
    print(f"NE={NE} ncols={Nk * Nb * Nb}", flush=True)
    lam, r = hip.production_fft_hilbert_transform(plan, hip.to_device(Gam), return_roundoff=True)
    ref, rr = orc.fft_hilbert(Gam, eps, wE)
    assert rel(hip.to_host(lam), ref) < TOL
    assert r < 1e-12 and abs(r - rr) < 1e-13
    SR, _, r2 = hip.retarded_self_energy_fft(plan, hip.to_device(SL), hip.to_device(SG), return_roundoff=True)
    SRo, _, ro = orc.retarded_self_energy_fft(SL, SG, eps, wE)
    assert rel(hip.to_host(SR), SRo) < TOL
    assert abs(r2 - ro) < 1e-13
    if NE == 8193:
        # the discrete principal-value sum itself, O(NE^2) per column
        assert rel(hip.to_host(lam), orc.direct_hilbert(Gam, eps, wE)) < 1e-12


@pytest.mark.parametrize("n", [30721, 70001])
def test_long_axis_dawson(hip, n):
    """Gaussian columns on 30721 and 70001 nodes (L = 65536, 262144):
    Dawson's integral plus the trapezoid rule's omitted-node term, the
    closed form that fixes the kernel's sign and 1/(2 pi) scale."""
    from test_oracle_anchors import dawson_columns
    x, w, G, exact = dawson_columns(n)
    ncol = G.shape[2]
    G = np.asfortranarray(G.reshape(n, ncol, 1, 1, order="F"))
    exact = exact.reshape(n, ncol, 1, 1, order="F")
    g = hip.Grid(x, w, np.ones(ncol), 1)
    plan = hip.FFTHilbertPlan(g, ncol, verify_roundoff=True)
    assert plan.path == "hbm"
    L, r = hip.production_fft_hilbert_transform(plan, hip.to_device(G), return_roundoff=True)
    assert np.max(np.abs(hip.to_host(L) - exact)) < 1e-13 * np.max(np.abs(exact))
    assert r < 1e-12


@pytest.mark.parametrize("NE", [2048, 3000, 5000, 8192])
def test_fourstep_equals_one_workgroup_path(hip, NE):
    rng = np.random.default_rng(3 + NE)
    Nk, Nb = 3, 2
    eps, wE = _uniform(NE)
    Gam, SL, SG, Rb = (hip.to_device(x) for x in _fields(rng, NE, Nk, Nb, 4))
    g = hip.Grid(eps, wE, np.ones(Nk), Nb)
    plan = hip.FFTHilbertPlan(g, verify_roundoff=True)
    assert plan.path == "lds"
    out = {}
    for path in ("lds", "hbm"):
        plan.set_path(path)
        assert plan.path == path
        lam, r = hip.production_fft_hilbert_transform(plan, Gam, return_roundoff=True)
        SR, _, r2 = hip.retarded_self_energy_fft(plan, SL, SG, return_roundoff=True)
        import torch
        Lb, Gb = torch.zeros_like(SL), torch.zeros_like(SG)
        hip.retarded_self_energy_adjoint(plan, Rb, Lb, Gb)
        out[path] = [hip.to_host(v) for v in (lam, SR, Lb, Gb)]
        assert r < 1e-13 and r2 < 1e-13
    for a, b in zip(out["lds"], out["hbm"]):
        assert rel(a, b) < 1e-13


def test_fourstep_several_chunks(hip):
    """2200 columns at L = 4096 with the witness take three chunks of the
    four-step scratch (1024 columns per 128 MiB, the last one partial); the
    result equals the one-workgroup path's."""
    rng = np.random.default_rng(11)
    NE, Nk, Nb = 2048, 550, 2
    eps, wE = _uniform(NE)
    SL, SG = (hip.to_device(x) for x in _fields(rng, NE, Nk, Nb, 2))
    g = hip.Grid(eps, wE, np.ones(Nk), Nb)
    plan = hip.FFTHilbertPlan(g, verify_roundoff=True)
    res = {}
    for path in ("lds", "hbm"):
        plan.set_path(path)
        SR, _, r = hip.retarded_self_energy_fft(plan, SL, SG, return_roundoff=True)
        res[path] = (hip.to_host(SR), r)
        del SR
    assert rel(res["lds"][0], res["hbm"][0]) < 1e-13
    assert res["hbm"][1] < 1e-13 and abs(res["hbm"][1] - res["lds"][1]) < 1e-14


def test_long_axis_adjoint_dot(hip):
    import torch
    rng = np.random.default_rng(12)
    NE, Nk, Nb = 30721, 2, 2
    eps, wE = _uniform(NE)
    g = hip.Grid(eps, wE, np.ones(Nk), Nb)
    plan = hip.FFTHilbertPlan(g, verify_roundoff=False)
    SL, SG, Rb = (hip.to_device(x) for x in _fields(rng, NE, Nk, Nb, 3))
    SR, _ = hip.retarded_self_energy_fft(plan, SL, SG)
    Lb, Gb = torch.zeros_like(SL), torch.zeros_like(SG)
    hip.retarded_self_energy_adjoint(plan, Rb, Lb, Gb)
    lhs = _ip(Rb, SR)
    rhs = _ip(Lb, SL) + _ip(Gb, SG)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_path_limits(hip):
    from qclhip import _lib
    eps, wE = _uniform(100)
    g = hip.Grid(eps, wE, np.ones(1), 1)
    plan = hip.FFTHilbertPlan(g)  # L = 256
    with pytest.raises(_lib.UnsupportedError):
        plan.set_path("hbm")
    eps, wE = _uniform(9000)
    g = hip.Grid(eps, wE, np.ones(1), 1)
    plan = hip.FFTHilbertPlan(g)  # L = 32768
    with pytest.raises(_lib.UnsupportedError):
        plan.set_path("lds")
    eps, wE = _uniform(131073)
    g = hip.Grid(eps, wE, np.ones(1), 1)
    with pytest.raises(_lib.UnsupportedError):
        hip.FFTHilbertPlan(g)
