"""Anderson mixing on the device against the oracle (production_mixing.jl:
36-317): the history append with its inner products, the combination, the
causality guard (every failure code, first-failure order, N_b up to 24) and
full SCBA iterations with Anderson mixing (FP64 bar 1e-12; the coefficient
solve, history and reset decisions must agree exactly)."""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx
from test_oracle_anderson import _causal_block, _families

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _rand_fields(rng, k, shape=(7, 3, 5, 5)):
    return [np.asfortranarray(rng.normal(size=shape) + 1j * rng.normal(size=shape)) for _ in range(k)]


def test_append_and_combine_vs_oracle(hip, orc):
    import torch
    rng = np.random.default_rng(1)
    nf = 6
    state, cand = _rand_fields(rng, nf), _rand_fields(rng, nf)
    olds = [_rand_fields(rng, nf) for _ in range(2)]
    dev = lambda xs: [hip.to_device(x) for x in xs]
    ds, dc = dev(state), dev(cand)
    saved = [torch.empty_like(t) for t in ds]
    resid = [torch.empty_like(t) for t in ds]
    dots = hip.anderson_append(ds, dc, saved, resid, [dev(o) for o in olds])
    for a, b in zip(saved, ds):
        assert torch.equal(a, b)
    for r, s_, c in zip(resid, ds, dc):
        assert torch.equal(r, c - s_)
    R = [np.asfortranarray(c - s_) for c, s_ in zip(cand, state)]
    for h in range(2):
        for f in range(nf):
            ref = orc.anderson_inner([olds[h][f]], [R[f]])
            assert abs(dots[f, h] - ref) <= 1e-13 * abs(ref) + 1e-12
    assert abs(dots[:, 2].sum() - orc.anderson_inner(R, R)) <= 1e-13 * orc.anderson_inner(R, R)
    coef = [0.3, -0.5, 1.2]
    X = [state, olds[0], olds[1]]
    Rh = [cand, olds[1], state]
    out = [torch.empty_like(t) for t in ds]
    hip.anderson_combine(coef, 0.7, [dev(x) for x in X], [dev(r) for r in Rh], out)
    ref = [np.zeros_like(x) for x in state]
    orc.anderson_combine(coef, 0.7, X, Rh, ref)
    for o, r in zip(out, ref):
        assert float(rel(hip.to_host(o), r)) < 1e-15


@pytest.mark.parametrize("n", [5, 24])
@pytest.mark.parametrize("code", [0, 1, 2, 3, 4, 5])
def test_guard_codes_vs_oracle(hip, orc, n, code):
    rng = np.random.default_rng(10 + code + n)
    R, L, G = _causal_block(rng, n)
    if code == 1:
        L = L.copy(); L[1, 2] = np.nan
    elif code == 2:
        R = R + 1.5j * np.eye(n) * np.abs(np.linalg.eigvalsh(-(R - R.conj().T) / 2j)).max()
    elif code == 3:
        B = rng.normal(size=(n, n)); L = L + 0.1 * (B + B.T)
    elif code == 4:
        w, V = np.linalg.eigh(-1j * L)
        w[0] = -0.2 * abs(w).max()
        L = 1j * (V @ np.diag(w) @ V.conj().T)
    elif code == 5:
        C = rng.normal(size=(n, n)); R = R + 1e-3 * (C + C.T) - 1e-3j * np.eye(n)
    fams = _families([_causal_block(rng, n), (R, L, G)], NE=4, Nk=3)
    ref = orc.anderson_guard(fams)
    g = hip.Grid(np.linspace(0, 1, 4), np.full(4, 0.25), np.full(3, 1 / 3), n)
    got = hip.anderson_guard(g, [tuple(hip.to_device(x) for x in f) for f in fams])
    assert got == ref
    assert got == ((0, 0) if code == 0 else (12 + 1, code))


def test_guard_first_failure_order_vs_oracle(hip, orc):
    rng = np.random.default_rng(3)
    fams = _families([_causal_block(rng, 4) for _ in range(3)], NE=5, Nk=2)
    fams[2][0][4, 1] = fams[2][0][4, 1] + 5j * np.eye(4)
    fams[1][2][3, 0, 1, 1] = np.inf
    fams[1][0][3, 1] = fams[1][0][3, 1] + 5j * np.eye(4)
    ref = orc.anderson_guard(fams)
    g = hip.Grid(np.linspace(0, 1, 5), np.full(5, 0.2), np.full(2, 0.5), 4)
    got = hip.anderson_guard(g, [tuple(hip.to_device(x) for x in f) for f in fams])
    assert got == ref == (10 + 3 * 2 + 0 + 1, 1)


def test_scba_anderson_iterations_vs_oracle(hip, orc):
    from qclhip import scba
    prob = fx.make_problem(N=4, NE=48, Nk=3, lo_kind="low_rank")
    full, state, _ = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"], mixing="anderson", anderson_depth=4)
    dev = hip.to_device
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                      dev(state["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                       scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    runner = orc.ScbaOracle(full, state)
    mixer = orc.AndersonOracle(depth=4)
    ws = scba.ScbaWorkspace(p)
    reasons = []
    for it in range(7):
        ro = runner.iterate_anderson(mixer)
        rd = scba.scba_iteration(p, s, ws)
        assert abs(rd["lam"] - ro["lam"]) < TOL * abs(ro["lam"]), it
        assert float(rel(rd["populations"], ro["populations"])) < TOL, it
        assert abs(rd["r_Sigma"] - ro["r_Sigma"]) <= 1e-10 * ro["r_Sigma"], it
        a, b = rd["anderson"], ro["anderson"]
        assert a["reset_reason"] == b["reset_reason"], (it, a["reset_reason"], b["reset_reason"])
        assert a["guard_first_failure"] == b["first_failure"], it
        assert np.allclose(a["coefficients"], b["coefficients"], rtol=1e-8, atol=1e-10), it
        reasons.append(a["reset_reason"])
    for i, m in enumerate(prob["mechanisms"]):
        for k, key in enumerate(("scatR", "scatL", "scatG")):
            err = float(rel(hip.to_host(s.scattering[m["name"]].parts()[k]), state[key][i]))
            assert err < 1e-11, (m["name"], key)
    assert "none" in reasons
