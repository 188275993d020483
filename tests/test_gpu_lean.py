"""The memory-lean SCBA iteration (one candidate family live at a time, the
residual suite split into a state pass and an r_K pass, totals folded across
mixing calls) against the standard device iteration (bit for bit) and
against the oracle with the five mechanisms of the full-iteration workload."""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu

TOL = 1e-12
ROW_KEYS = ("lam", "number", "holes", "Jbar", "roundoff", "r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma",
            "r_lambda")


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _setup(hip, prob, state, full):
    from qclhip import scba
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    return p


def _state(hip, prob, state):
    from qclhip import scba
    dev = lambda x: hip.to_device(x)
    return scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                         dev(state["scatG"][i]))
                                 for i, m in enumerate(prob["mechanisms"])},
                          scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                          scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))


@pytest.mark.parametrize("N,NE", [(5, 48), (30, 45)])
def test_lean_equals_standard_iteration(hip, N, NE):
    import torch
    from qclhip import scba
    prob = fx.make_problem(N=N, NE=NE, Nk=3)
    full, state, green = fx.oracle_state(prob)
    p = _setup(hip, prob, state, full)
    s_std, s_lean = _state(hip, prob, state), _state(hip, prob, state)
    ws_std, ws_lean = scba.ScbaWorkspace(p), scba.LeanWorkspace(p)
    for it in range(3):
        a = scba.scba_iteration(p, s_std, ws_std)
        b = scba.scba_iteration_lean(p, s_lean, ws_lean)
        for k in ROW_KEYS:
            assert a[k] == b[k], (it, k, a[k], b[k])
        assert np.array_equal(a["populations"], b["populations"])
    for k in ("GR", "GL", "GG"):
        assert torch.equal(getattr(s_std.green, k), getattr(s_lean.green, k)), k
    assert s_lean.green.A is None
    for m in p.mechanisms:
        for x, y in zip(s_std.scattering[m.name].parts(), s_lean.scattering[m.name].parts()):
            assert torch.equal(x, y), m.name
    for x, y in zip(s_std.plus.parts() + s_std.minus.parts(), s_lean.plus.parts() + s_lean.minus.parts()):
        assert torch.equal(x, y)
    for x, y in zip(ws_std.totals.parts(), ws_lean.totals.parts()):
        assert torch.equal(x, y)


def test_lean_five_mechanisms_vs_oracle(hip, orc):
    """Acoustic, alloy (momentum independent), IFR (separable), LO (low rank,
    thermal), impurity (low rank): three lean iterations against the oracle."""
    from qclhip import scba
    prob = fx.make_problem5(N=6, NE=128, Nk=4)
    full, state, green = fx.oracle_state(prob)
    p = _setup(hip, prob, state, full)
    s = _state(hip, prob, state)
    ws = scba.LeanWorkspace(p)
    runner = orc.ScbaOracle(full, state)
    for it in range(3):
        ro = runner.iterate(measure=True)
        rd = scba.scba_iteration_lean(p, s, ws)
        assert abs(rd["lam"] - ro["lam"]) < TOL * abs(ro["lam"])
        assert abs(rd["Jbar"] - ro["Jbar"]) < TOL * max(abs(ro["Jbar"]), 1e-300) + 1e-14
        assert float(rel(rd["populations"], ro["populations"])) < TOL
        assert rd["roundoff"] < 1e-12
        assert rd["r_D"] < 1e-13 and ro["r_D"] < 1e-13
        for k in ("r_A", "r_K", "r_Sigma"):
            assert abs(rd[k] - ro[k]) <= 1e-9 * ro[k] + 1e-13, (it, k, rd[k], ro[k])
        for k in ("r_PSD", "r_caus", "r_lambda"):
            assert abs(rd[k] - ro[k]) <= 1e-12, (it, k, rd[k], ro[k])
    for k in ("GR", "GL", "GG"):
        assert float(rel(hip.to_host(getattr(s.green, k)), state[k])) < TOL, k
    for i, m in enumerate(prob["mechanisms"]):
        for c, key in enumerate(("scatR", "scatL", "scatG")):
            assert float(rel(hip.to_host(s.scattering[m["name"]].parts()[c]), state[key][i])) < TOL, (m["name"], c)


def _random_state(hip, rng, NE, Nk, Nb):
    import torch
    mk = lambda: np.asfortranarray(0.3 * (rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb))))
    return [hip.to_device(mk()) for _ in range(9)]


@pytest.mark.parametrize("Nb", [4, 17, 33])
def test_residual_passes_equal_suite(hip, Nb):
    """State pass (A given or formed from G^R) and r_K pass reproduce the
    terms of the whole suite bit for bit, on random (non-PSD) states."""
    rng = np.random.default_rng(50 + Nb)
    NE, Nk = 21, 2
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=Nb)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], Nb)
    h = hip.h_to_device(w["h"])
    GR, GL, GG, SR, SL, SG, C1, C2, _ = _random_state(hip, rng, NE, Nk, Nb)
    A = hip.spectral_function(g, GR)
    st = hip.GreenState(GR, GL, GG, A, None)
    full = hip.residual_suite(g, h, st, [(SR, SL, SG)], [C1, C2])
    part = hip.residual_state(g, h, st, [(SR, SL, SG)])
    noA = hip.residual_state(g, h, hip.GreenState(GR, GL, GG, None, None), [(SR, SL, SG)])
    rK = hip.residual_keldysh(g, st, [C1, C2])
    for k in ("r_D", "r_A", "r_PSD", "r_caus"):
        assert part[k] == full[k] and noA[k] == full[k], k
    assert part["r_K"] == 0.0 and rK == full["r_K"]
    assert full["r_PSD"] > 0 and full["r_K"] > 0


def test_mix_accumulate_flags(hip):
    """Totals folded over two calls equal one call's totals; a candidate
    total sums the candidates; a total aliasing a pair is refused."""
    import torch
    from qclhip import ArgumentError
    gen = torch.Generator().manual_seed(3)
    mk = lambda: torch.randn(4096, dtype=torch.complex128, generator=gen).cuda()
    olds = [mk() for _ in range(3)]
    cands = [mk() for _ in range(3)]
    o1 = [x.clone() for x in olds]
    t_one = torch.empty_like(olds[0])
    r1 = hip.mix_fields_residual(list(zip(o1, cands)), 0.3, totals=[(t_one, [0, 1, 2])])
    o2 = [x.clone() for x in olds]
    t_two = torch.empty_like(olds[0])
    csum = torch.empty_like(olds[0])
    r2 = hip.mix_fields_residual(list(zip(o2[:2], cands[:2])), 0.3,
                                 totals=[(t_two, [0, 1]), (csum, [0, 1], "candidate")])
    r2 += hip.mix_fields_residual([(o2[2], cands[2])], 0.3,
                                  totals=[(t_two, [0], "accumulate"), (csum, [0], "accumulate,candidate")])
    assert torch.equal(t_one, t_two) and r1 == r2
    for x, y in zip(o1, o2):
        assert torch.equal(x, y)
    assert torch.equal(csum, (cands[0] + cands[1]) + cands[2])
    with pytest.raises(ArgumentError):
        hip.mix_fields_residual([(o2[0], cands[0])], 0.3, totals=[(o2[0], [0])])
    with pytest.raises(ArgumentError):
        hip.mix_fields_residual([(o2[0], cands[0])], 0.3, totals=[(t_two, [0], "sum")])
