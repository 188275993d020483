"""The full-iteration workload (a six-layer GaAs/Al0.15Ga0.85As period; five
mechanisms: acoustic and alloy momentum independent, IFR separable, LO low
rank and thermal, impurity low rank) at reduced (N, NE, Nk): lean device
iterations against the oracle in FP64, and the FP32+refine GF solve inside
the iteration against the oracle's FP64 observables."""
import math

import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def period_problem(orc, N=8, NE=256, Nk=6, Nz=128):
    syn = fx.syn
    w = syn.two_well_period(N=N, NE=NE, Nk=Nk, Nz=Nz)
    kern = {name: orc.make_kernel(kind, a, b, c, rank=r) for name, (kind, a, b, c, r) in w["kernels"].items()}
    hw = syn.HW_LO_EV / syn.E0_EV
    base = dict(NE=NE, Nk=Nk, Nb=N, gs=2, eps=w["eps"], wE=w["wE"], wk=w["wk"], h=np.asfortranarray(w["h"]),
                Tplus=np.asfortranarray(w["Tplus"]), Tminus=np.asfortranarray(w["Tminus"]), Ep=w["Ep"], hwLO=hw,
                NLO=1.0 / math.expm1(hw / w["kbt"]), alpha_mix=0.3, kbt=w["kbt"], eta_seed=0.05, energy_chunk=16)
    green, _ = fx.oracle_seed(base)
    peak = lambda name: float(np.abs(orc.contract(kern[name], green["GL"], w["wk"], 1.0, 16)).max())
    qK = syn.calibrated_qk(peak, w)
    mech = []
    for name in syn.PERIOD_MECHANISMS:
        kind, a, b, c, r = w["kernels"][name]
        mech.append(dict(name=name, kind=kind, a=a, b=b, c=c, rank=r, qK=qK[name], is_lo=name == "LO"))
    return dict(base, mechanisms=mech)


def _device(hip, prob, state, full, precision="fp64"):
    from qclhip import scba
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    g.set_precision(precision)
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    dev = lambda x: hip.to_device(x)
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                      dev(state["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                       scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    return p, s, scba.LeanWorkspace(p)


def test_period_lean_iterations_vs_oracle(hip, orc):
    from qclhip import scba
    prob = period_problem(orc)
    full, state, green = fx.oracle_state(prob)
    p, s, ws = _device(hip, prob, state, full)
    runner = orc.ScbaOracle(full, state)
    for it in range(3):
        ro = runner.iterate(measure=True)
        rd = scba.scba_iteration_lean(p, s, ws)
        assert abs(rd["lam"] - ro["lam"]) < TOL * abs(ro["lam"])
        assert abs(rd["Jbar"] - ro["Jbar"]) < TOL * abs(ro["Jbar"]) + 1e-15, (rd["Jbar"], ro["Jbar"])
        assert float(rel(rd["populations"], ro["populations"])) < TOL
        assert rd["roundoff"] < 1e-12 and rd["r_D"] < 1e-13
        for k in ("r_A", "r_K", "r_Sigma"):
            assert abs(rd[k] - ro[k]) <= 1e-9 * ro[k] + 1e-13, (it, k, rd[k], ro[k])
        for k in ("r_PSD", "r_caus", "r_lambda"):
            assert abs(rd[k] - ro[k]) <= 1e-12, (it, k, rd[k], ro[k])
    assert abs(ro["Jbar"]) > 0
    for k in ("GR", "GL", "GG"):
        assert float(rel(hip.to_host(getattr(s.green, k)), state[k])) < TOL, k
    for i, m in enumerate(prob["mechanisms"]):
        for c, key in enumerate(("scatR", "scatL", "scatG")):
            assert float(rel(hip.to_host(s.scattering[m["name"]].parts()[c]), state[key][i])) < TOL, (m["name"], c)


def test_period_fp32_refine_iteration_vs_fp64_oracle(hip, orc):
    """FP32 factorization with FP64 refinement in the iteration's GF solve;
    mixing, normalization and observables in FP64: every observable of two
    iterations within 1e-8 of the oracle's FP64 values."""
    from qclhip import scba
    prob = period_problem(orc)
    full, state, green = fx.oracle_state(prob)
    p, s, ws = _device(hip, prob, state, full, precision="fp32_refine")
    runner = orc.ScbaOracle(full, state)
    for it in range(2):
        ro = runner.iterate(measure=True)
        rd = scba.scba_iteration_lean(p, s, ws)
        assert abs(rd["Jbar"] - ro["Jbar"]) <= 1e-8 * abs(ro["Jbar"]), (it, rd["Jbar"], ro["Jbar"])
        assert abs(rd["lam"] - ro["lam"]) <= 1e-10 * abs(ro["lam"])
        assert float(rel(rd["populations"], ro["populations"])) < 1e-8
        assert rd["r_D"] < 1e-12
    for k in ("GR", "GL", "GG"):
        assert float(rel(hip.to_host(getattr(s.green, k)), state[k])) < 1e-8, k
    assert p.grid.refine_fallbacks() == 0
