"""Batched SCBA iterations over uneven instances (LeanBatch: one batched GF
solve, then each instance's lean tail; captured as one HIP graph per
iteration) against the same instances iterated one at a time (bit for bit)
and against the oracle, in both launch modes of the batched solve."""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu

SHAPES = ((3, 40, 3), (5, 64, 2), (9, 33, 3), (17, 50, 2))  # (N, NE, Nk)
KEYS = ("lam", "number", "holes", "Jbar", "roundoff", "r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma", "r_lambda")


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _instance(hip, N, NE, Nk, seed):
    from qclhip import scba
    prob = fx.make_problem(N=N, NE=NE, Nk=Nk, seed=seed)
    full, state, green = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], N)
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    dev = lambda x: hip.to_device(x)

    def fresh():
        return scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                             dev(state["scatG"][i]))
                                     for i, m in enumerate(prob["mechanisms"])},
                              scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                              scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    return p, fresh, full, state


@pytest.mark.parametrize("mode", ["persistent", "per_instance"])
def test_lean_batch_graph_equals_one_by_one(hip, orc, mode):
    import torch
    from qclhip import scba
    inst = [_instance(hip, N, NE, Nk, 40 + i) for i, (N, NE, Nk) in enumerate(SHAPES)]
    ps = [x[0] for x in inst]
    s_one = [x[1]() for x in inst]
    s_bat = [x[1]() for x in inst]
    w_one = [scba.LeanWorkspace(p) for p in ps]
    w_bat = [scba.LeanWorkspace(p) for p in ps]
    runners = [orc.ScbaOracle(x[2], x[3]) for x in inst]
    for p, s, w in zip(ps, s_bat, w_bat):
        scba.scba_iteration_lean(p, s, w)       # first iteration: sums the families, sizes scratch
    lb = scba.LeanBatch(ps, s_bat, w_bat)
    hip.batch_set_mode(lb.batch, mode)
    rows_b = [lb.capture() and [s.history[-1] for s in s_bat]]  # capture() runs one eager batched iteration
    for _ in range(2):
        rows_b.append(lb.replay())
    rows_o = []
    for it in range(4):
        rows_o.append([scba.scba_iteration_lean(p, s, w) for p, s, w in zip(ps, s_one, w_one)])
    for k in range(3):  # batched iterations 2..4 against one-by-one iterations 2..4
        for i in range(len(ps)):
            a, b = rows_o[k + 1][i], rows_b[k][i]
            for key in KEYS:
                assert a[key] == b[key], (mode, k, i, key, a[key], b[key])
            assert np.array_equal(a["populations"], b["populations"])
    for i, (p, a, b) in enumerate(zip(ps, s_one, s_bat)):
        for x, y in zip((a.green.GR, a.green.GL, a.green.GG), (b.green.GR, b.green.GL, b.green.GG)):
            assert torch.equal(x, y), i
        for m in p.mechanisms:
            for x, y in zip(a.scattering[m.name].parts(), b.scattering[m.name].parts()):
                assert torch.equal(x, y), (i, m.name)
    for i, r in enumerate(runners):  # four oracle iterations
        for _ in range(4):
            ro = r.iterate(measure=True)
        rb = rows_b[-1][i]
        assert abs(rb["lam"] - ro["lam"]) < 1e-12 * abs(ro["lam"])
        assert abs(rb["Jbar"] - ro["Jbar"]) < 1e-12 * abs(ro["Jbar"]) + 1e-14
        assert float(rel(rb["populations"], ro["populations"])) < 1e-12


def test_batch_autotune_picks_a_mode(hip):
    from qclhip import scba
    inst = [_instance(hip, N, NE, Nk, 60 + i) for i, (N, NE, Nk) in enumerate(SHAPES)]
    ps = [x[0] for x in inst]
    ss = [x[1]() for x in inst]
    ws = [scba.LeanWorkspace(p) for p in ps]
    for p, s, w in zip(ps, ss, ws):
        scba.scba_iteration_lean(p, s, w)
    lb = scba.LeanBatch(ps, ss, ws)
    t = lb.autotune()
    assert t["mode"] in ("persistent", "per_instance") and t["persistent_ms"] > 0 and t["per_instance_ms"] > 0
    assert hip.batch_mode(lb.batch) == t["mode"]
    lb.capture()
    lb.replay()
