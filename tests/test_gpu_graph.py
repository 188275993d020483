"""A captured SCBA iteration (HIP graph) replays exactly the eager iteration:
several replays, with unrelated eager work and synchronizations between them,
each equal to the eager run from the same state (rows and fields)."""
import copy

import numpy as np
import pytest

import scba_fixture as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _device_state(hip, prob, state):
    from qclhip import scba
    dev = lambda x: hip.to_device(x)
    return scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                         dev(state["scatG"][i]))
                                 for i, m in enumerate(prob["mechanisms"])},
                          scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                          scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))


@pytest.mark.parametrize("lo_kind", ["low_rank", "dense"])
def test_graph_replays_equal_eager(hip, lo_kind):
    import torch
    from qclhip import scba
    prob = fx.make_problem(N=4, NE=64, Nk=3, lo_kind=lo_kind)
    full, state, green = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    n_rep = 4
    s_eager = _device_state(hip, prob, state)
    ws_e = scba.ScbaWorkspace(p)
    eager = [scba.scba_iteration(p, s_eager, ws_e) for _ in range(1 + n_rep)]
    s_graph = _device_state(hip, prob, state)
    ws_g = scba.ScbaWorkspace(p)
    graph = scba.IterationGraph(p, s_graph, ws_g)  # runs eager iteration 1, captures the next
    scratch = torch.zeros_like(s_graph.green.GR)
    for k in range(n_rep):
        row = graph.replay()
        ref = eager[1 + k]
        for key in ("lam", "number", "holes", "Jbar", "roundoff", "r_D", "r_A", "r_K", "r_PSD", "r_caus",
                    "r_Sigma", "r_lambda"):
            assert row[key] == ref[key], (k, key, row[key], ref[key])
        assert np.array_equal(row["populations"], ref["populations"]), k
        # unrelated eager work and a synchronization between replays
        hip.spectral_function(g, s_graph.green.GR, A=scratch)
        hip.number_functional(g, s_graph.green.GL)
        torch.cuda.synchronize()
    for k in ("GR", "GL", "GG"):
        assert torch.equal(getattr(s_graph.green, k), getattr(s_eager.green, k)), k
    for m in p.mechanisms:
        for a, b in zip(s_graph.scattering[m.name].parts(), s_eager.scattering[m.name].parts()):
            assert torch.equal(a, b), m.name
    for a, b in zip(s_graph.plus.parts() + s_graph.minus.parts(), s_eager.plus.parts() + s_eager.minus.parts()):
        assert torch.equal(a, b)


def test_deferred_results_match_eager(hip):
    import torch
    w = fx.syn.gf_workload(N=6, NE=40, Nk=3)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 6)
    inj = fx.syn.injection_family(w)
    plus = fx.syn.diag_fields_device(inj)
    scat = fx.syn.diag_fields_device(w)
    h = hip.h_to_device(w["h"])
    st, info = hip.current_green_production(g, h, [plus, scat], 1.0, spectral=False)
    J = hip.boundary_flux_bar(g, plus[1], plus[2], st.GL, st.GG)
    pops = hip.state_populations(g, st.GL, 1.0)
    num = hip.number_functional(g, st.GL)
    buf = torch.zeros(32, dtype=torch.float64, device="cuda")
    g.ctx.defer(buf)
    try:
        st2, dinfo = hip.current_green_production(g, h, [plus, scat], 1.0, spectral=False)
        dJ = hip.boundary_flux_bar(g, plus[1], plus[2], st2.GL, st2.GG)
        dp = hip.state_populations(g, st2.GL, 1.0)
        dn = hip.number_functional(g, st2.GL)
    finally:
        g.ctx.defer(None)
    g.ctx.synchronize()
    host = buf.cpu().numpy()
    vals = dinfo.read(host)
    assert list(vals) == [info["lam"], info["number"], info["holes"], info["a"], info["c"]]
    assert dJ.read(host) == J and dn.read(host) == num
    assert np.array_equal(dp.read(host), pops)


def test_mixing_totals_equal_family_sums(hip):
    """The totals written by the mixing pass and loaded as one family by the
    next GF solve give bitwise the iterations that sum the families at every
    solve (the workspace's cache cleared before each iteration)."""
    import torch
    from qclhip import scba
    prob = fx.make_problem(N=5, NE=48, Nk=3)
    full, state, green = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    s_c, s_f = _device_state(hip, prob, state), _device_state(hip, prob, state)
    ws_c, ws_f = scba.ScbaWorkspace(p), scba.ScbaWorkspace(p)
    for it in range(4):
        ws_f.totals_valid = False  # sum the families at this solve
        rc, rf = scba.scba_iteration(p, s_c, ws_c), scba.scba_iteration(p, s_f, ws_f)
        assert ws_c.totals_valid
        for key in ("lam", "number", "Jbar", "r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma"):
            assert rc[key] == rf[key], (it, key)
        for a, b in zip(ws_c.totals.parts(), ws_f.totals.parts()):
            assert torch.equal(a, b), it
    for k in ("GR", "GL", "GG"):
        assert torch.equal(getattr(s_c.green, k), getattr(s_f.green, k)), k


def test_replay_refused_after_scratch_grows(hip):
    """A graph keeps the scratch address of its capture: after an eager call
    on the same context reallocates the scratch (a Poisson solve needs far more
    than a number functional), launch is refused with ArgumentError instead of
    replaying into freed memory; a new capture works again."""
    import ctypes as ct
    import torch
    from qclhip import _lib
    ctx = hip.Context(0)  # its own scratch, sized only by this test
    w = fx.syn.gf_workload(N=4, NE=32, Nk=2)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 4, ctx=ctx)
    SR, SL, SG = fx.syn.diag_fields_device(w)
    st, _ = hip.current_green_production(g, hip.h_to_device(w["h"]), [(SR, SL, SG)], 1.0, spectral=False)
    ref = hip.number_functional(g, st.GL)  # eager: sizes the scratch
    side = torch.cuda.Stream()
    base = torch.cuda.current_stream()
    buf = torch.zeros(4, dtype=torch.float64, device="cuda")

    def capture():
        torch.cuda.synchronize()
        ctx.set_stream(side)
        try:
            ctx.defer(buf)
            ctx.graph_begin()
            try:
                hip.number_functional(g, st.GL)
            finally:
                graph = ctx.graph_end()
                ctx.defer(None)
        finally:
            ctx.set_stream(base)
        return graph

    def replay(graph):
        buf.zero_()
        torch.cuda.synchronize()
        ctx.set_stream(side)
        try:
            graph.launch()
            ctx.synchronize()
        finally:
            ctx.set_stream(base)
        return float(buf[0].item())

    graph = capture()
    assert replay(graph) == ref
    # grow the scratch: the bordered Poisson system of N_z = 1024 needs ~8 MB
    Nz = 1024
    dev = lambda x: torch.as_tensor(x, device="cuda", dtype=torch.float64)
    nbar = dev(np.full(Nz, 1.0))
    U = torch.empty(Nz, dtype=torch.float64, device="cuda")
    out = (ct.c_double * 3)()
    args = (dev(np.full(Nz, 12.0)), dev(np.full(Nz, 0.1)), dev(np.full(Nz, 1.0)))
    ctx.call("qclhip_periodic_poisson", Nz, *(ct.c_void_p(t.data_ptr()) for t in args), ct.c_double(1.0),
             ct.c_void_p(nbar.data_ptr()), ct.c_void_p(U.data_ptr()), out)
    with pytest.raises(_lib.ArgumentError, match="reallocated"):
        replay(graph)
    graph = capture()
    assert replay(graph) == ref
    del graph, g
    ctx.close()
