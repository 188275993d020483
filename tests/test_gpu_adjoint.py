"""Adjoint (A19) of the GF-solve and SCBA operators.

No reference counterpart exists, so the reverse operators are pinned by what
the mathematics fixes: dot-product identities <ybar, J x> = <J^H ybar, x>
for every linear operator, central differences of the forward operators for
the nonlinear current_green, and - for the full fixed-point adjoint - central
differences of the ORACLE's converged SCBA current J(h +- eps dh) against the
device gradient. Inner products are real: <a, b> = Re sum conj(a) b.
"""
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


def _ip(a, b):
    import torch
    return float(torch.sum(torch.conj(a) * b).real)


def _rand(hip, rng, NE, Nk, Nb, scale=1.0):
    x = scale * (rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb)))
    return hip.to_device(np.asfortranarray(x))


def _grid(hip, NE=40, Nk=3, Nb=5, seed=1):
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=seed)
    return hip.Grid(w["eps"], w["wE"], w["wk"], Nb), w


@pytest.mark.parametrize("delta", [0.37, -0.37, 0.0, 2.5, 1e-3])
def test_shift_transpose_dot(hip, delta):
    rng = np.random.default_rng(5)
    g, w = _grid(hip)
    plan = hip.EnergyShiftPlan(g, delta)
    X, Y = _rand(hip, rng, g.NE, g.Nk, g.Nb), _rand(hip, rng, g.NE, g.Nk, g.Nb)
    lhs = _ip(Y, hip.apply_energy_shift(plan, X))
    rhs = _ip(hip.shift_transpose(plan, Y), X)
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + 1.0)


@pytest.mark.parametrize("kind", ["dense", "momentum_independent", "separable_ifr", "low_rank"])
def test_kernel_adjoint_dot(hip, kind):
    rng = np.random.default_rng(6)
    Nb, Nk = 3, 4
    g, w = _grid(hip, NE=24, Nk=Nk, Nb=Nb)
    N2, Nc = Nb * Nb, Nk * Nb * Nb
    cm = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    if kind == "dense":
        K = hip.ProductionKernel(g, kind, cm(Nc, Nc))
    elif kind == "momentum_independent":
        K = hip.ProductionKernel(g, kind, cm(N2, N2))
    elif kind == "separable_ifr":
        K = hip.ProductionKernel(g, kind, cm(N2, 2), cm(2, N2), cm(Nk, Nk), rank=2)
    else:
        K = hip.ProductionKernel(g, kind, cm(Nc, 3), cm(3, Nc), rank=3)
    Ka = K.adjoint()
    G, Sb = _rand(hip, rng, g.NE, Nk, Nb), _rand(hip, rng, g.NE, Nk, Nb)
    lhs = _ip(Sb, hip.production_static_contraction(K, G, 0.7))
    rhs = _ip(hip.production_static_contraction(Ka, Sb, 0.7), G)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_retarded_self_energy_adjoint_dot(hip):
    import torch
    rng = np.random.default_rng(7)
    g, w = _grid(hip, NE=100, Nk=2, Nb=3)
    plan = hip.FFTHilbertPlan(g, g.ncols, verify_roundoff=False)
    SL, SG, Rb = (_rand(hip, rng, g.NE, g.Nk, g.Nb) for _ in range(3))
    SR, _ = hip.retarded_self_energy_fft(plan, SL, SG)
    Lb, Gb = torch.zeros_like(SL), torch.zeros_like(SG)
    hip.retarded_self_energy_adjoint(plan, Rb, Lb, Gb)
    lhs = _ip(Rb, SR)
    rhs = _ip(Lb, SL) + _ip(Gb, SG)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


@pytest.mark.parametrize("delta", [0.5, -0.5])
def test_embedding_adjoint_dot(hip, delta):
    import torch
    rng = np.random.default_rng(8)
    g, w = _grid(hip, NE=48, Nk=3, Nb=7)
    plan = hip.EnergyShiftPlan(g, delta)
    T = rng.normal(size=(7, 7)) + 1j * rng.normal(size=(7, 7))
    Td = hip.matrix_to_device(T)
    X, Sb = _rand(hip, rng, g.NE, g.Nk, g.Nb), _rand(hip, rng, g.NE, g.Nk, g.Nb)
    S = hip.embedding_component_production(g, Td, X, plan)
    Xb = torch.zeros_like(X)
    hip.embedding_adjoint(g, Td, plan, Sb, Xb)
    lhs, rhs = _ip(Sb, S), _ip(Xb, X)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_boundary_flux_adjoint(hip):
    """J is bilinear in (S+, G): the central difference is exact."""
    import torch
    rng = np.random.default_rng(9)
    g, w = _grid(hip, NE=32, Nk=3, Nb=4)
    x = [_rand(hip, rng, g.NE, g.Nk, g.Nb) for _ in range(4)]
    d = [_rand(hip, rng, g.NE, g.Nk, g.Nb) for _ in range(4)]
    bars = [torch.zeros_like(v) for v in x]
    hip.boundary_flux_adjoint(g, 1.0, *x, *bars)
    t = 1e-3
    Jp = hip.boundary_flux_bar(g, *[a + t * b for a, b in zip(x, d)])
    Jm = hip.boundary_flux_bar(g, *[a - t * b for a, b in zip(x, d)])
    fd = (Jp - Jm) / (2 * t)
    ad = sum(_ip(b, v) for b, v in zip(bars, d))
    assert abs(fd - ad) <= 1e-10 * (abs(fd) + 1e-3)


@pytest.mark.parametrize("NE", [48, 45])
@pytest.mark.parametrize("branch", ["a", "c"])
def test_current_green_adjoint_fd(hip, branch, NE):
    """Directional derivative of F = <W<, G<> + <W>, G>> + <WR, G^R> through
    current_green (Dyson, Keldysh, paired map on both branches) against
    central differences in the families and h; NE = 45 leaves a partial last
    tile of the reverse kernel."""
    import torch
    rng = np.random.default_rng(10)
    prob = fx.make_problem(N=5, NE=NE, Nk=3)
    full, state, green = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    fams_h = list(zip([state["embR"]] + state["scatR"], [state["embL"]] + state["scatL"],
                      [state["embG"]] + state["scatG"]))
    fams = [tuple(hip.to_device(x) for x in f) for f in fams_h]
    h = hip.h_to_device(prob["h"])
    st, info = hip.current_green_production(g, h, fams, 1.0, spectral=False)
    # target below (a branch) or above (c branch) the raw number
    target = info["number"] * (0.8 if branch == "a" else 1.05)
    st, info = hip.current_green_production(g, h, fams, target, spectral=False)
    assert (info["c"] == 0.0) == (branch == "a")
    WR, WL, WG = (_rand(hip, rng, g.NE, g.Nk, g.Nb) for _ in range(3))
    (SRb, SLb, SGb), hb = hip.current_green_adjoint(g, fams, target, info, st, WL, WG, GRbar=WR)

    def F(fs, hh):
        s2, _ = hip.current_green_production(g, hh, fs, target, spectral=False)
        return _ip(WR, s2.GR) + _ip(WL, s2.GL) + _ip(WG, s2.GG)

    scale = float(torch.linalg.vector_norm(fams[1][0])) / np.sqrt(fams[1][0].numel())
    dR, dL, dG = (_rand(hip, rng, g.NE, g.Nk, g.Nb, scale) for _ in range(3))
    # S< and S> stay anti-Hermitian per block (the number functional's realness gate)
    dL = dL - torch.conj(dL.transpose(0, 1))
    dG = dG - torch.conj(dG.transpose(0, 1))
    dh = hip.h_to_device(0.01 * (rng.normal(size=prob["h"].shape) + 1j * rng.normal(size=prob["h"].shape)))
    t = 1e-5
    fp = [(fams[0][0] + t * dR, fams[0][1] + t * dL, fams[0][2] + t * dG)] + fams[1:]
    fm = [(fams[0][0] - t * dR, fams[0][1] - t * dL, fams[0][2] - t * dG)] + fams[1:]
    fd = (F(fp, h + t * dh) - F(fm, h - t * dh)) / (2 * t)
    ad = _ip(SRb, dR) + _ip(SLb, dL) + _ip(SGb, dG) + _ip(hb, dh)
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def _device_problem(hip, prob, full, state, alpha=None):
    from qclhip import scba
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"] if alpha is None else alpha, verify_roundoff=False)
    dev = lambda x: hip.to_device(x)
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                      dev(state["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                       scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    return g, p, s


def test_candidates_adjoint_dot(hip):
    """The candidate map (contractions, LO shifts, Hilbert, embedding) is
    linear in (G^R, G<, G>): <mu, cand(G)> = <cand^H(mu), G>."""
    import torch
    from qclhip import adjoint, scba
    rng = np.random.default_rng(13)
    prob = fx.make_problem(N=4, NE=64, Nk=3)
    full, state, green = fx.oracle_state(prob)
    g, p, s = _device_problem(hip, prob, full, state)
    G = [_rand(hip, rng, g.NE, g.Nk, g.Nb) for _ in range(3)]
    gs = hip.GreenState(G[0], G[1], G[2], None, None)
    scat = {m.name: scba.new_family(g) for m in p.mechanisms}
    scba._candidate_scattering(p, gs, scat)
    plus, minus = scba.new_family(g), scba.new_family(g)
    scba._candidate_embedding(p, gs, plus, minus)
    mu = adjoint._Families(p)
    lhs = 0.0
    for fam, out in zip(mu.all(), [plus, minus] + [scat[m.name] for m in p.mechanisms]):
        for x, y in zip(fam.parts(), out.parts()):
            x.copy_(_rand(hip, rng, g.NE, g.Nk, g.Nb))
            lhs += _ip(x, y)
    bars = [torch.zeros_like(x) for x in G]
    adjoint.ScbaAdjoint(p).candidates_vjp(mu, bars)
    rhs = sum(_ip(b, x) for b, x in zip(bars, G))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + 1.0)


def test_scba_map_adjoint_fd(hip):
    """One application of the SCBA map F (current_green, then candidates):
    <mu, dF[ds]> by central differences against <F^H mu, ds>."""
    import torch
    from qclhip import adjoint, scba
    rng = np.random.default_rng(14)
    prob = fx.make_problem(N=4, NE=64, Nk=3)
    full, state, green = fx.oracle_state(prob)
    g, p, s = _device_problem(hip, prob, full, state)
    fams = [s.plus, s.minus] + [s.scattering[m.name] for m in p.mechanisms]

    def F(ff):
        parts = [f.parts() for f in ff]
        gs, info = hip.current_green_production(g, p.h, parts, p.target, spectral=False)
        scat = {m.name: scba.new_family(g) for m in p.mechanisms}
        scba._candidate_scattering(p, gs, scat)
        plus, minus = scba.new_family(g), scba.new_family(g)
        scba._candidate_embedding(p, gs, plus, minus)
        return [plus, minus] + [scat[m.name] for m in p.mechanisms], gs, info

    out0, gs0, info0 = F(fams)
    mu = adjoint._Families(p)
    for fam in mu.all():
        for x in fam.parts():
            x.copy_(_rand(hip, rng, g.NE, g.Nk, g.Nb))
    bars = [torch.zeros_like(x) for x in (gs0.GR, gs0.GL, gs0.GG)]
    A = adjoint.ScbaAdjoint(p)
    A.candidates_vjp(mu, bars)
    parts = [f.parts() for f in fams]
    (SRb, SLb, SGb), _ = hip.current_green_adjoint(g, parts, p.target, info0, gs0, bars[1], bars[2],
                                                   GRbar=bars[0], want_hbar=False)
    ds = []
    for f in fams:
        dR = 1e-2 * _rand(hip, rng, g.NE, g.Nk, g.Nb)
        dL = 1e-2 * _rand(hip, rng, g.NE, g.Nk, g.Nb)
        dG = 1e-2 * _rand(hip, rng, g.NE, g.Nk, g.Nb)
        dL = dL - torch.conj(dL.transpose(0, 1))
        dG = dG - torch.conj(dG.transpose(0, 1))
        ds.append((dR, dL, dG))
    ad = sum(_ip(SRb, d[0]) + _ip(SLb, d[1]) + _ip(SGb, d[2]) for d in ds)
    t = 1e-4

    def shifted(sign):
        return [scba.Family(*(x + sign * t * dx for x, dx in zip(f.parts(), d))) for f, d in zip(fams, ds)]

    op, _, _ = F(shifted(1.0))
    om, _, _ = F(shifted(-1.0))
    fd = 0.0
    for fam, a, b in zip(mu.all(), op, om):
        for x, y, z in zip(fam.parts(), a.parts(), b.parts()):
            fd += _ip(x, (y - z) / (2 * t))
    assert abs(fd - ad) <= 1e-6 * abs(fd)


def _oracle_converged_J(full, state0, h, tol=1e-15, max_iter=4000):
    from oracle import oracle as orc
    f = dict(full, h=np.asfortranarray(h))
    st = copy.deepcopy(state0)
    runner = orc.ScbaOracle(f, st)
    prev = None
    for it in range(max_iter):
        J = runner.iterate()["Jbar"]
        if prev is not None and it > 5 and abs(J - prev) <= tol * abs(J):
            break
        prev = J
    return J


def test_scba_fixed_point_adjoint_vs_oracle_fd(hip, orc):
    """dJ/dh of the converged SCBA current on the device (fixed-point adjoint)
    against central differences of the oracle's converged current."""
    from qclhip import adjoint, scba
    prob = fx.make_problem(N=4, NE=64, Nk=3, alpha_mix=0.5)
    full, state, green = fx.oracle_state(prob)
    state0 = copy.deepcopy(state)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"], verify_roundoff=False)
    dev = lambda x: hip.to_device(x)
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                      dev(state["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                       scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    iters = adjoint.converge(p, s, tol=1e-15)
    assert iters < 5000
    hbar, meta = adjoint.ScbaAdjoint(p).gradient(s, tol=1e-13)
    print("adjoint iterations", meta["iterations"], meta["history"], flush=True)
    assert meta["change"] <= 1e-13
    hb = hbar.cpu().numpy().transpose(2, 1, 0)  # (Nk, Nb, Nb)
    rng = np.random.default_rng(12)
    h0 = prob["h"]
    for kind in ("hermitian", "general"):
        d = rng.normal(size=h0.shape) + 1j * rng.normal(size=h0.shape)
        if kind == "hermitian":
            d = d + np.conj(np.transpose(d, (0, 2, 1)))
        d *= 1e-3
        t = 1e-3
        Jp = _oracle_converged_J(full, state0, h0 + t * d)
        Jm = _oracle_converged_J(full, state0, h0 - t * d)
        fd = (Jp - Jm) / (2 * t)
        ad = float(np.real(np.sum(np.conj(hb) * d)))
        assert abs(fd - ad) <= 2e-6 * abs(fd), kind
