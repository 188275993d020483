"""Residual suite and self-energy residual on the device against the oracle
(residuals.jl:537-865): random states whose residuals are O(1) (relative bar
1e-10: the device's Householder/Sturm eigenvalues and the oracle's Jacobi
ones agree to rounding), closed
forms (G = 0, exact inverses, known negative and causal-violating
eigenvalues), non-finite blocks (Inf), every tile class N_b = 1..48, an SCBA
state (near-zero residuals, absolute bar) and the deferred-results mode."""
import numpy as np
import pytest

import scba_fixture as fx

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
KEYS = ("r_D", "r_A", "r_K", "r_PSD", "r_caus")


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _rand(rng, shape, scale=1.0):
    return np.asfortranarray(scale * (rng.normal(size=shape) + 1j * rng.normal(size=shape)))


def _random_state(NE, Nk, Nb, seed, nfam=2, ncand=2):
    rng = np.random.default_rng(seed)
    shape = (NE, Nk, Nb, Nb)
    s = dict(eps=np.linspace(-0.3, 0.4, NE), h=_rand(rng, (Nk, Nb, Nb), 0.2))
    for k in ("GR", "GL", "GG", "A"):
        s[k] = _rand(rng, shape)
    s["fams"] = [[_rand(rng, shape, 0.1) for _ in range(3)] for _ in range(nfam)]
    s["cands"] = [_rand(rng, shape, 0.1) for _ in range(ncand)]
    return s


def _oracle(orc, s):
    tot = [np.asfortranarray(sum(f[k] for f in s["fams"])) for k in range(3)]
    cand = np.asfortranarray(sum(s["cands"]))
    return orc.residual_suite(s["eps"], s["h"], s["GR"], s["GL"], s["GG"], s["A"], *tot, cand)


def _device(hip, s):
    import torch
    NE, Nk, Nb = s["GR"].shape[:3]
    g = hip.Grid(s["eps"], np.full(NE, 0.01), np.full(Nk, 1.0 / Nk), Nb)
    green = hip.GreenState(*(hip.to_device(s[k]) for k in ("GR", "GL", "GG", "A")),
                           torch.zeros((Nk, NE), dtype=torch.float64, device="cuda"))
    fams = [tuple(hip.to_device(x) for x in f) for f in s["fams"]]
    cands = [hip.to_device(x) for x in s["cands"]]
    return g, hip.h_to_device(s["h"]), green, fams, cands


def _gpu(hip, s):
    g, h, green, fams, cands = _device(hip, s)
    return hip.residual_suite(g, h, green, fams, cands)


@pytest.mark.parametrize("Nb", [1, 2, 3, 5, 8, 9, 16, 17, 24, 31, 33, 40, 48])
def test_random_states_vs_oracle(hip, orc, Nb):
    s = _random_state(NE=7, Nk=3, Nb=Nb, seed=100 + Nb)
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    for k in KEYS:
        assert ref[k] > 0.0, k  # the random state violates every gate
        assert abs(got[k] - ref[k]) <= 1e-10 * ref[k], (k, got[k], ref[k])


@pytest.mark.parametrize("Nb", [5, 24, 32])
def test_many_unsettled_blocks_vs_oracle(hip, orc, Nb):
    # every block of a random state fails its Cholesky screens: the list of
    # unsettled blocks outgrows the eigenvalue pass's grid (1024 waves)
    NE = 1201 if Nb == 5 else 520
    s = _random_state(NE=NE, Nk=2, Nb=Nb, seed=300 + Nb, nfam=1, ncand=1)
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    for k in KEYS:
        assert abs(got[k] - ref[k]) <= 1e-10 * ref[k], (k, got[k], ref[k])


def test_unsettled_blocks_among_settled(hip, orc):
    # a consistent state with a few blocks of A given a negative eigenvalue:
    # settled and unsettled blocks in the same tiles, the maximum from one
    s, _ = _scba_state(fx.make_problem(N=10, NE=48, Nk=3))
    rng = np.random.default_rng(9)
    Q, _ = np.linalg.qr(rng.normal(size=(10, 10)) + 1j * rng.normal(size=(10, 10)))
    for (e, m, neg) in ((3, 0, -0.2), (17, 1, -0.05), (40, 2, -0.4), (41, 2, -0.1)):
        d = np.linspace(0.1, 1.0, 10)
        d[0] = neg
        s["A"][e, m] = Q @ np.diag(d) @ Q.conj().T
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    assert ref["r_PSD"] > 0.3
    for k in ("r_PSD", "r_caus", "r_A"):
        assert abs(got[k] - ref[k]) <= 1e-10 * ref[k] + 1e-13, (k, got[k], ref[k])


def test_single_family_and_many_candidates(hip, orc):
    s = _random_state(NE=9, Nk=2, Nb=6, seed=5, nfam=1, ncand=8)
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    for k in KEYS:
        assert abs(got[k] - ref[k]) <= 1e-10 * ref[k], k


def _zero_state(NE=3, Nk=2, Nb=4):
    z = lambda: np.zeros((NE, Nk, Nb, Nb), np.complex128, order="F")
    return dict(eps=np.linspace(0, 1, NE), h=np.zeros((Nk, Nb, Nb), np.complex128, order="F"),
                GR=z(), GL=z(), GG=z(), A=z(), fams=[[z(), z(), z()]], cands=[z()])


def test_zero_green_function(hip):
    r = _gpu(hip, _zero_state())
    assert r["r_D"] == pytest.approx(1.0, rel=1e-15)
    assert r["r_A"] == 0.0 and r["r_K"] == 0.0 and r["r_PSD"] == 0.0 and r["r_caus"] == 0.0


def test_exact_inverse(hip):
    rng = np.random.default_rng(5)
    s = _zero_state(NE=5, Nk=3, Nb=12)
    for e in range(5):
        for m in range(3):
            S = 0.1 * (rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12)))
            s["fams"][0][0][e, m] = S
            s["GR"][e, m] = np.linalg.inv(s["eps"][e] * np.eye(12) - S)
    assert _gpu(hip, s)["r_D"] < 1e-14


@pytest.mark.parametrize("field,factor", [("A", 1.0), ("GL", 1j), ("GG", -1j)])
def test_negative_eigenvalue_closed_form(hip, field, factor):
    rng = np.random.default_rng(7)
    s = _zero_state()
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    s[field][1, 0] = factor * (Q @ np.diag([-0.3, 0.5, 1.0, 2.0]) @ Q.conj().T)
    assert _gpu(hip, s)["r_PSD"] == pytest.approx((0.3 - 64 * 4 * EPS * 2.0) / 2.0, rel=1e-13)


def test_causality_closed_form(hip):
    s = _zero_state()
    Q, _ = np.linalg.qr(np.random.default_rng(2).normal(size=(4, 4)) + 0j)
    s["fams"][0][0][2, 1] = 1j * (Q @ np.diag([0.2, -0.5, -1.0, 0.0]) @ Q.T)
    assert _gpu(hip, s)["r_caus"] == pytest.approx(0.2 / (1.0 + 1e-14), rel=1e-13)
    s["fams"][0][0][2, 1] = -1j * np.eye(4)
    assert _gpu(hip, s)["r_caus"] == 0.0


def test_spectral_and_keldysh_closed_form(hip):
    s = _zero_state()
    I = np.eye(4)
    s["GR"][:, :] = I
    s["A"][:, :] = 2 * I
    s["fams"][0][2][:, :] = -1.5j * I
    s["fams"][0][1][:, :] = 1.5j * I
    s["GL"][:, :] = 2j * I
    s["cands"][0][:, :] = 1j * I
    r = _gpu(hip, s)
    assert r["r_A"] == pytest.approx(0.5, rel=1e-14)
    assert r["r_K"] == pytest.approx(0.5, rel=1e-14)


@pytest.mark.parametrize("where", ["GR", "SR", "A", "cand"])
def test_nonfinite_block_gives_inf(hip, orc, where):
    s = _random_state(NE=4, Nk=2, Nb=5, seed=11)
    target = {"GR": s["GR"], "SR": s["fams"][1][0], "A": s["A"], "cand": s["cands"][0]}[where]
    target[2, 1, 3, 0] = np.nan
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    for k in KEYS:
        assert (got[k] == np.inf) == (ref[k] == np.inf), (k, got[k], ref[k])
        if ref[k] != np.inf:
            assert abs(got[k] - ref[k]) <= 1e-10 * ref[k], k
    assert any(got[k] == np.inf for k in KEYS)


def _scba_state(prob):
    full, state, green = fx.oracle_state(prob)
    from oracle import oracle as orc
    fams = [[state["embR"], state["embL"], state["embG"]]] + \
        [[state["scatR"][i], state["scatL"][i], state["scatG"][i]] for i in range(len(prob["mechanisms"]))]
    new = orc.current_green(prob["eps"], prob["h"], [f[0] for f in fams], [f[1] for f in fams],
                            [f[2] for f in fams], prob["wE"], prob["wk"], prob["gs"], full["target"])
    new = dict(GR=new["GR"], GL=new["GL"], GG=new["GG"], A=orc.spectral(new["GR"]))
    scat, plus, minus, emb = fx.oracle_candidates(prob, full["kernels"], new)
    cands = [emb[1]] + [c[1] for c in scat]
    return dict(eps=prob["eps"], h=prob["h"], fams=fams, cands=cands, **new), (emb, plus, minus, scat, state)


@pytest.mark.parametrize("N", [4, 10])
def test_scba_state_vs_oracle(hip, orc, N):
    prob = fx.make_problem(N=N, NE=48, Nk=3)
    s, _ = _scba_state(prob)
    ref = _oracle(orc, s)
    got = _gpu(hip, s)
    # a consistent state: r_D, r_A at rounding level, r_PSD and r_caus ~0
    assert ref["r_D"] < 1e-13 and got["r_D"] < 1e-13
    for k in ("r_A", "r_K"):
        assert abs(got[k] - ref[k]) <= 1e-10 * ref[k] + 1e-13, k
    for k in ("r_PSD", "r_caus"):
        assert abs(got[k] - ref[k]) <= 1e-12, k


def test_selfenergy_residual_vs_oracle(hip, orc):
    prob = fx.make_problem(N=6, NE=40, Nk=3)
    s, (emb, plus, minus, scat, state) = _scba_state(prob)
    old = [state["embR"], state["embL"], state["embG"]]
    ref = orc.selfenergy_residual(old, emb)
    got = hip.selfenergy_residual([hip.to_device(x) for x in old], [hip.to_device(x) for x in emb])
    assert ref > 0 and abs(got - ref) <= 1e-12 * ref
    dev = [hip.to_device(x) for x in emb]
    assert hip.selfenergy_residual(dev, dev) == 0.0
    bad = [x.copy() for x in emb]
    bad[2][0, 0, 0, 0] = np.inf
    assert hip.selfenergy_residual([hip.to_device(x) for x in old], [hip.to_device(x) for x in bad]) == np.inf
    assert orc.selfenergy_residual(old, bad) == np.inf


def test_deferred_matches_eager(hip, orc):
    import torch
    s = _random_state(NE=6, Nk=2, Nb=7, seed=21)
    g, h, green, fams, cands = _device(hip, s)
    eager = hip.residual_suite(g, h, green, fams, cands)
    sig = hip.selfenergy_residual(fams[0], fams[1])
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    g.ctx.defer(buf)
    try:
        d1 = hip.residual_suite(g, h, green, fams, cands)
        d2 = hip.selfenergy_residual(fams[0], fams[1])
    finally:
        g.ctx.defer(None)
    g.ctx.synchronize()
    host = buf.cpu().numpy()
    assert list(d1.read(host)) == [eager[k] for k in KEYS]
    assert d2.read(host) == sig


def test_mix_fields_residual_vs_oracle(hip, orc):
    """Fused mixing + self-energy residuals: mixed fields equal the plain
    mixing; each pair's ratio and the couple (sum) ratio equal the oracle's
    selfenergy residual of the same (old, candidate) sums."""
    import torch
    rng = np.random.default_rng(31)
    shape = (9, 3, 5, 5)
    olds = [_rand(rng, shape) for _ in range(5)]
    cands = [o + _rand(rng, shape, 0.1 * (k + 1)) for k, o in enumerate(olds)]
    dev_o = [hip.to_device(x) for x in olds]
    dev_o2 = [t.clone() for t in dev_o]
    dev_c = [hip.to_device(x) for x in cands]
    partner = [-1, 3, 4, 1, 2]  # couples (1,3) and (2,4)
    ratios = hip.mix_fields_residual(list(zip(dev_o, dev_c)), 0.3, partner)
    hip.mix_fields(list(zip(dev_o2, dev_c)), 0.3)
    for a, b in zip(dev_o, dev_o2):
        assert torch.equal(a, b)
    z = np.zeros(shape, np.complex128)
    single = lambda o, c: orc.selfenergy_residual([o, z, z], [c, z, z])  # one component: R slot
    refs = [single(o, c) for o, c in zip(olds, cands)]
    refs += [single(olds[1] + olds[3], cands[1] + cands[3]), single(olds[2] + olds[4], cands[2] + cands[4])]
    assert len(ratios) == 7
    for got, ref in zip(ratios, refs):
        assert abs(got - ref) <= 1e-12 * ref, (got, ref)
    bad = [t.clone() for t in dev_c]
    bad[4][0, 0, 0, 0] = float("nan")
    r = hip.mix_fields_residual(list(zip([t.clone() for t in dev_o2], bad)), 0.3, partner)
    assert r[4] == np.inf and r[6] == np.inf and r[0] < np.inf


def test_mix_fields_residual_totals(hip):
    """The mixing pass grouped by totals: mixed fields and ratios equal the
    ungrouped pass bitwise; each total equals the left-fold sum of its
    members' mixed fields bitwise (what the next solve's family sum loads);
    the argument checks reject a pair outside every total and a couple split
    over two totals."""
    import torch
    rng = np.random.default_rng(37)
    shape = (11, 3, 4, 4)
    olds = [_rand(rng, shape) for _ in range(7)]
    cands = [o + _rand(rng, shape, 0.1 * (k + 1)) for k, o in enumerate(olds)]
    dev_o = [hip.to_device(x) for x in olds]
    dev_o2 = [t.clone() for t in dev_o]
    dev_c = [hip.to_device(x) for x in cands]
    partner = [-1, 3, -1, 1, -1, 6, 5]  # couples (1,3) and (5,6)
    members = [[3, 1, 0], [6, 2, 5, 4]]  # summation order deliberately not ascending
    tots = [torch.empty_like(dev_o[0]) for _ in members]
    r1 = hip.mix_fields_residual(list(zip(dev_o, dev_c)), 0.4, partner, totals=list(zip(tots, members)))
    r2 = hip.mix_fields_residual(list(zip(dev_o2, dev_c)), 0.4, partner)
    assert r1 == r2
    for a, b in zip(dev_o, dev_o2):
        assert torch.equal(a, b)
    for t, m in zip(tots, members):
        ref = dev_o[m[0]].clone()
        for k in m[1:]:
            ref = ref + dev_o[k]
        assert torch.equal(t, ref)
    with pytest.raises(hip._lib.ArgumentError):
        hip.mix_fields_residual(list(zip(dev_o, dev_c)), 0.4, partner, totals=[(tots[0], [3, 1, 0])])
    with pytest.raises(hip._lib.ArgumentError):
        hip.mix_fields_residual(list(zip(dev_o, dev_c)), 0.4, partner,
                                totals=[(tots[0], [1, 0, 2]), (tots[1], [3, 4, 5, 6])])


@pytest.mark.parametrize("engine", ["vector", "matrix", "fused"])
def test_current_green_totals(hip, engine):
    """qclhip_current_green_totals: the Green functions equal those of
    qclhip_current_green bitwise, and the totals equal the family sums
    bitwise (kernel by-product in the FP64 vector path, a summation pass
    otherwise)."""
    import torch
    w = fx.syn.gf_workload(N=9, NE=40, Nk=3)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 9)
    g.set_engine(engine)
    inj = fx.syn.injection_family(w)
    fams = [fx.syn.diag_fields_device(inj), fx.syn.diag_fields_device(w)]
    fams.append(tuple(0.37 * x for x in fams[1]))  # a third physical family
    h = hip.h_to_device(w["h"])
    st1, i1 = hip.current_green_production(g, h, fams, 1.0)
    tot = tuple(torch.empty_like(st1.GR) for _ in range(3))
    st2, i2 = hip.current_green_production(g, h, fams, 1.0, totals=tot)
    for k in ("GR", "GL", "GG", "A"):
        assert torch.equal(getattr(st1, k), getattr(st2, k)), k
    assert i1 == i2
    for k in range(3):
        ref = torch.empty_like(tot[k])
        hip.sum_fields([f[k] for f in fams], ref)
        assert torch.equal(tot[k], ref), k


def _screened_state(hip, N, NE, Nk, indefinite):
    """A GF solve on two diagonal physical families (an injecting one and the
    workload's), with indefinite: a random Hermitian part added to i S< of
    the scattering family (so the raw products -i G<raw fail their screens),
    through current_green_production(raw_screen=...)."""
    import torch
    syn = fx.syn
    w = syn.gf_workload(N=N, NE=NE, Nk=Nk)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N)
    SR, SL, SG = syn.diag_fields_host(w)
    if indefinite:
        rng = np.random.default_rng(N)
        M = rng.normal(size=SL.shape) + 1j * rng.normal(size=SL.shape)
        M = 0.5 * (M + np.conj(np.swapaxes(M, 2, 3)))  # Hermitian blocks
        SL = np.asfortranarray(SL - 1j * 2e-4 * M)  # S< stays anti-Hermitian
    fams = [tuple(hip.to_device(x) for x in syn.diag_fields_host(syn.injection_family(w))),
            (hip.to_device(SR), hip.to_device(SL), hip.to_device(SG))]
    tot = tuple(torch.empty_like(fams[0][0]) for _ in range(3))
    scr = hip.screen_buffer(g)
    st, info = hip.current_green_production(g, hip.h_to_device(w["h"]), fams, 1.0, spectral=False, totals=tot,
                                            raw_screen=scr)
    target = 0.9 * info["number"]
    st, info = hip.current_green_production(g, hip.h_to_device(w["h"]), fams, target, state=st, spectral=False,
                                            totals=tot, raw_screen=scr)
    cands = [0.3 * fams[1][1], fams[0][1]]
    return g, hip.h_to_device(w["h"]), st, tot, cands, scr


@pytest.mark.parametrize("N,indefinite", [(5, False), (24, False), (24, True), (28, True), (32, False), (32, True)])
def test_prescreened_suite_equals_suite(hip, N, indefinite):
    """The suite and its state pass with the raw-product screens taken from
    the GF solve's Keldysh kernel equal the plain entries bit for bit, on
    states whose raw screens settle and on ones whose raw screens fail (those
    blocks get their eigenvalues from products formed again). At N = 28 and
    32 the prescreened pass runs the four-waves-per-block kernel, whose every
    per-block value must equal the one-wave kernel's."""
    g, h, st, tot, cands, scr = _screened_state(hip, N, NE=37, Nk=3, indefinite=indefinite)
    flags = scr.cpu().numpy()
    if indefinite:
        assert np.count_nonzero(flags & 0x20) > 0  # unsettled -i G<raw screens
    else:
        assert not np.any(flags)
    plain = hip.residual_suite(g, h, st, [tot], cands)
    pre = hip.residual_suite(g, h, st, [tot], cands, raw_screen=scr)
    assert plain == pre, (plain, pre)
    if indefinite:
        assert plain["r_PSD"] > 0.0
    plain = hip.residual_state(g, h, st, [tot])
    pre = hip.residual_state(g, h, st, [tot], raw_screen=scr)
    assert plain == pre, (plain, pre)


def test_screened_scba_iteration_equals_plain(hip):
    """SCBA iterations whose GF solve screens the raw products for the suite
    (the default on the FP64 vector engine) equal iterations without, bit for
    bit, in every residual and field."""
    import torch
    from qclhip import scba
    from test_gpu_adjoint import _device_problem
    prob = fx.make_problem(N=10, NE=48, Nk=3)
    full, state, _ = fx.oracle_state(prob)
    rows, fields = [], []
    for use in (True, False):
        g, p, s = _device_problem(hip, prob, full, state)
        ws = scba.ScbaWorkspace(p)
        if not use:
            ws.raw_screen = None
        r = [scba.scba_iteration(p, s, ws) for _ in range(3)]
        rows.append(r)
        fields.append([t.clone() for t in (s.green.GR, s.green.GL, s.green.GG)])
    for a, b in zip(rows[0], rows[1]):
        for k in ("r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma", "lam", "Jbar"):
            assert a[k] == b[k], k
    for x, y in zip(*fields):
        assert torch.equal(x, y)


def test_screened_gf_solve_limits(hip):
    """Raw-product screens come from the FP64 vector Keldysh kernels up to
    Nb = 32 only; elsewhere the screened GF solve is refused (and the SCBA
    iteration leaves the raw products to the residual suite)."""
    from qclhip import _lib, scba
    with pytest.raises(_lib.UnsupportedError):
        _screened_state(hip, 40, NE=37, Nk=3, indefinite=False)
    w = fx.syn.gf_workload(N=6, NE=16, Nk=2)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 6)
    fams = [fx.syn.diag_fields_device(w)]
    for engine in ("matrix", "fused"):
        g.set_engine(engine)
        with pytest.raises(_lib.UnsupportedError):
            hip.current_green_production(g, hip.h_to_device(w["h"]), fams, 1.0, raw_screen=hip.screen_buffer(g))
        assert scba._screen_buffer(g, type("W", (), {"raw_screen": hip.screen_buffer(g)})()) is None
