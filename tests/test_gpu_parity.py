"""Parity of the HIP operators with the CPU oracle (FP64, <= 1e-12 relative).

Every call goes through libqclhip.so's C-ABI. Sizes are small enough for the
oracle to finish in seconds; the full-size batch is checked on a sample of
blocks (blocks are independent) and through size-independent identities.
"""
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


def _grid(hip, eps, wE, wk, Nb):
    return hip.Grid(eps, wE, wk, Nb)


def _random_fields(rng, NE, Nk, Nb, n, scale=0.01):
    return [np.asfortranarray(scale * (rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb))))
            for _ in range(n)]


def test_scalar_lorentzian_gpu(hip, orc):
    eps = np.linspace(-1.0, 1.0, 2001)
    g = _grid(hip, eps, np.full(2001, eps[1] - eps[0]), np.ones(1), 1)
    h = hip.h_to_device(np.array([[[0.17]]]))
    sigR = hip.to_device(np.full((2001, 1, 1, 1), -0.04j))
    GR, scale = hip.retarded_green_production(g, h, sigR)
    A = hip.to_host(hip.spectral_function(g, GR))[:, 0, 0, 0]
    err = float(rel(A.real, 0.08 / ((eps - 0.17) ** 2 + 0.04 ** 2)))
    assert err < 2e-13


@pytest.mark.parametrize("NE", [24, 21])
@pytest.mark.parametrize("Nb", [1, 2, 3, 5, 8, 12, 16, 20, 24, 31, 32, 40, 48])
def test_dyson_keldysh_dense_sigma(hip, orc, Nb, NE):
    """Dense (non-diagonal) self-energies at every tile size of the kernel;
    NE = 21 leaves a partial last tile of 8 (Dyson) and 4 (Keldysh) energies."""
    rng = np.random.default_rng(100 + Nb + NE)
    Nk = 3
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=Nb)
    SR, SL, SG = fx.syn.diag_fields_host(w)
    dR, dL, dG = _random_fields(rng, NE, Nk, Nb, 3, scale=0.01)
    SR = np.asfortranarray(SR + dR)
    SL = np.asfortranarray(SL + dL)
    SG = np.asfortranarray(SG + dG)
    GR_o, scale_o = orc.retarded_green(w["eps"], w["h"], SR)
    GL_o = orc.keldysh(GR_o, SL)
    GG_o = orc.keldysh(GR_o, SG)
    g = _grid(hip, w["eps"], w["wE"], w["wk"], Nb)
    import torch
    traces = torch.empty(2 * NE * Nk * 2, dtype=torch.float64, device="cuda")
    GR, GL, GG = hip.dyson_keldysh(g, hip.h_to_device(w["h"]), hip.to_device(SR), hip.to_device(SL),
                                   hip.to_device(SG), traces=traces, check_status=True)
    err = float(rel(hip.to_host(GR), GR_o))
    assert err < TOL
    err = float(rel(hip.to_host(GL), GL_o))
    assert err < TOL
    err = float(rel(hip.to_host(GG), GG_o))
    assert err < TOL
    tr = traces.cpu().numpy().view(np.complex128).reshape(2, Nk, NE)
    err = float(rel(tr[0].T, np.einsum("emaa->em", GL_o)))
    assert err < TOL
    err = float(rel(tr[1].T, np.einsum("emaa->em", GG_o)))
    assert err < TOL


@pytest.mark.parametrize("Nb", [5, 24])
def test_dyson_keldysh_strided_equals_contiguous(hip, Nb):
    """The launch pair on fields with padded planes gives the contiguous
    results bit for bit and leaves the padding untouched; NE = 21 leaves
    partial last tiles."""
    import torch
    rng = np.random.default_rng(40 + Nb)
    NE, Nk = 21, 3
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=Nb)
    SR, SL, SG = fx.syn.diag_fields_host(w)
    dR, dL, dG = _random_fields(rng, NE, Nk, Nb, 3, scale=0.01)
    sig = [hip.to_device(np.asfortranarray(x + d)) for x, d in ((SR, dR), (SL, dL), (SG, dG))]
    g = _grid(hip, w["eps"], w["wE"], w["wk"], Nb)
    h = hip.h_to_device(w["h"])
    GR, GL, GG = hip.dyson_keldysh(g, h, *sig, check_status=True)
    plane = NE * Nk + 37
    fields = [hip.strided_field(g, plane) for _ in range(6)]
    for f, x in zip(fields[:3], sig):
        f.copy_(x)
    for f in fields[3:]:
        f._base[:, NE * Nk:] = 1.5  # padding the kernels must not touch
    hip.dyson_keldysh_strided(g, plane, h, *fields)
    g.ctx.synchronize()
    for got, ref in zip(fields[3:], (GR, GL, GG)):
        assert torch.equal(got, ref)
        assert bool((got._base[:, NE * Nk:] == 1.5).all())
    import ctypes
    from qclhip import ArgumentError
    with pytest.raises(ArgumentError):
        hip.strided_field(g, NE * Nk - 1)
    with pytest.raises(ArgumentError):  # the library's own check
        g.ctx.call("qclhip_dyson_keldysh_strided", g.handle, ctypes.c_int64(NE * Nk - 1), *[None] * 4,
                   ctypes.c_double(0.0), *[None] * 5)


def test_retarded_green_scale_and_seed_eta(hip, orc):
    w = fx.syn.gf_workload(N=6, NE=32, Nk=2, seed=4)
    g = _grid(hip, w["eps"], w["wE"], w["wk"], 6)
    GR, scale = hip.retarded_green_production(g, hip.h_to_device(w["h"]), None, eta=0.05)
    GR_o, scale_o = orc.retarded_green(w["eps"], w["h"], None, eta=0.05)
    err = float(rel(hip.to_host(GR), GR_o))
    assert err < TOL
    err = float(rel(scale.cpu().numpy().T, scale_o))
    assert err < 1e-15


def test_singular_block_raises(hip, orc):
    from qclhip import SingularException
    eps = np.array([0.5, 1.0, 2.0])
    g = _grid(hip, eps, np.ones(3), np.ones(2), 2)
    h = np.zeros((2, 2, 2), np.complex128)
    h[1] = np.eye(2)  # D = (e - 1) I is zero only at (e=1, m=1): one-based block 1 + 3*1 + 1 = 5
    with pytest.raises(SingularException) as err:
        hip.retarded_green_production(g, hip.h_to_device(h), None)
    assert err.value.block == 5
    # a rank-deficient (non-zero) block: exact zero pivot
    h2 = np.zeros((2, 2, 2), np.complex128)
    h2[0] = 0.5 * np.eye(2) - np.ones((2, 2))  # D at (e=0, m=0) is [[1,1],[1,1]]
    with pytest.raises(SingularException) as err:
        hip.retarded_green_production(g, hip.h_to_device(h2), None)
    assert err.value.block == 1


def test_current_green_vs_oracle(hip, orc):
    prob = fx.make_problem(N=5, NE=48, Nk=3)
    full, state, green = fx.oracle_state(prob)
    fams_o = ([state["embR"]] + state["scatR"], [state["embL"]] + state["scatL"], [state["embG"]] + state["scatG"])
    ref = orc.current_green(prob["eps"], prob["h"], *fams_o, prob["wE"], prob["wk"], 2, full["target"])
    g = _grid(hip, prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    fams = [tuple(hip.to_device(x) for x in (R, L, G)) for R, L, G in zip(*fams_o)]
    st, info = hip.current_green_production(g, hip.h_to_device(prob["h"]), fams, full["target"])
    for k in ("GR", "GL", "GG", "A"):
        err = float(rel(hip.to_host(getattr(st, k)), ref[k]))
        assert err < TOL, k
    for k in ("lam", "number", "holes"):
        assert abs(info[k] - ref[k]) <= TOL * abs(ref[k]), k


@pytest.mark.parametrize("NE", [48, 45])
@pytest.mark.parametrize("Nb", [5, 20, 31, 48])
def test_current_green_spectral_shapes(hip, orc, Nb, NE):
    """current_green's A (stored by the Keldysh kernel from its G^R tiles)
    and the other outputs against the oracle on both Keldysh code paths
    (register prefetch for N <= 24, none above) and with partial last tiles
    (NE = 45); two dense families summed on load."""
    rng = np.random.default_rng(700 + Nb + NE)
    Nk = 2
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=Nb + 3)
    fams_h = []
    for src in (fx.syn.injection_family(w), w):
        R, L, G = fx.syn.diag_fields_host(src)
        dR, dL, dG = _random_fields(rng, NE, Nk, Nb, 3, scale=0.002)
        fams_h.append((np.asfortranarray(R + dR), np.asfortranarray(L + dL - np.conj(dL.transpose(0, 1, 3, 2))),
                       np.asfortranarray(G + dG - np.conj(dG.transpose(0, 1, 3, 2)))))
    probe = orc.current_green(w["eps"], w["h"], *zip(*fams_h), w["wE"], w["wk"], 2, 1.0)
    for target in (0.9 * probe["number"], 1.05 * probe["number"]):  # both branches of the paired map
        ref = orc.current_green(w["eps"], w["h"], *zip(*fams_h), w["wE"], w["wk"], 2, target)
        g = _grid(hip, w["eps"], w["wE"], w["wk"], Nb)
        fams = [tuple(hip.to_device(x) for x in f) for f in fams_h]
        st, info = hip.current_green_production(g, hip.h_to_device(w["h"]), fams, target)
        for k in ("GR", "GL", "GG", "A"):
            err = float(rel(hip.to_host(getattr(st, k)), ref[k]))
            assert err < TOL, (k, err)
        A_sep = hip.to_host(hip.spectral_function(g, st.GR))
        assert np.array_equal(hip.to_host(st.A), A_sep)  # the arithmetic of the spectral kernel
        for k in ("lam", "number", "holes"):
            assert abs(info[k] - ref[k]) <= TOL * abs(ref[k]), k


@pytest.mark.parametrize("Nb", [5, 31])
def test_current_green_scalar_lesser(hip, orc, Nb):
    """occupation_normalization = :scalar_lesser (production_state.jl:54-61):
    G< = G<raw/lambda, G> = G^R - G^A + G<, against the oracle's composition;
    lambda <= 0 raises DomainError."""
    w = fx.syn.gf_workload(N=Nb, NE=45, Nk=2, seed=Nb + 11)
    fams_h = [fx.syn.diag_fields_host(fx.syn.injection_family(w)), fx.syn.diag_fields_host(w)]
    probe = orc.current_green(w["eps"], w["h"], *zip(*fams_h), w["wE"], w["wk"], 2, 1.0)
    target = 0.8 * probe["number"]
    ref = orc.current_green_scalar_lesser(w["eps"], w["h"], *zip(*fams_h), w["wE"], w["wk"], 2, target)
    g = _grid(hip, w["eps"], w["wE"], w["wk"], Nb)
    fams = [tuple(hip.to_device(x) for x in f) for f in fams_h]
    h = hip.h_to_device(w["h"])
    st, info = hip.current_green_production(g, h, fams, target, occupation_normalization="scalar_lesser")
    for k in ("GR", "GL", "GG", "A"):
        err = float(rel(hip.to_host(getattr(st, k)), ref[k]))
        assert err < TOL, (k, err)
    assert abs(info["lam"] - ref["lam"]) <= TOL * ref["lam"]
    assert abs(info["number"] - ref["number"]) <= TOL * abs(ref["number"])
    from qclhip import DomainError, ArgumentError
    neg = [tuple(hip.to_device(-x) if k else hip.to_device(x) for k, x in enumerate(f)) for f in fams_h]
    with pytest.raises(DomainError):  # S< and S> negated: N[G<raw] < 0
        hip.current_green_production(g, h, neg, target, occupation_normalization="scalar_lesser")
    with pytest.raises(ArgumentError):
        hip.current_green_production(g, h, fams, target, occupation_normalization="scalar")


def test_number_functional_gpu(hip, orc):
    rng = np.random.default_rng(3)
    NE, Nk, Nb = 40, 3, 4
    eps = np.linspace(0, 1, NE)
    wE, wk = rng.uniform(0.5, 1, NE), rng.uniform(0.5, 1, Nk)
    G = np.asfortranarray(1j * rng.normal(size=(NE, Nk, Nb, Nb)))
    g = _grid(hip, eps, wE, wk, Nb)
    v = hip.number_functional(g, hip.to_device(G))
    assert abs(v - orc.number_functional(G, wE, wk, 2)) < 1e-13 * abs(v)
    from qclhip import DomainError
    with pytest.raises(DomainError):
        hip.number_functional(g, hip.to_device(np.asfortranarray(G + 1.0)))


@pytest.mark.parametrize("delta", [0.37, -0.37, 0.0, 2.5])
def test_shift_gpu(hip, orc, delta):
    rng = np.random.default_rng(8)
    NE, Nk, Nb = 50, 2, 3
    eps = np.linspace(-1, 1, NE)
    X = _random_fields(rng, NE, Nk, Nb, 1)[0]
    g = _grid(hip, eps, np.ones(NE), np.ones(Nk), Nb)
    Y = hip.to_host(hip.apply_energy_shift(hip.EnergyShiftPlan(g, delta), hip.to_device(X)))
    err = rel(Y, orc.apply_shift(orc.shift_plan(eps, delta), X))
    assert err < 1e-15


@pytest.mark.parametrize("kind", ["dense", "momentum_independent", "separable_ifr", "low_rank"])
def test_contractions_gpu(hip, orc, kind):
    rng = np.random.default_rng(21)
    prob = fx.make_problem(N=4, NE=40, Nk=3, lo_kind="dense")
    m = {x["kind"]: x for x in prob["mechanisms"]}
    if kind == "low_rank":
        l, r = fx.syn.low_rank_factors(rng, 4, 3, r=3)
        m["low_rank"] = dict(kind="low_rank", a=l, b=r, c=None, rank=3)
    mech = m[kind]
    G = _random_fields(rng, 40, 3, 4, 1, scale=1.0)[0]
    ko = orc.make_kernel(kind, mech["a"], mech["b"], mech["c"], rank=mech["rank"])
    g = _grid(hip, prob["eps"], prob["wE"], prob["wk"], 4)
    kd = hip.ProductionKernel(g, kind, mech["a"], mech["b"], mech["c"], rank=mech["rank"])
    S = hip.production_static_contraction(kd, hip.to_device(G), qK=0.7, alpha=1.3)
    ref = orc.contract(ko, G, prob["wk"], 0.7, alpha=1.3)
    err = float(rel(hip.to_host(S), ref))
    assert err < TOL
    S2 = hip.production_static_contraction(kd, hip.to_device(G), qK=0.7, alpha=0.5, beta=1.0, S=S)
    ref2 = orc.contract(ko, G, prob["wk"], 0.7, alpha=0.5, beta=1.0, S=ref.copy(order="F"))
    err = float(rel(hip.to_host(S2), ref2))
    assert err < TOL


@pytest.mark.parametrize("N,J", [(32, 6), (24, 16), (8, 16)])
def test_separable_large_factor_gpu(hip, orc, N, J):
    """Separable (IFR) kernels whose J x N^2 factor exceeds 64 KiB of LDS
    (N=32, six interfaces: 96 KiB; N=24, J=16: 144 KiB) run on 1024-thread
    workgroups; a small one on 256 threads; all against the oracle, with beta."""
    rng = np.random.default_rng(60 + N + J)
    NE, Nk = 20, 4
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=N)
    left, right, radial = fx.syn.separable_factors(rng, N, Nk, J=J, scale=0.05)
    G = _random_fields(rng, NE, Nk, N, 1, scale=1.0)[0]
    S0 = _random_fields(rng, NE, Nk, N, 1, scale=1.0)[0]
    ko = orc.make_kernel("separable_ifr", left, right, radial, rank=J)
    g = _grid(hip, w["eps"], w["wE"], w["wk"], N)
    kd = hip.ProductionKernel(g, "separable_ifr", left, right, radial, rank=J)
    S = hip.production_static_contraction(kd, hip.to_device(G), qK=0.7, alpha=1.1, beta=0.5, S=hip.to_device(S0))
    ref = orc.contract(ko, G, w["wk"], 0.7, alpha=1.1, beta=0.5, S=S0.copy(order="F"))
    assert float(rel(hip.to_host(S), ref)) < TOL


@pytest.mark.parametrize("lo_kind", ["dense", "low_rank"])
def test_lo_contraction_gpu(hip, orc, lo_kind):
    """Both LO paths: one contraction per component then the shifts (dense),
    and the shifts on the reduced coefficients before the expansion (low rank)."""
    rng = np.random.default_rng(5)
    prob = fx.make_problem(N=4, NE=64, Nk=3, lo_kind=lo_kind)
    mech = prob["mechanisms"][2]
    GL, GG = _random_fields(rng, 64, 3, 4, 2, scale=1.0)
    eps = prob["eps"]
    ko = orc.make_kernel(lo_kind, mech["a"], mech["b"], mech["c"], rank=mech["rank"])
    ref = orc.lo_contraction(ko, GL, GG, prob["wk"], 0.9, prob["NLO"],
                             orc.shift_plan(eps, prob["hwLO"]), orc.shift_plan(eps, -prob["hwLO"]))
    g = _grid(hip, eps, prob["wE"], prob["wk"], 4)
    kd = hip.ProductionKernel(g, lo_kind, mech["a"], mech["b"], mech["c"], rank=mech["rank"])
    SL, SG = hip.production_lo_contraction(kd, hip.to_device(GL), hip.to_device(GG), 0.9, prob["NLO"],
                                           hip.EnergyShiftPlan(g, prob["hwLO"]),
                                           hip.EnergyShiftPlan(g, -prob["hwLO"]))
    err = float(rel(hip.to_host(SL), ref[0]))
    assert err < TOL
    err = float(rel(hip.to_host(SG), ref[1]))
    assert err < TOL


@pytest.mark.parametrize("NE", [2, 3, 37, 64, 100, 513, 2048, 4096])
def test_fft_hilbert_gpu(hip, orc, NE):
    rng = np.random.default_rng(NE)
    Nk, Nb = 2, 2
    eps = np.linspace(-1.3, 0.9, NE)
    wE = np.full(NE, eps[1] - eps[0])
    wE[[0, -1]] /= 2
    Gam = _random_fields(rng, NE, Nk, Nb, 1, scale=1.0)[0]
    ref, rr = orc.fft_hilbert(Gam, eps, wE)
    g = _grid(hip, eps, wE, np.ones(Nk), Nb)
    plan = hip.FFTHilbertPlan(g, verify_roundoff=True)
    lam, r = hip.production_fft_hilbert_transform(plan, hip.to_device(Gam), return_roundoff=True)
    err = float(rel(hip.to_host(lam), ref))
    assert err < TOL
    assert r < 1e-12 and abs(r - rr) < 1e-13
    if NE <= 100:
        err = float(rel(hip.to_host(lam), orc.direct_hilbert(Gam, eps, wE)))
        assert err < 2e-13
    SL, SG = _random_fields(rng, NE, Nk, Nb, 2, scale=1.0)
    SRo, _, ro = orc.retarded_self_energy_fft(SL, SG, eps, wE)
    SR, _, r2 = hip.retarded_self_energy_fft(plan, hip.to_device(SL), hip.to_device(SG), return_roundoff=True)
    err = float(rel(hip.to_host(SR), SRo))
    assert err < TOL
    assert abs(r2 - ro) < 1e-13


@pytest.mark.parametrize("Nb", [5, 24, 32, 40, 48])
def test_embedding_gpu(hip, orc, Nb):
    """Every tile size of k_sandwich, including the large blocks whose rows
    are split over several waves (N > 24); NE = 61 leaves partial tiles."""
    rng = np.random.default_rng(12 + Nb)
    NE, Nk = 61, 2
    eps = np.linspace(-1, 2, NE)
    X = _random_fields(rng, NE, Nk, Nb, 1, scale=1.0)[0]
    T = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
    g = _grid(hip, eps, np.ones(NE), np.ones(Nk), Nb)
    for delta in (0.5, -0.5):
        ref = orc.embedding_component(T, X, orc.shift_plan(eps, delta))
        S = hip.embedding_component_production(g, hip.matrix_to_device(T), hip.to_device(X),
                                               hip.EnergyShiftPlan(g, delta))
        err = float(rel(hip.to_host(S), ref))
        assert err < TOL


def test_observables_gpu(hip, orc):
    rng = np.random.default_rng(13)
    NE, Nk, Nb = 30, 3, 4
    eps = np.linspace(-1.5, 1.0, NE)
    wE, wk = np.full(NE, eps[1] - eps[0]), rng.uniform(0.1, 1, Nk)
    SL, SG, GL, GG = _random_fields(rng, NE, Nk, Nb, 4, scale=1.0)
    g = _grid(hip, eps, wE, wk, Nb)
    J = hip.boundary_flux_bar(g, *(hip.to_device(x) for x in (SL, SG, GL, GG)))
    Jo = orc.boundary_flux_bar(eps, wE, wk, 2, SL, SG, GL, GG)
    assert abs(J - Jo) < TOL * max(1.0, abs(Jo))
    p = hip.state_populations(g, hip.to_device(GL), 1.7)
    err = float(rel(p, orc.state_populations(wE, wk, 2, 1.7, GL)))
    assert err < TOL


def test_seed_gpu(hip, orc):
    w = fx.syn.gf_workload(N=4, NE=128, Nk=3, seed=6)
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], None, eta=0.03)
    q_o = orc.seed_weights(GR_o, w["wE"], w["wk"], 2)
    target = 0.4 * q_o.sum()
    mu_o = orc.seed_mu(w["eps"], q_o, w["kbt"], target)
    GL_o = orc.seed_lesser(GR_o, w["eps"], mu_o, w["kbt"])
    g = _grid(hip, w["eps"], w["wE"], w["wk"], 4)
    st, mu = hip.seed_green_production(g, hip.h_to_device(w["h"]), 0.03, w["kbt"], target)
    assert abs(mu - mu_o) < 1e-12 * max(1, abs(mu_o))
    err = float(rel(hip.to_host(st.GL), GL_o))
    assert err < TOL
    assert abs(hip.number_functional(g, st.GL) - target) < 1e-12 * target


@pytest.mark.parametrize("lo_kind", ["dense", "low_rank"])
def test_scba_iterations_gpu(hip, orc, lo_kind):
    """Three full production iterations, device driver vs oracle driver."""
    from qclhip import scba
    prob = fx.make_problem(N=4, NE=64, Nk=3, lo_kind=lo_kind)
    full, state, green = fx.oracle_state(prob)
    g = _grid(hip, prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
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
    runner = orc.ScbaOracle(full, state)
    ws = scba.ScbaWorkspace(p)
    for it in range(3):
        ro = runner.iterate(measure=True)
        rd = scba.scba_iteration(p, s, ws)
        assert abs(rd["lam"] - ro["lam"]) < TOL * abs(ro["lam"])
        assert abs(rd["Jbar"] - ro["Jbar"]) < TOL * max(abs(ro["Jbar"]), 1e-300) + 1e-14
        err = float(rel(rd["populations"], ro["populations"]))
        assert err < TOL
        assert rd["roundoff"] < 1e-12
        # the fixed-point measurement (production_iteration.jl:102-166)
        assert rd["r_D"] < 1e-13 and ro["r_D"] < 1e-13
        for k in ("r_A", "r_K", "r_Sigma"):
            assert abs(rd[k] - ro[k]) <= 1e-9 * ro[k] + 1e-13, (it, k, rd[k], ro[k])
        for k in ("r_PSD", "r_caus", "r_lambda"):
            assert abs(rd[k] - ro[k]) <= 1e-12, (it, k, rd[k], ro[k])
    for k in ("GR", "GL", "GG"):
        err = float(rel(hip.to_host(getattr(s.green, k)), state[k]))
        assert err < TOL, k
    for i, m in enumerate(prob["mechanisms"]):
        err = float(rel(hip.to_host(s.scattering[m["name"]].R), state["scatR"][i]))
        assert err < TOL
        err = float(rel(hip.to_host(s.scattering[m["name"]].L), state["scatL"][i]))
        assert err < TOL
    err = float(rel(hip.to_host(s.embedding().R), state["embR"]))
    assert err < TOL


def test_gf_batch_full_size_sampled(hip, orc):
    """The synthetic GF-solve workload at full size (N=24, NE=2048, Nk=64):
    sampled blocks equal the oracle, and N[G<] equals the target."""
    import torch
    w = fx.syn.gf_workload()
    N, NE, Nk = w["N"], w["NE"], w["Nk"]
    g = _grid(hip, w["eps"], w["wE"], w["wk"], N)
    SR, SL, SG = fx.syn.diag_fields_device(w)
    emb = [torch.zeros_like(SR) for _ in range(3)]
    target = 1.0
    st, info = hip.current_green_production(g, hip.h_to_device(w["h"]), [tuple(emb), (SR, SL, SG)], target)
    assert abs(hip.number_functional(g, st.GL) - target) < 1e-12
    rng = np.random.default_rng(0)
    es = np.sort(rng.choice(NE, 24, replace=False))
    ms = np.sort(rng.choice(Nk, 4, replace=False))
    sub = dict(w)
    SRh = np.zeros((len(es), len(ms), N, N), np.complex128, order="F")
    SLh = np.zeros_like(SRh)
    SGh = np.zeros_like(SRh)
    for a in range(N):
        gam = w["gamma"][np.ix_(es, ms, [a])][:, :, 0]
        f = w["f"][es][:, None]
        SRh[:, :, a, a] = -0.5j * gam
        SLh[:, :, a, a] = 1j * f * gam
        SGh[:, :, a, a] = -1j * (1 - f) * gam
    GR_o, _ = orc.retarded_green(w["eps"][es], np.asfortranarray(w["h"][ms]), SRh)
    GL_o, GG_o = orc.keldysh(GR_o, SLh), orc.keldysh(GR_o, SGh)
    orc.normalize_pair(GL_o, GG_o, info["a"], info["c"])
    GRd = st.GR.cpu().numpy()  # (b, a, m, e)
    GLd = st.GL.cpu().numpy()
    for i, e in enumerate(es):
        for j, m in enumerate(ms):
            err = float(rel(GRd[:, :, m, e].T, GR_o[i, j]))
            assert err < TOL
            err = float(rel(GLd[:, :, m, e].T, GL_o[i, j]))
            assert err < TOL


def test_current_equals_landauer_transmission_gpu(hip):
    """The device current functional on the rotated three-site chain equals
    the closed-form Landauer current (physics/cavity_current_and_spectral_
    oracles.jl:5-36), independently of the oracle: Dyson, both Keldysh
    products and the boundary flux all run on the device."""
    from test_oracle_anchors import landauer_chain
    for fLR, bound in ((None, None), ((0.4, 0.4), 3e-14)):
        c = landauer_chain(fLR=fLR)
        g = _grid(hip, c["eps"], c["wE"], c["wk"], 3)
        GR, GL, GG = hip.dyson_keldysh(g, hip.h_to_device(c["h"]), *(hip.to_device(c[k]) for k in ("SR", "SL", "SG")))
        J = hip.boundary_flux_bar(g, hip.to_device(c["plusL"]), hip.to_device(c["plusG"]), GL, GG)
        if bound is None:
            assert abs(J - c["J"]) <= 2e-13 * abs(c["J"]), (J, c["J"])
        else:
            assert abs(J) < bound


@pytest.mark.parametrize("n", [481, 1921])
def test_hilbert_sign_and_scale_dawson_gpu(hip, n):
    """The device FFT Hilbert transform of Gaussian columns equals Dawson's
    integral plus the trapezoid rule's omitted-node term (closed form, no
    oracle), with the 2L round-off witness below 1e-12."""
    from test_oracle_anchors import dawson_columns
    x, w, G, exact = dawson_columns(n)
    ncol = G.shape[2]
    G = np.asfortranarray(G.reshape(n, ncol, 1, 1, order="F"))
    exact = exact.reshape(n, ncol, 1, 1, order="F")
    g = _grid(hip, x, w, np.ones(ncol), 1)
    plan = hip.FFTHilbertPlan(g, ncol, verify_roundoff=True)
    L, r = hip.production_fft_hilbert_transform(plan, hip.to_device(G), return_roundoff=True)
    scale = np.max(np.abs(exact))
    assert np.max(np.abs(hip.to_host(L) - exact)) < 1e-13 * scale
    assert r < 1e-12


@pytest.mark.parametrize("N,Nk,NE", [(5, 7, 70), (3, 2, 200), (8, 3, 33)])
def test_matrix_core_gemm_dense_odd_sizes(hip, orc, N, Nk, NE):
    """The dense contraction on the FP64 matrix-core GEMM at sizes that leave
    partial 64 x 64 tiles and partial 16-deep stages (Nc = Nk N^2 = 175, 18,
    192), with beta, against the oracle; the device-built kernel equals the
    host-built one bit for bit."""
    import torch
    rng = np.random.default_rng(80 + N + Nk)
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=N)
    Nc = Nk * N * N
    F = np.asfortranarray(rng.normal(size=(Nc, Nc)) + 1j * rng.normal(size=(Nc, Nc))) * 0.1
    G = _random_fields(rng, NE, Nk, N, 1, scale=1.0)[0]
    S0 = _random_fields(rng, NE, Nk, N, 1, scale=1.0)[0]
    g = _grid(hip, w["eps"], w["wE"], w["wk"], N)
    kd = hip.ProductionKernel(g, "dense", F)
    kdev = hip.ProductionKernel.dense_from_device(g, torch.from_numpy(np.ascontiguousarray(F.T)).cuda())
    ko = orc.make_kernel("dense", F)
    S = hip.production_static_contraction(kd, hip.to_device(G), qK=0.7, alpha=1.2, beta=0.4, S=hip.to_device(S0))
    S2 = hip.production_static_contraction(kdev, hip.to_device(G), qK=0.7, alpha=1.2, beta=0.4, S=hip.to_device(S0))
    ref = orc.contract(ko, G, w["wk"], 0.7, alpha=1.2, beta=0.4, S=S0.copy(order="F"))
    assert float(rel(hip.to_host(S), ref)) < TOL
    assert torch.equal(S, S2)


def test_matrix_core_gemm_low_rank_wide(hip, orc):
    """A low-rank kernel of rank 20 (above the streaming kernels' 16) runs
    as two matrix-core GEMMs (NE x r x Nc, then NE x Nc x r)."""
    rng = np.random.default_rng(91)
    N, Nk, NE, r = 4, 5, 90, 20
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=3)
    l, rr = fx.syn.low_rank_factors(rng, N, Nk, r=r, scale=0.05)
    G = _random_fields(rng, NE, Nk, N, 1, scale=1.0)[0]
    g = _grid(hip, w["eps"], w["wE"], w["wk"], N)
    kd = hip.ProductionKernel(g, "low_rank", l, rr, rank=r)
    S = hip.production_static_contraction(kd, hip.to_device(G), qK=0.9)
    ref = orc.contract(orc.make_kernel("low_rank", l, rr, rank=r), G, w["wk"], 0.9)
    assert float(rel(hip.to_host(S), ref)) < TOL
