"""Pins of the oracle's residual suite (residuals.jl:537-865) and of its
Hermitian eigenvalue routine by cases whose values the mathematics fixes:
numpy's eigvalsh, G = 0 (r_D = 1 exactly), exact inverses (r_D ~ eps),
blocks with a known negative eigenvalue (r_PSD closed form), a retarded
self-energy with a known positive Im eigenvalue (r_caus closed form),
constant-diagonal states (r_A = r_K = 1/2)."""
import numpy as np
import pytest

from oracle import oracle as orc

EPS = np.finfo(float).eps


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 24, 33, 48])
@pytest.mark.parametrize("scale", [1e-8, 1.0, 1e8])
def test_hermitian_eigvals_vs_numpy(n, scale):
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = scale * (X + X.conj().T)
    w = orc.hermitian_eigvals(H)
    ref = np.linalg.eigvalsh(H)
    assert np.max(np.abs(w - ref)) <= 64 * n * EPS * np.max(np.abs(ref))


def test_hermitian_eigvals_degenerate():
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)) + 1j * rng.normal(size=(6, 6)))
    lam = np.array([-1.0, -1.0, 0.0, 2.0, 2.0, 2.0])
    w = orc.hermitian_eigvals(Q @ np.diag(lam) @ Q.conj().T)
    assert np.max(np.abs(w - lam)) < 1e-14
    assert np.array_equal(orc.hermitian_eigvals(np.eye(5)), np.ones(5))


def _zero_state(NE=3, Nk=2, Nb=4):
    z = lambda: np.zeros((NE, Nk, Nb, Nb), np.complex128, order="F")
    return dict(eps=np.linspace(0, 1, NE), h=np.zeros((Nk, Nb, Nb), np.complex128, order="F"),
                GR=z(), GL=z(), GG=z(), A=z(), SR=z(), SL=z(), SG=z(), cand=z())


def _suite(s):
    return orc.residual_suite(s["eps"], s["h"], s["GR"], s["GL"], s["GG"], s["A"], s["SR"], s["SL"], s["SG"],
                              s["cand"])


def test_zero_green_function_gives_unit_dyson_residual():
    r = _suite(_zero_state())
    assert r["r_D"] == pytest.approx(1.0, rel=1e-15)  # ||-I|| / (0 + sqrt(Nb))
    assert r["r_PSD"] == 0.0 and r["r_caus"] == 0.0


def test_exact_inverse_has_eps_dyson_residual():
    rng = np.random.default_rng(5)
    s = _zero_state()
    NE, Nk, Nb = 3, 2, 4
    for e in range(NE):
        for m in range(Nk):
            S = 0.1 * (rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb)))
            s["SR"][e, m] = S
            D = s["eps"][e] * np.eye(Nb) - s["h"][m] - S
            s["GR"][e, m] = np.linalg.inv(D)
    assert _suite(s)["r_D"] < 1e-14


@pytest.mark.parametrize("field,factor", [("A", 1.0), ("GL", 1j), ("GG", -1j)])
def test_negative_eigenvalue_gives_closed_form_psd(field, factor):
    rng = np.random.default_rng(7)
    s = _zero_state()
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    lam = np.array([-0.3, 0.5, 1.0, 2.0])
    # f X must equal Q diag(lam) Q^dagger: f = 1 for A, -i for G<, i for G>
    s[field][1, 0] = factor * (Q @ np.diag(lam) @ Q.conj().T)
    budget = 64 * 4 * EPS * 2.0
    assert _suite(s)["r_PSD"] == pytest.approx((0.3 - budget) / 2.0, rel=1e-14)


def test_raw_keldysh_product_positivity():
    """The raw product G^R S< G^R^dagger is checked before the paired map."""
    s = _zero_state()
    s["GR"][0, 1] = np.eye(4)
    s["SL"][0, 1] = 1j * np.diag([-0.25, 1.0, 1.0, 1.0])  # -i S< has eigenvalue -0.25
    r = _suite(s)
    # Gamma = i(S> - S<) = diag(-0.25, 1, 1, 1) is itself indefinite: same ratio
    assert r["r_PSD"] == pytest.approx((0.25 - 64 * 4 * EPS) / 1.0, rel=1e-14)


def test_causality_closed_form():
    s = _zero_state()
    s["SR"][2, 1] = 1j * np.diag([0.2, -0.5, -1.0, 0.0])
    r = _suite(s)
    assert r["r_caus"] == pytest.approx(0.2 / (1.0 + 1e-14), rel=1e-14)
    s["SR"][2, 1] = -1j * np.eye(4)
    assert _suite(s)["r_caus"] == 0.0


def test_spectral_and_keldysh_closed_form():
    s = _zero_state()
    NE, Nk, Nb = 3, 2, 4
    I = np.eye(Nb)
    for e in range(NE):
        for m in range(Nk):
            s["GR"][e, m] = I
            s["A"][e, m] = 2 * I
            s["SG"][e, m] = -1.5j * I     # Gamma = i(S> - S<) = 3 I
            s["SL"][e, m] = 1.5j * I
            s["GL"][e, m] = 2j * I
            s["cand"][e, m] = 1j * I      # G cand G^dagger = i I
    r = _suite(s)
    assert r["r_A"] == pytest.approx(0.5, rel=1e-14)
    assert r["r_K"] == pytest.approx(0.5, rel=1e-14)


def test_selfenergy_residual_closed_form():
    rng = np.random.default_rng(9)
    old = [rng.normal(size=(5, 2, 3, 3)) + 1j * rng.normal(size=(5, 2, 3, 3)) for _ in range(3)]
    cand = [old[0], 2 * old[1], old[2] + 0.0]
    assert orc.selfenergy_residual(old, cand) == pytest.approx(0.5, rel=1e-15)
    assert orc.selfenergy_residual(old, old) == 0.0


def test_nonfinite_entries_give_inf():
    """The reference's canonical reductions (residuals.jl:495-512, 606-610):
    a non-finite block makes the affected residuals Inf."""
    s = _zero_state()
    s["GR"][:, :] = np.eye(4)
    s["GR"][1, 1, 2, 0] = np.nan
    r = _suite(s)
    assert r["r_D"] == np.inf and r["r_A"] == np.inf and r["r_PSD"] == np.inf
    s = _zero_state()
    s["SR"][0, 0, 0, 0] = np.inf
    r = _suite(s)
    assert r["r_caus"] == np.inf and r["r_D"] == np.inf
    old = [np.zeros((2, 1, 2, 2), np.complex128) for _ in range(3)]
    cand = [x.copy() for x in old]
    cand[1][1, 0, 1, 1] = np.nan
    assert orc.selfenergy_residual(old, cand) == np.inf
