"""Pins of the oracle's Anderson mixing (production_mixing.jl:36-317): the
coefficient solve against the constrained least-squares closed form, the
guard codes on constructed blocks (each failure mode, first-failure order),
and the mixed SCBA iteration (history, reset, convergence against linear
mixing)."""
import numpy as np
import pytest

import scba_fixture as fx
from oracle import oracle as orc


def _residual_gram(rng, count, n=40):
    R = rng.normal(size=(count, n))
    return R @ R.T, R


def test_coefficients_single_and_kkt():
    assert np.array_equal(orc.anderson_coefficients(np.array([[3.0]]), 1e-12), [1.0])
    rng = np.random.default_rng(1)
    G, R = _residual_gram(rng, 4)
    c = orc.anderson_coefficients(G, 0.0)
    # min ||sum c_i R_i|| subject to sum c_i = 1: c = G^-1 1 / (1^T G^-1 1)
    x = np.linalg.solve(G, np.ones(4))
    assert np.allclose(c, x / x.sum(), rtol=1e-10, atol=1e-12)
    assert abs(c.sum() - 1) < 1e-14


def test_coefficients_orthogonal_and_rank_deficient():
    norms = np.array([1.0, 2.0, 4.0])
    c = orc.anderson_coefficients(np.diag(norms**2), 0.0)
    w = 1 / norms**2
    assert np.allclose(c, w / w.sum(), rtol=1e-14)
    rng = np.random.default_rng(2)
    R = rng.normal(size=(2, 30))
    R = np.vstack([R, R[0]])  # a repeated residual: rank 2
    c = orc.anderson_coefficients(R @ R.T, 1e-12)
    assert np.all(np.isfinite(c)) and abs(c.sum() - 1) < 1e-12
    # the optimum is that of the two distinct residuals
    G2 = R[:2] @ R[:2].T
    d = np.linalg.solve(G2, np.ones(2))
    d /= d.sum()
    assert np.linalg.norm(c @ R) == pytest.approx(np.linalg.norm(d @ R[:2]), rel=1e-8)


def _causal_block(rng, n):
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Y = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A, Bm = X @ X.conj().T, Y @ Y.conj().T  # PSD
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    L, G = 1j * A, -1j * Bm          # -i S< = A, i S> = B
    R = H - 0.5j * (A + Bm)          # S^R - S^A = S> - S<, Im S^R <= 0
    return R, L, G


def _families(blocks, NE=2, Nk=3):
    n = blocks[0][0].shape[0]
    fams = []
    for R, L, G in blocks:
        f = [np.zeros((NE, Nk, n, n), np.complex128, order="F") for _ in range(3)]
        for k, X in enumerate((R, L, G)):
            f[k][:, :] = X
        fams.append(f)
    return fams


@pytest.mark.parametrize("code", [0, 1, 2, 3, 4, 5])
def test_guard_codes(code):
    rng = np.random.default_rng(10 + code)
    n = 5
    R, L, G = _causal_block(rng, n)
    if code == 1:
        L = L.copy(); L[1, 2] = np.nan
    elif code == 2:
        R = R + 0.5j * np.eye(n) * np.abs(np.linalg.eigvalsh(-(R - R.conj().T) / 2j)).max() * 3
    elif code == 3:
        B = rng.normal(size=(n, n)); L = L + 0.1 * (B + B.T)  # -i S< gains an anti-Hermitian part
    elif code == 4:
        w, V = np.linalg.eigh(-1j * L)
        w[0] = -0.2 * abs(w).max()  # -i S< Hermitian (code 3 passes) but indefinite
        L = 1j * (V @ np.diag(w) @ V.conj().T)
    elif code == 5:
        C = rng.normal(size=(n, n)); R = R + 1e-3 * (C + C.T)  # Hermitian error: Im S^R unchanged
        R = R - 1e-3j * np.eye(n)
    fams = _families([(R, L, G)])
    first, got = orc.anderson_guard(fams)
    if code == 0:
        assert (first, got) == (0, 0)
    else:
        assert got == code and first == 1


def test_guard_first_failure_order():
    rng = np.random.default_rng(3)
    blocks = [_causal_block(rng, 4) for _ in range(2)]
    fams = _families(blocks, NE=3, Nk=2)
    fams[1][0][2, 1] = fams[1][0][2, 1] + 5j * np.eye(4)  # family 2, e = 3, m = 2 (code 2)
    fams[1][1][1, 0, 0, 0] = np.inf                        # family 2, e = 2, m = 1 (code 1): first
    first, code = orc.anderson_guard(fams)
    assert (first, code) == (6 + 1 * 2 + 0 + 1, 1)


def _runner(lo_kind="low_rank", alpha=0.3):
    prob = fx.make_problem(N=4, NE=48, Nk=3, lo_kind=lo_kind, alpha_mix=alpha)
    full, state, _ = fx.oracle_state(prob)
    return orc.ScbaOracle(full, state)


def test_anderson_iteration_accelerates_and_stays_causal():
    lin, acc = _runner(), _runner()
    mixer = orc.AndersonOracle(depth=4)
    rl = [lin.iterate(measure=True)["r_Sigma"] for _ in range(6)]
    rows = [acc.iterate_anderson(mixer) for _ in range(6)]
    ra = [r["r_Sigma"] for r in rows]
    assert ra[0] == pytest.approx(rl[0], rel=1e-12)  # the first step is the damped candidate step
    assert sum(r["anderson"]["accepted"] for r in rows) >= 4
    assert ra[-1] < 0.5 * rl[-1], (ra, rl)
    for r in rows:  # a guard failure falls back to x + damping r, which passed the second guard
        a = r["anderson"]
        assert (a["first_failure"] == 0) == (a["reset_reason"] != "causality_guard")
