"""Pins of the oracle's outer-loop operators (the Poisson step that closes
the SCBA loop): the electron density against the number functional (sum
rule for an orthonormal basis), the bordered periodic Poisson solve against
the discrete-Laplacian eigenvectors (cosine densities), zeta = v.b/sP and
numpy's dense solve, the Hamiltonian projection against Phi^dagger c Phi =
c I, and the potential mixing."""
import numpy as np
import pytest

from oracle import oracle as orc


def _basis(rng, Nz, Nb, dx):
    Q, _ = np.linalg.qr(rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb)))
    return np.asfortranarray(Q), np.asfortranarray(Q / np.sqrt(dx))  # Phi (orthonormal), chi


def _lesser(rng, NE, Nk, Nb):
    X = rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb))
    return np.asfortranarray(1j * np.einsum("emab,emcb->emac", X, X.conj()))  # -i G< PSD


def test_density_sum_rule_equals_number_functional():
    rng = np.random.default_rng(1)
    NE, Nk, Nb, Nz, dx = 12, 3, 5, 40, 0.25
    Phi, chi = _basis(rng, Nz, Nb, dx)
    GL = _lesser(rng, NE, Nk, Nb)
    wE, wk = rng.uniform(0.5, 1.5, NE), rng.uniform(0.5, 1.5, Nk)
    n = orc.electron_density(GL, wE, wk, chi, 2.0)
    assert np.all(n > 0)
    assert np.sum(dx * n) == pytest.approx(orc.number_functional(GL, wE, wk, 2), rel=1e-13)


def test_density_single_coherence():
    # G< with one coherence (a, b) = (0, 1) only: n_g = gs/(2 pi) Re(-i chi_g0 I_01 conj(chi_g1))
    rng = np.random.default_rng(2)
    NE, Nk, Nb, Nz = 4, 2, 3, 7
    GL = np.zeros((NE, Nk, Nb, Nb), np.complex128, order="F")
    GL[:, :, 0, 1] = 0.3 + 0.7j
    wE, wk = np.full(NE, 0.5), np.full(Nk, 0.25)
    chi = np.asfortranarray(rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb)))
    I01 = (0.3 + 0.7j) * NE * Nk * 0.125
    ref = 2.0 / (2 * np.pi) * np.real(-1j * chi[:, 0] * I01 * np.conj(chi[:, 1]))
    assert np.allclose(orc.electron_density(GL, wE, wk, chi, 2.0), ref, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("k", [1, 3])
def test_poisson_cosine_eigenvector(k):
    N, dx, lam = 64, 0.5, 3.0
    g = np.arange(N)
    ND = np.full(N, 2.0)
    nbar = ND - 0.1 * np.cos(2 * np.pi * k * g / N)
    U, zeta, rP, neut = orc.periodic_poisson(np.ones(N), np.full(N, dx), ND, lam, nbar)
    mu = (2 - 2 * np.cos(2 * np.pi * k / N)) / dx**2
    b = -lam * (ND - nbar)
    assert np.max(np.abs(U - b / mu)) <= 1e-12 * np.max(np.abs(b / mu))
    assert abs(zeta) < 1e-14 and rP < 1e-13 and neut < 1e-15


def test_poisson_vs_numpy_bordered_solve():
    rng = np.random.default_rng(4)
    N, dx, lam = 48, 0.3, 1.7
    er = rng.uniform(8, 14, N)
    ND = rng.uniform(0, 2, N)
    nbar = rng.uniform(0, 2, N)
    U, zeta, rP, neut = orc.periodic_poisson(er, np.full(N, dx), ND, lam, nbar)
    face = 2 * er * np.roll(er, -1) / (er + np.roll(er, -1))
    L = np.zeros((N, N))
    for i in range(N):
        L[i, i] = (face[i - 1] + face[i]) / dx**2
        L[i, i - 1] = -face[i - 1] / dx**2
        L[i, (i + 1) % N] = -face[i] / dx**2
    b = -lam * (ND - nbar)
    sP = np.max(np.sum(np.abs(L), axis=1))
    v = np.full(N, 1 / np.sqrt(N))
    M = np.block([[L / sP, v[:, None]], [v[None, :], np.zeros((1, 1))]])
    x = np.linalg.solve(M, np.concatenate([b / sP, [0.0]]))
    assert np.max(np.abs(U - x[:N])) <= 1e-10 * np.max(np.abs(x[:N]))
    assert zeta == pytest.approx(v @ b / sP, rel=1e-10)  # v^T L = 0
    assert abs(np.sum(U)) < 1e-10 * np.max(np.abs(U))   # the border row
    # L U - b = -zeta sP v: the Poisson residual is the non-neutral part
    assert rP == pytest.approx(abs(zeta) * sP / (np.linalg.norm(b) + lam * 1e-14), rel=1e-9)
    assert neut == pytest.approx(abs(np.sum(dx * (ND - nbar))) / np.sum(dx * ND), rel=1e-12)
    nbar_neutral = nbar - np.mean(nbar - ND)  # sum (N_D - nbar) = 0
    U2, zeta2, rP2, neut2 = orc.periodic_poisson(er, np.full(N, dx), ND, lam, nbar_neutral)
    assert abs(zeta2) < 1e-13 and rP2 < 1e-12 and neut2 < 1e-14


def test_projection_constant_potential():
    rng = np.random.default_rng(5)
    Nz, Nb, Nk = 30, 6, 4
    Phi, _ = _basis(rng, Nz, Nb, 1.0)
    S = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
    Minv = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
    kin = np.array([0.0, 0.1, 0.5, 2.0])
    h = orc.project_hamiltonians(Phi, np.full(Nz, 0.37), S, kin, Minv)
    for m in range(Nk):
        assert np.allclose(h[m], S + 0.37 * np.eye(Nb) + kin[m] * Minv, rtol=0, atol=1e-14)
    U = rng.normal(size=Nz)
    h = orc.project_hamiltonians(Phi, U, S, kin, Minv)
    assert np.allclose(h[2], S + Phi.conj().T @ np.diag(U) @ Phi + kin[2] * Minv, rtol=0, atol=1e-13)


def test_mix_potential():
    U = np.array([1.0, 2.0, -1.0])
    r = orc.mix_potential(U, np.array([1.5, 2.0, -3.0]), 0.25)
    assert r == 2.0 and np.allclose(U, [1.125, 2.0, -1.5], rtol=0, atol=0)
