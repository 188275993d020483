"""The Poisson step closing the SCBA loop, on the device against the oracle
(observables.jl:1-26, poisson.jl:12-65, basis.jl:288-308,
production_solver.jl:1044-1047): electron density, bordered periodic
Poisson solve (neutral and non-neutral densities, N_z up to 1024),
Hamiltonian projection and potential mixing (FP64 bar 1e-12)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _lesser(rng, NE, Nk, Nb):
    X = rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb))
    return np.asfortranarray(1j * np.einsum("emab,emcb->emac", X, X.conj()))


@pytest.mark.parametrize("Nb,Nz", [(3, 17), (8, 256), (24, 512)])
def test_density_vs_oracle(hip, orc, Nb, Nz):
    rng = np.random.default_rng(Nb)
    NE, Nk = 40, 3
    GL = _lesser(rng, NE, Nk, Nb)
    eps = np.linspace(-0.2, 0.5, NE)
    wE, wk = np.full(NE, eps[1] - eps[0]), rng.uniform(0.5, 1.5, Nk)
    chi = rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb))
    g = hip.Grid(eps, wE, wk, Nb)
    n = hip.electron_density(g, hip.to_device(GL), hip.basis_to_device(chi), 2.0)
    ref = orc.electron_density(GL, wE, wk, chi, 2.0)
    assert float(rel(n.cpu().numpy(), ref)) < 1e-13


def _poisson_inputs(rng, N, neutral):
    er = rng.uniform(8, 14, N)
    ND = np.where((np.arange(N) % 7) < 2, rng.uniform(1, 2, N), 0.0)
    nbar = rng.uniform(0, 1, N)
    if neutral:
        nbar = nbar * (ND.sum() / nbar.sum())
    return er, np.full(N, 0.2), ND, nbar


@pytest.mark.parametrize("N,neutral", [(3, True), (64, False), (512, True), (512, False), (1024, True)])
def test_poisson_vs_oracle(hip, orc, N, neutral):
    import torch
    rng = np.random.default_rng(N)
    er, wx, ND, nbar = _poisson_inputs(rng, N, neutral)
    lam = 2.5
    U, zeta, rP, neut = hip.periodic_poisson(er, wx, ND, lam, torch.as_tensor(nbar, device="cuda"))
    Uo, zo, rPo, no = orc.periodic_poisson(er, wx, ND, lam, nbar)
    assert float(rel(U.cpu().numpy(), Uo)) < 1e-12
    assert abs(zeta - zo) <= 1e-12 * max(abs(zo), 1e-300) + 1e-15
    assert abs(rP - rPo) <= 1e-10 * rPo + 1e-14
    assert abs(neut - no) <= 1e-12 * no + 1e-16
    if neutral:
        assert rP < 1e-12 and abs(zeta) < 1e-12


def test_projection_vs_oracle(hip, orc):
    import torch
    rng = np.random.default_rng(9)
    Nz, Nb, Nk = 512, 24, 7
    Phi = rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb))
    U = rng.normal(size=Nz)
    S = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
    Minv = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
    kin = rng.uniform(0, 3, Nk)
    h = hip.project_hamiltonians(hip.basis_to_device(Phi), torch.as_tensor(U, device="cuda"),
                                 hip.matrix_to_device(S), kin, hip.matrix_to_device(Minv))
    ref = orc.project_hamiltonians(Phi, U, S, kin, Minv)
    got = np.asfortranarray(h.cpu().numpy().T)  # device (Nb,Nb,Nk) -> reference (Nk,Nb,Nb)
    assert float(rel(got, ref)) < 1e-13


def test_mix_potential_vs_oracle(hip, orc):
    import torch
    rng = np.random.default_rng(3)
    U = rng.normal(size=700)
    cand = U + rng.normal(size=700)
    Ud = torch.as_tensor(U, device="cuda")
    r = hip.mix_potential(Ud, torch.as_tensor(cand, device="cuda"), 0.4)
    ro = orc.mix_potential(U, cand, 0.4)
    assert r == ro
    assert np.array_equal(Ud.cpu().numpy(), U)
