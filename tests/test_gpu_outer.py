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


def test_outer_step_after_scba_vs_oracle(hip, orc):
    """Two SCBA iterations on device and oracle, then one Poisson step on
    each: density, mixed Hartree energy, residuals and the next h agree."""
    import torch
    import scba_fixture as fx
    from qclhip import scba, outer
    prob = fx.make_problem(N=4, NE=64, Nk=3)
    full, state, _ = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"])
    dev = hip.to_device
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                      dev(state["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                       scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    runner = orc.ScbaOracle(full, state)
    ws = scba.ScbaWorkspace(p)
    for _ in range(2):
        runner.iterate()
        scba.scba_iteration(p, s, ws)
    rng = np.random.default_rng(12)
    Nz, Nb, dx = 96, prob["Nb"], 0.25
    Q, _ = np.linalg.qr(rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb)))
    Phi, chi = Q, Q / np.sqrt(dx)
    er = rng.uniform(10, 13, Nz)
    ND = np.where(np.arange(Nz) < 12, 4.0, 0.0)
    S = np.diag(np.linspace(0, 0.3, Nb)) + 0j
    Minv = np.eye(Nb) + 0j
    kin = np.array([0.0, 0.05, 0.2])
    U0 = 0.01 * np.cos(2 * np.pi * np.arange(Nz) / Nz)
    op = outer.OuterProblem(torch.as_tensor(er, device="cuda"), torch.full((Nz,), dx, device="cuda", dtype=torch.float64),
                            torch.as_tensor(ND, device="cuda"), 1.3, hip.basis_to_device(chi),
                            hip.basis_to_device(Phi), hip.matrix_to_device(S), torch.as_tensor(kin, device="cuda"),
                            hip.matrix_to_device(Minv), gs=2.0, alpha_P=0.4, ND2D=float(dx * ND.sum()), Lp=Nz * dx)
    U = torch.as_tensor(U0, device="cuda")
    row = outer.poisson_step(op, g, s.green.GL, U)
    n_o = orc.electron_density(state["GL"], prob["wE"], prob["wk"], chi, 2.0)
    assert float(rel(row["nbar"], n_o)) < 1e-12
    cand, zeta, rP, neut = orc.periodic_poisson(er, np.full(Nz, dx), ND, 1.3, n_o)
    U_o = U0.copy()
    rU = orc.mix_potential(U_o, cand, 0.4)
    assert float(rel(U.cpu().numpy(), U_o)) < 1e-11
    assert abs(row["r_U"] - rU) <= 1e-11 * rU
    assert abs(row["zeta"] - zeta) <= 1e-11 * abs(zeta) + 1e-15 and abs(row["r_P"] - rP) <= 1e-10 * rP + 1e-14
    h_o = orc.project_hamiltonians(Phi, U_o, S, kin, Minv)
    assert float(rel(np.asfortranarray(row["h"].cpu().numpy().T), h_o)) < 1e-12


def test_interchange_problem_runs_like_the_in_memory_one(hip, orc, tmp_path):
    """A problem written to disk and read back runs the SCBA iteration of the
    oracle on the in-memory problem it was written from (1e-12)."""
    import torch
    import scba_fixture as fx
    from qclhip import interchange as ic, scba
    prob = fx.make_problem(N=4, NE=64, Nk=3)
    full, state, _ = fx.oracle_state(prob)
    arrays = {k: prob[k] for k in ("eps", "wE", "wk", "h", "Tplus", "Tminus")}
    scalars = dict(gs=2, target=full["target"], Ep=prob["Ep"], hwLO=prob["hwLO"], NLO=prob["NLO"],
                   kbt=prob["kbt"], alpha_mix=prob["alpha_mix"])
    ic.save_problem(tmp_path, arrays, scalars, prob["mechanisms"])
    p, op = ic.device_problem(ic.load_problem(tmp_path))
    assert op is None
    dev = hip.to_device
    mk = lambda: scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                               dev(state["scatG"][i]))
                                       for i, m in enumerate(prob["mechanisms"])},
                                scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                                scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
    s1 = mk()
    r1 = scba.scba_iteration(p, s1, scba.ScbaWorkspace(p))
    ro = orc.ScbaOracle(full, state).iterate()
    assert abs(r1["lam"] - ro["lam"]) < 1e-12 * abs(ro["lam"])
    assert float(rel(r1["populations"], ro["populations"])) < 1e-12
