"""Gradients of the converged SCBA current with respect to a potential
profile on a real-space grid and to layer thicknesses (the differentiation
workload, reduced in size): the device fixed-point adjoint dJ/dh, carried to
dJ/dU by the chain rule of project_hamiltonians (dh/dU_g = Phi[g,:]^*
Phi[g,:]^T, basis.jl:300) and to dJ/dL by central differences of the
assembly h(L), against central differences of the ORACLE's converged
current J(h(U +- t dU)), J(h(L +- d))."""
import copy

import numpy as np
import pytest

import scba_fixture as fx
from test_gpu_adjoint import _oracle_converged_J

pytestmark = pytest.mark.gpu

NZ = 12


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _basis(rng, Nb):
    Q, _ = np.linalg.qr(rng.normal(size=(NZ, Nb)) + 1j * rng.normal(size=(NZ, Nb)))
    return np.asfortranarray(Q)  # orthonormal columns on the NZ grid points


def _potential_of_thicknesses(L, z, height=0.08, width=0.35):
    """A smooth layered potential: barriers of the given height between the
    interfaces cumsum(L) (odd layers are barriers), logistic edges."""
    edges = np.concatenate([[0.0], np.cumsum(L)])
    s = lambda x: 1.0 / (1.0 + np.exp(-x / width))
    U = np.zeros_like(z)
    for j in range(0, len(L), 2):
        U += height * (s(z - edges[j]) - s(z - edges[j + 1]))
    return U


@pytest.fixture(scope="module")
def converged(hip, orc):
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
    assert adjoint.converge(p, s, tol=1e-15) < 5000
    hbar, meta = adjoint.ScbaAdjoint(p).gradient(s, tol=1e-13)
    assert meta["change"] <= 1e-13
    Nb, Nk = prob["Nb"], prob["Nk"]
    kin = np.array([prob["h"][m][0, 0] - prob["h"][0][0, 0] for m in range(Nk)]).real  # Cbar kappa_m^2 shifts
    S = np.asfortranarray(prob["h"][0])
    return dict(prob=prob, full=full, state0=state0, hbar=hbar, S=S, kin=kin, Minv=np.eye(Nb, dtype=np.complex128),
                Phi=_basis(np.random.default_rng(77), Nb))


def test_potential_gradient_vs_oracle_fd(hip, orc, converged):
    """dJ/dU from api.potential_gradient in two directions of a potential on
    a 12-point grid, h(U) by the oracle's project_hamiltonians
    (basis.jl:288-308); a transposed or unconjugated chain rule fails it."""
    c = converged
    h0 = orc.project_hamiltonians(c["Phi"], np.zeros(NZ), c["S"], c["kin"], c["Minv"])
    assert np.abs(h0 - c["prob"]["h"]).max() < 1e-14
    gU = hip.potential_gradient(c["hbar"], c["Phi"])
    assert gU.shape == (NZ,)
    rng = np.random.default_rng(31)
    for k in range(2):
        dU = rng.normal(size=NZ) * 1e-3
        t = 0.3  # central-difference truncation O(t^2) well below the bound
        Jp = _oracle_converged_J(c["full"], c["state0"], orc.project_hamiltonians(c["Phi"], t * dU, c["S"], c["kin"],
                                                                                  c["Minv"]))
        Jm = _oracle_converged_J(c["full"], c["state0"], orc.project_hamiltonians(c["Phi"], -t * dU, c["S"],
                                                                                  c["kin"], c["Minv"]))
        fd = (Jp - Jm) / (2 * t)
        ad = float(gU @ dU)
        assert abs(fd - ad) <= 2e-6 * abs(fd), (k, fd, ad)


def test_thickness_gradient_vs_oracle_fd(hip, orc, converged):
    """dJ/dL_i of four layer thicknesses that shape a barrier potential on
    the grid: adjoint.thickness_gradient (chain rule through central
    differences of the assembly) against central differences of the
    oracle's converged current."""
    from qclhip import adjoint
    c = converged
    z = np.linspace(0.0, 4.0, NZ)
    L0 = np.array([0.7, 1.1, 0.9, 1.2])
    assemble = lambda L: orc.project_hamiltonians(c["Phi"], _potential_of_thicknesses(L, z), c["S"], c["kin"],
                                                  c["Minv"])
    # the reference point: the converged state belongs to h(U = 0), so take
    # h(L) relative to h(L0) (a fixed offset does not change dh/dL)
    off = assemble(L0) - c["prob"]["h"]
    shifted = lambda L: assemble(L) - off
    gL = adjoint.thickness_gradient(c["hbar"], shifted, L0, step=1e-4)
    for i in range(4):
        d = np.zeros(4)
        d[i] = 1e-3
        Jp = _oracle_converged_J(c["full"], c["state0"], shifted(L0 + d))
        Jm = _oracle_converged_J(c["full"], c["state0"], shifted(L0 - d))
        fd = (Jp - Jm) / (2 * d[i])
        assert abs(fd - gL[i]) <= 1e-5 * abs(fd), (i, fd, gL[i])


def test_adjoint_checkpoints(hip, converged):
    """Gradients formed at adjoint checkpoints: the one at the last iteration
    equals the returned gradient bitwise, and earlier ones approach it."""
    import torch
    from qclhip import adjoint, scba
    c = converged
    prob = c["prob"]
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, c["full"]["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"], verify_roundoff=False)
    dev = lambda x: hip.to_device(x)
    st = c["state0"]
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(st["scatR"][i]), dev(st["scatL"][i]), dev(st["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(st["plusR"]), dev(st["plusL"]), dev(st["plusG"])),
                       scba.Family(dev(st["minusR"]), dev(st["minusL"]), dev(st["minusG"])))
    adjoint.converge(p, s, tol=1e-15)
    hbar, meta = adjoint.ScbaAdjoint(p).gradient(s, tol=0.0, max_iterations=60, checkpoints=(20, 40, 60))
    its = [it for it, _ in meta["checkpoints"]]
    assert its == [20, 40, 60]
    assert torch.equal(meta["checkpoints"][-1][1], hbar)
    err = [float(torch.linalg.norm(hb - c["hbar"]) / torch.linalg.norm(c["hbar"])) for _, hb in meta["checkpoints"]]
    assert err[2] < err[0], err


def test_shared_adjoint_families_equal_separate(hip, converged):
    """The linear adjoint iteration carries the families it keeps equal (minus
    and every mechanism's; plus's R) in one copy: after a fixed number of
    iterations the gradient, and the gradients at checkpoints, equal those of
    the iteration that carries every family on its own, bit for bit; the
    change measure (the same terms summed in another grouping) agrees to
    rounding."""
    import torch
    from qclhip import adjoint, scba
    c = converged
    prob = c["prob"]
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                         hip.matrix_to_device(prob["Tminus"]), mechs, c["full"]["target"], prob["Ep"], prob["hwLO"],
                         prob["NLO"], prob["alpha_mix"], verify_roundoff=False)
    dev = lambda x: hip.to_device(x)
    st = c["state0"]
    s = scba.ScbaState(None, {m["name"]: scba.Family(dev(st["scatR"][i]), dev(st["scatL"][i]), dev(st["scatG"][i]))
                              for i, m in enumerate(prob["mechanisms"])},
                       scba.Family(dev(st["plusR"]), dev(st["plusL"]), dev(st["plusG"])),
                       scba.Family(dev(st["minusR"]), dev(st["minusL"]), dev(st["minusG"])))
    adjoint.converge(p, s, tol=1e-15)
    adj = adjoint.ScbaAdjoint(p)
    runs = [adj.gradient(s, tol=0.0, max_iterations=40, checkpoints=(15,), share_families=sh) for sh in (True, False)]
    (h1, m1), (h2, m2) = runs
    assert torch.equal(h1, h2)
    assert torch.equal(m1["checkpoints"][0][1], m2["checkpoints"][0][1])
    assert m1["iterations"] == m2["iterations"] == 40
    assert abs(m1["change"] - m2["change"]) <= 1e-12 * m2["change"]


def test_anderson_adjoint_equals_linear(hip, orc):
    """The Anderson-accelerated adjoint iteration reaches the gradient of the
    damped linear one (the same fixed point) in fewer iterations, and a
    forward converged with Anderson mixing (the reference's policy, which
    restarts often here and is not faster) gives the same dJ/dh."""
    from qclhip import adjoint, scba
    prob = fx.make_problem(N=4, NE=64, Nk=3, alpha_mix=0.5)
    full, state, green = fx.oracle_state(prob)
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    mechs = [scba.Mechanism(m["name"], hip.ProductionKernel(g, m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]),
                            m["qK"], m["is_lo"]) for m in prob["mechanisms"]]
    dev = lambda x: hip.to_device(x)

    def run(mixing):
        p = scba.ScbaProblem(g, hip.h_to_device(prob["h"]), hip.matrix_to_device(prob["Tplus"]),
                             hip.matrix_to_device(prob["Tminus"]), mechs, full["target"], prob["Ep"], prob["hwLO"],
                             prob["NLO"], prob["alpha_mix"], verify_roundoff=False, mixing=mixing)
        s = scba.ScbaState(None, {m["name"]: scba.Family(dev(state["scatR"][i]), dev(state["scatL"][i]),
                                                          dev(state["scatG"][i]))
                                  for i, m in enumerate(prob["mechanisms"])},
                           scba.Family(dev(state["plusR"]), dev(state["plusL"]), dev(state["plusG"])),
                           scba.Family(dev(state["minusR"]), dev(state["minusL"]), dev(state["minusG"])))
        it = adjoint.converge(p, s, tol=1e-15)
        return p, s, it

    p_lin, s_lin, it_lin = run("linear")
    p_and, s_and, it_and = run("anderson")
    print("forward iterations linear", it_lin, "anderson", it_and, flush=True)
    h_lin, m_lin = adjoint.ScbaAdjoint(p_lin).gradient(s_lin, tol=1e-13)
    h_and, m_and = adjoint.ScbaAdjoint(p_lin).gradient(s_lin, tol=1e-13, acceleration="anderson")
    h_fa, m_fa = adjoint.ScbaAdjoint(p_and).gradient(s_and, tol=1e-13, acceleration="anderson")
    print("adjoint iterations linear", m_lin["iterations"], "anderson", m_and["iterations"], flush=True)
    ref = h_lin.cpu().numpy()
    for h in (h_and, h_fa):
        assert np.linalg.norm(h.cpu().numpy() - ref) <= 1e-9 * np.linalg.norm(ref)
    assert m_and["iterations"] < m_lin["iterations"]
