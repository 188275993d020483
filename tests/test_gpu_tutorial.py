"""The reference's tutorial problem (reference_parameters, tutorial_numerics,
acoustic-only scattering; restated in oracle/tutorial.py, whose equality with
Julia's assembly is unpinned here) solved by the device SCBA loop and by the
oracle: the seed at the sheet doping, then up to the tutorial's max_scba = 80
iterations with alpha = 0.35, every iteration's row compared, stopping when
the tutorial's fixed-point gates pass. Prints the per-iteration residual rows,
the current density (A/cm^2) and the occupations."""
import numpy as np
import pytest

from conftest import rel
import tutorial_fixture as tf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def test_tutorial_scba_vs_oracle(hip, orc):
    from oracle import tutorial as tu
    from qclhip import scba
    t = tu.tutorial_problem()
    prob = tf.oracle_problem(t)
    full, state, mu_o = tf.oracle_state(prob, t["target"])
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"])
    h = hip.h_to_device(prob["h"])
    seed, mu = hip.seed_green_production(g, h, prob["eta_seed"], prob["kbt"], t["target"])
    assert abs(mu - mu_o) <= 1e-12 * max(1.0, abs(mu_o))
    mech = [scba.Mechanism("acoustic", hip.ProductionKernel(g, "momentum_independent", t["acoustic"]), t["qK"])]
    p = scba.ScbaProblem(g, h, hip.matrix_to_device(prob["Tplus"]), hip.matrix_to_device(prob["Tminus"]), mech,
                         t["target"], prob["Ep"], prob["hwLO"], 0.0, prob["alpha_mix"])
    s = scba.initial_state(p, seed)
    ws = scba.ScbaWorkspace(p)
    runner = orc.ScbaOracle(full, state)
    gates = t["options"]["gates"]
    print("\ntutorial: N_b=3 N_E=49 N_k=5, acoustic only, alpha 0.35; seed mu %.9f (E0 = 0.1 eV)" % mu)
    print(" it        lambda            J (A/cm^2)      r_Sigma     r_K         r_lambda    r_A         r_D")
    converged = None
    for it in range(t["options"]["max_scba"]):
        ro = runner.iterate(measure=True)
        rd = scba.scba_iteration(p, s, ws)
        assert abs(rd["lam"] - ro["lam"]) <= 1e-11 * abs(ro["lam"]), (it, rd["lam"], ro["lam"])
        assert abs(rd["Jbar"] - ro["Jbar"]) <= 1e-10 * abs(ro["Jbar"]) + 1e-16, (it, rd["Jbar"], ro["Jbar"])
        assert float(rel(rd["populations"], ro["populations"])) < 1e-10, it
        for k in ("r_Sigma", "r_K", "r_A"):
            assert abs(rd[k] - ro[k]) <= 1e-8 * ro[k] + 1e-13, (it, k, rd[k], ro[k])
        assert rd["r_D"] < 1e-13 and rd["roundoff"] < t["options"]["r_roundoff"]
        print("%3d  %.12f  %+.9e  %.3e  %.3e  %.3e  %.3e  %.3e" % (
            it + 1, rd["lam"], rd["Jbar"] * t["J0_A_cm2"], rd["r_Sigma"], rd["r_K"], rd["r_lambda"], rd["r_A"],
            rd["r_D"]))
        if tf.gates_pass(rd, gates):
            converged = it + 1
            break
    last = s.history[-1]
    print("gates", gates, "passed at iteration", converged)
    print("J = %.9e A/cm^2, occupations %s" % (last["Jbar"] * t["J0_A_cm2"], np.array2string(last["populations"],
                                                                                            precision=9)))
    for k in ("GR", "GL", "GG"):
        assert float(rel(hip.to_host(getattr(s.green, k)), state[k])) < 1e-10, k
