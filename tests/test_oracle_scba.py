"""The oracle's production SCBA iteration against its literal (E0) twin and
the identities every iteration must satisfy."""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx


def _iterate(orc, prob, n):
    full, state, _ = fx.oracle_state(prob)
    runner = orc.ScbaOracle(full, state)
    rows = [runner.iterate() for _ in range(n)]
    return full, state, rows


def test_iteration_keeps_number_and_spectral_identity(orc):
    # production_state.jl:49-70: the paired map fixes N[G<] = target and keeps G> - G<;
    # Dyson makes G> - G< = G^R - G^A whenever Sigma^R - Sigma^A = Sigma> - Sigma<.
    prob = fx.make_problem(N=4, NE=64, Nk=3, lo_kind="dense")
    full, state, rows = _iterate(orc, prob, 3)
    for r in rows:
        assert np.isfinite(r["lam"]) and r["lam"] > 0
        assert np.isfinite(r["Jbar"]) and r["roundoff"] < 1e-12
    num = orc.number_functional(state["GL"], prob["wE"], prob["wk"], 2)
    assert abs(num - full["target"]) < 1e-12 * full["target"]
    GR = state["GR"]
    GA = np.conj(np.transpose(GR, (0, 1, 3, 2)))
    rA = rel(state["GG"] - state["GL"], GR - GA)
    assert rA < 1e-10
    # The populations sum to one (the diagonal of the number functional).
    assert abs(rows[-1]["populations"].sum() - 1.0) < 1e-12


def test_candidate_equals_literal_twin(orc):
    # numerics/physics_first_exact_scba_candidate_equivalence.jl:4-81: the
    # production candidate equals the literal one (six-index sums, dense shift
    # matrices, direct principal value) to 2e-11; embedding to 3e-14.
    prob = fx.make_problem(N=3, NE=40, Nk=2, lo_kind="dense", alpha_mix=1.0)
    full, state, rows = _iterate(orc, prob, 1)  # alpha = 1: the state now holds the candidates
    eps, wE, wk = prob["eps"], prob["wE"], prob["wk"]
    NE, Nk, Nb = prob["NE"], prob["Nk"], prob["Nb"]
    green = {k: state[k] for k in ("GR", "GL", "GG")}
    Wp, Wm = orc.shift_matrix(eps, prob["hwLO"]), orc.shift_matrix(eps, -prob["hwLO"])
    for i, m in enumerate(prob["mechanisms"]):
        K = orc.make_kernel("literal", fx.k6_of(m, Nk, Nb))
        if m["is_lo"]:
            N = prob["NLO"]
            lp = orc.apply_shift_dense(Wp, green["GL"])
            lm = orc.apply_shift_dense(Wm, green["GL"])
            gp = orc.apply_shift_dense(Wp, green["GG"])
            gm = orc.apply_shift_dense(Wm, green["GG"])
            SL = (N + 1) * orc.contract(K, lp, wk, 1.0) + N * orc.contract(K, lm, wk, 1.0)
            SG = (N + 1) * orc.contract(K, gm, wk, 1.0) + N * orc.contract(K, gp, wk, 1.0)
        else:
            SL = orc.contract(K, green["GL"], wk, 1.0)
            SG = orc.contract(K, green["GG"], wk, 1.0)
        Gam = np.asfortranarray(1j * (SG - SL))
        SR = orc.direct_hilbert(Gam, eps, wE) - 0.5j * Gam
        assert rel(state["scatL"][i], SL) < 2e-11
        assert rel(state["scatG"][i], SG) < 2e-11
        assert rel(state["scatR"][i], SR) < 2e-11
    Ep = prob["Ep"]
    for key, comp in (("GR", "R"), ("GL", "L"), ("GG", "G")):
        for sign, T, name in ((1, prob["Tplus"], "plus"), (-1, prob["Tminus"], "minus")):
            X = orc.apply_shift_dense(orc.shift_matrix(eps, sign * Ep), green[key])
            expect = np.zeros_like(X)
            for e in range(NE):
                for mm in range(Nk):
                    expect[e, mm] = T @ X[e, mm] @ T.conj().T
            assert rel(state[name + comp], expect) < 3e-14


def test_lo_shift_then_contract_commutes(orc):
    # The device applies the LO shifts after the contraction (S K[G] = K[S G]);
    # the oracle follows the reference order. Both agree to round-off.
    prob = fx.make_problem(N=3, NE=48, Nk=2, lo_kind="dense")
    m = prob["mechanisms"][2]
    K = orc.make_kernel("dense", m["a"])
    rng = np.random.default_rng(2)
    G = np.asfortranarray(rng.normal(size=(48, 2, 3, 3)) + 1j * rng.normal(size=(48, 2, 3, 3)))
    plan = orc.shift_plan(prob["eps"], prob["hwLO"])
    a = orc.contract(K, orc.apply_shift(plan, G), prob["wk"], 1.0)
    b = orc.apply_shift(plan, orc.contract(K, G, prob["wk"], 1.0))
    assert rel(a, b) < 1e-14


def test_measured_iteration_composes_the_residual_operators(orc):
    """production_iteration.jl:102-166: the measurement of one iteration is
    the residual suite of the new Green state against the totals of this
    iteration and the unmixed candidate total, the largest self-energy
    residual over embedding, plus, minus and every mechanism, and |lambda-1|;
    restated here from the separately pinned operators."""
    prob = fx.make_problem(N=4, NE=48, Nk=3, lo_kind="low_rank")
    full, state, _ = fx.oracle_state(prob)
    M = len(prob["mechanisms"])
    fams = [[state["embR"], state["embL"], state["embG"]]] + \
        [[state["scatR"][i], state["scatL"][i], state["scatG"][i]] for i in range(M)]
    new = orc.current_green(prob["eps"], prob["h"], [f[0] for f in fams], [f[1] for f in fams],
                            [f[2] for f in fams], prob["wE"], prob["wk"], prob["gs"], full["target"])
    green = dict(GR=new["GR"], GL=new["GL"], GG=new["GG"], A=orc.spectral(new["GR"]))
    scat, plus, minus, emb = fx.oracle_candidates(prob, full["kernels"], green)
    tot = [np.asfortranarray(sum(f[k] for f in fams)) for k in range(3)]
    candL = np.asfortranarray(emb[1] + sum(c[1] for c in scat))
    ref = orc.residual_suite(prob["eps"], prob["h"], green["GR"], green["GL"], green["GG"], green["A"], *tot, candL)
    pairs = [(fams[0], emb), ([state["plusR"], state["plusL"], state["plusG"]], plus),
             ([state["minusR"], state["minusL"], state["minusG"]], minus)] + \
        [(fams[1 + i], scat[i]) for i in range(M)]
    ref["r_Sigma"] = max(orc.selfenergy_residual(o, c) for o, c in pairs)
    ref["r_lambda"] = abs(new["lam"] - 1.0)
    row = orc.ScbaOracle(full, state).iterate(measure=True)
    for k, v in ref.items():
        assert abs(row[k] - v) <= 1e-12 * abs(v) + 1e-15, (k, row[k], v)
    assert row["r_D"] < 1e-13 and row["r_Sigma"] > 0
