"""Synthetic SCBA problems shared by the oracle and GPU tests.

A seeded banded Hamiltonian on a uniform energy grid with three mechanisms
in the reference's representations: an elastic momentum-independent
covariance map (acoustic-like), a separable IFR-like map, and a thermal LO
map (low rank, or dense for the tiny cases), plus the bulk embedding with a
coupling T+ (T- = T+^dagger).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as orc
from qclhip import synthetic as syn


def make_problem(N=4, NE=64, Nk=3, seed=17, lo_kind="low_rank", alpha_mix=0.3):
    w = syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=seed)
    rng = np.random.default_rng(seed + 1)
    mech = []
    block = syn.covariance_block(rng, N, nvert=2, scale=0.08)
    mech.append(dict(name="acoustic", kind="momentum_independent", a=block, b=None, c=None, rank=0,
                     qK=1.0, is_lo=False))
    left, right, radial = syn.separable_factors(rng, N, Nk, J=2, scale=0.05)
    mech.append(dict(name="IFR", kind="separable_ifr", a=left, b=right, c=radial, rank=2, qK=1.0,
                     is_lo=False))
    if lo_kind == "dense":
        F = syn.dense_covariance_kernel(rng, N, Nk, nvert=2, scale=0.05)
        mech.append(dict(name="LO", kind="dense", a=F, b=None, c=None, rank=0, qK=1.0, is_lo=True))
    else:
        l, r = syn.low_rank_factors(rng, N, Nk, r=3, scale=0.05)
        mech.append(dict(name="LO", kind="low_rank", a=l, b=r, c=None, rank=3, qK=1.0, is_lo=True))
    T = 0.05 * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    kbt = w["kbt"]
    hwLO = 0.36
    prob = dict(NE=NE, Nk=Nk, Nb=N, gs=2, eps=w["eps"], wE=w["wE"], wk=w["wk"],
                h=np.asfortranarray(w["h"]), Tplus=np.asfortranarray(T), Tminus=np.asfortranarray(T.conj().T),
                Ep=0.5, hwLO=hwLO, NLO=1.0 / math.expm1(hwLO / kbt), alpha_mix=alpha_mix, mechanisms=mech,
                kbt=kbt, eta_seed=0.05, energy_chunk=16)
    return prob


def make_problem5(N=6, NE=128, Nk=4, seed=29, alpha_mix=0.3):
    """The five mechanisms of the full-iteration workload in their device
    representations: acoustic and alloy momentum-independent, IFR separable,
    LO low-rank and thermal, ionized impurity low-rank."""
    prob = make_problem(N=N, NE=NE, Nk=Nk, seed=seed, alpha_mix=alpha_mix)
    rng = np.random.default_rng(seed + 2)
    mech = [m for m in prob["mechanisms"] if m["name"] in ("acoustic", "IFR", "LO")]
    mech.insert(1, dict(name="alloy", kind="momentum_independent", a=syn.covariance_block(rng, N, 3, 0.04),
                        b=None, c=None, rank=0, qK=1.0, is_lo=False))
    l, r = syn.low_rank_factors(rng, N, Nk, r=2, scale=0.04)
    mech.append(dict(name="impurity", kind="low_rank", a=l, b=r, c=None, rank=2, qK=1.0, is_lo=False))
    prob["mechanisms"] = mech
    return prob


def oracle_kernels(prob):
    return [orc.make_kernel(m["kind"], m["a"], m["b"], m["c"], rank=m["rank"]) for m in prob["mechanisms"]]


def oracle_seed(prob, mu=1.0):
    """Seed Green state at a fixed mu; the target number is the seed's number."""
    GR, scale = orc.retarded_green(prob["eps"], prob["h"], None, eta=prob["eta_seed"])
    q = orc.seed_weights(GR, prob["wE"], prob["wk"], prob["gs"])
    target = float(np.sum(q * np.array([orc.fermi(e, mu, prob["kbt"]) for e in prob["eps"]])))
    GL = orc.seed_lesser(GR, prob["eps"], mu, prob["kbt"])
    A = orc.spectral(GR)
    GG = np.asfortranarray((GR - np.conj(np.transpose(GR, (0, 1, 3, 2)))) + GL)
    return dict(GR=GR, GL=GL, GG=GG, A=A, scale=scale), target


def oracle_candidates(prob, kernels, green):
    """Fresh candidate families from a Green state (production_solver.jl:117-141)."""
    eps, wE, wk = prob["eps"], prob["wE"], prob["wk"]
    pLO, mLO = orc.shift_plan(eps, prob["hwLO"]), orc.shift_plan(eps, -prob["hwLO"])
    pE, mE = orc.shift_plan(eps, prob["Ep"]), orc.shift_plan(eps, -prob["Ep"])
    scat = []
    for m, k in zip(prob["mechanisms"], kernels):
        if m["is_lo"]:
            SL, SG = orc.lo_contraction(k, green["GL"], green["GG"], wk, m["qK"], prob["NLO"], pLO, mLO,
                                        prob["energy_chunk"])
        else:
            SL = orc.contract(k, green["GL"], wk, m["qK"], prob["energy_chunk"])
            SG = orc.contract(k, green["GG"], wk, m["qK"], prob["energy_chunk"])
        SR, _, _ = orc.retarded_self_energy_fft(SL, SG, eps, wE)
        scat.append([SR, SL, SG])
    plus = [orc.embedding_component(prob["Tplus"], green[k], pE) for k in ("GR", "GL", "GG")]
    minus = [orc.embedding_component(prob["Tminus"], green[k], mE) for k in ("GR", "GL", "GG")]
    emb = [np.asfortranarray(a + b) for a, b in zip(plus, minus)]
    return scat, plus, minus, emb


def oracle_state(prob):
    kernels = oracle_kernels(prob)
    green, target = oracle_seed(prob)
    scat, plus, minus, emb = oracle_candidates(prob, kernels, green)
    NE, Nk, Nb = prob["NE"], prob["Nk"], prob["Nb"]
    state = dict(scatR=[s[0] for s in scat], scatL=[s[1] for s in scat], scatG=[s[2] for s in scat],
                 plusR=plus[0], plusL=plus[1], plusG=plus[2], minusR=minus[0], minusL=minus[1],
                 minusG=minus[2], embR=emb[0], embL=emb[1], embG=emb[2],
                 GR=orc.field(NE, Nk, Nb), GL=orc.field(NE, Nk, Nb), GG=orc.field(NE, Nk, Nb),
                 A=orc.field(NE, Nk, Nb), scale=np.zeros((NE, Nk), order="F"))
    full = dict(prob, target=target, kernels=kernels, qK=[m["qK"] for m in prob["mechanisms"]],
                is_lo=[int(m["is_lo"]) for m in prob["mechanisms"]])
    return full, state, green


def k6_of(m, Nk, Nb):
    """Literal six-index tensor K[m,m',a,c,d,b] of a mechanism's representation."""
    N2, Nc = Nb * Nb, Nk * Nb * Nb
    if m["kind"] == "momentum_independent":
        F = np.zeros((Nc, Nc), np.complex128)
        for q in range(N2):
            for qp in range(N2):
                F[q * Nk:(q + 1) * Nk, qp * Nk:(qp + 1) * Nk] = m["a"][q, qp]
    elif m["kind"] == "separable_ifr":
        core = m["a"] @ m["b"]
        F = np.zeros((Nc, Nc), np.complex128)
        for q in range(N2):
            for qp in range(N2):
                F[q * Nk:(q + 1) * Nk, qp * Nk:(qp + 1) * Nk] = m["c"] * core[q, qp]
    elif m["kind"] == "low_rank":
        F = m["a"] @ m["b"]
    else:
        F = m["a"]
    K = np.zeros((Nk, Nk, Nb, Nb, Nb, Nb), np.complex128, order="F")
    for a in range(Nb):
        for b in range(Nb):
            for c in range(Nb):
                for d in range(Nb):
                    K[:, :, a, c, d, b] = F[np.arange(Nk)[:, None] + Nk * (a + Nb * b),
                                            np.arange(Nk)[None, :] + Nk * (c + Nb * d)]
    return K
