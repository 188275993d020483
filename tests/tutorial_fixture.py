"""The acoustic-only tutorial problem (oracle/tutorial.py's restatement of
the reference's assembly) as oracle and device SCBA problems."""
import numpy as np

from oracle import oracle as orc
from oracle import tutorial as tu

import scba_fixture as fx


def oracle_problem(t):
    g = t["grids"]
    b = t["basis"]
    mech = [dict(name="acoustic", kind="momentum_independent", a=t["acoustic"], b=None, c=None, rank=0, qK=t["qK"],
                 is_lo=False)]
    return dict(NE=len(g["eps"]), Nk=len(g["kappa"]), Nb=t["numerics"]["N_b"], gs=2, eps=g["eps"], wE=g["wE"],
                wk=g["wk"], h=np.asfortranarray(t["h"]), Tplus=np.asfortranarray(b["Tplus"]),
                Tminus=np.asfortranarray(b["Tminus"]), Ep=t["Ep"], hwLO=0.367, NLO=0.0,
                alpha_mix=t["options"]["alpha"], mechanisms=mech, kbt=t["kbt"], eta_seed=t["eta_seed"],
                energy_chunk=16)


def oracle_state(prob, target):
    """Seed at the chemical potential whose number is the sheet doping
    (seed.jl:151-197), then the first candidates as the state."""
    GR, scale = orc.retarded_green(prob["eps"], prob["h"], None, eta=prob["eta_seed"])
    q = orc.seed_weights(GR, prob["wE"], prob["wk"], prob["gs"])
    mu = orc.seed_mu(prob["eps"], q, prob["kbt"], target)
    GL = orc.seed_lesser(GR, prob["eps"], mu, prob["kbt"])
    GG = np.asfortranarray((GR - np.conj(np.transpose(GR, (0, 1, 3, 2)))) + GL)
    green = dict(GR=GR, GL=GL, GG=GG, A=orc.spectral(GR), scale=scale)
    kernels = fx.oracle_kernels(prob)
    scat, plus, minus, emb = fx.oracle_candidates(prob, kernels, green)
    NE, Nk, Nb = prob["NE"], prob["Nk"], prob["Nb"]
    state = dict(scatR=[s[0] for s in scat], scatL=[s[1] for s in scat], scatG=[s[2] for s in scat],
                 plusR=plus[0], plusL=plus[1], plusG=plus[2], minusR=minus[0], minusL=minus[1],
                 minusG=minus[2], embR=emb[0], embL=emb[1], embG=emb[2],
                 GR=orc.field(NE, Nk, Nb), GL=orc.field(NE, Nk, Nb), GG=orc.field(NE, Nk, Nb),
                 A=orc.field(NE, Nk, Nb), scale=np.zeros((NE, Nk), order="F"))
    full = dict(prob, target=target, kernels=kernels, qK=[m["qK"] for m in prob["mechanisms"]],
                is_lo=[int(m["is_lo"]) for m in prob["mechanisms"]])
    return full, state, mu


def gates_pass(row, gates):
    return all(row[k] <= v for k, v in gates.items())
