"""Momentum-independent candidates on one momentum's columns: the compact
contraction, the transform of those columns (each standing for Nk in the
round-off witness) and the broadcast give the full-field route's S<, S> and
S^R bit for bit (production_operators.jl:535-595, 842-884), and its witness
to rounding."""
import numpy as np
import pytest
import torch

import scba_fixture as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _fields(rng, NE, Nk, N, scale):
    return [np.asfortranarray(scale * (rng.normal(size=(NE, Nk, N, N)) + 1j * rng.normal(size=(NE, Nk, N, N))))
            for _ in range(2)]


@pytest.mark.parametrize("N,NE,Nk,path", [(5, 40, 3, "auto"), (24, 300, 4, "auto"), (8, 2100, 2, "hbm")])
def test_compact_candidate_equals_full(hip, N, NE, Nk, path):
    rng = np.random.default_rng(300 + N + NE)
    eps = np.linspace(-0.2, 0.4, NE)
    g = hip.Grid(eps, np.full(NE, eps[1] - eps[0]), rng.uniform(0.5, 1.5, Nk), N)
    K = hip.ProductionKernel(g, "momentum_independent", fx.syn.covariance_block(rng, N, 2, 0.04))
    GL, GG = (hip.to_device(x) for x in _fields(rng, NE, Nk, N, 0.05))
    full = hip.FFTHilbertPlan(g, g.ncols)
    comp = hip.FFTHilbertPlan(g, N * N).set_multiplicity(Nk)
    if path != "auto":
        full.set_path(path)
        comp.set_path(path)
    SL = hip.production_static_contraction(K, GL, qK=0.7)
    SG = hip.production_static_contraction(K, GG, qK=0.7)
    SR, _, r_full = hip.retarded_self_energy_fft(full, SL, SG, return_roundoff=True)
    YL = hip.momentum_independent_contraction_compact(K, GL, qK=0.7)
    YG = hip.momentum_independent_contraction_compact(K, GG, qK=0.7)
    YR, _, r_comp = hip.retarded_self_energy_fft(comp, YL, YG, return_roundoff=True)
    for Y, S in ((YL, SL), (YG, SG), (YR, SR)):
        B = hip.broadcast_momenta(g, Y)
        assert torch.equal(B, S)
    assert r_full > 0.0
    assert abs(r_comp - r_full) <= 1e-12 * r_full


def test_compact_contraction_argument_errors(hip):
    from qclhip import _lib
    rng = np.random.default_rng(7)
    N, NE, Nk = 4, 16, 2
    eps = np.linspace(0.0, 0.3, NE)
    g = hip.Grid(eps, np.full(NE, eps[1] - eps[0]), np.ones(Nk), N)
    K = hip.ProductionKernel(g, "momentum_independent", fx.syn.covariance_block(rng, N, 2, 0.04))
    G = hip.to_device(_fields(rng, NE, Nk, N, 0.05)[0])
    with pytest.raises(_lib.ArgumentError):  # the adjoint form scales the broadcast by momentum
        hip.momentum_independent_contraction_compact(K.adjoint(), G)
    l, r = fx.syn.low_rank_factors(rng, N, Nk, r=2)
    with pytest.raises(_lib.ArgumentError):
        hip.momentum_independent_contraction_compact(hip.ProductionKernel(g, "low_rank", l, r, rank=2), G)
    with pytest.raises(_lib.DimensionMismatch):
        hip.broadcast_momenta(g, G)


def test_compact_mix_equals_broadcast_mix(hip):
    """A mixing pass that reads one-momentum candidates for every momentum
    leaves the same mixed fields, accumulated totals, candidate sum and
    ratios, bit for bit, as the pass over their broadcast fields."""
    rng = np.random.default_rng(11)
    N, NE, Nk = 6, 50, 3
    eps = np.linspace(-0.2, 0.4, NE)
    g = hip.Grid(eps, np.full(NE, eps[1] - eps[0]), np.ones(Nk), N)
    field = lambda: hip.to_device(_fields(rng, NE, Nk, N, 0.05)[0])
    olds = [field() for _ in range(3)]
    Ys = [hip.to_device(_fields(rng, NE, 1, N, 0.05)[0]) for _ in range(3)]
    tots = [field() for _ in range(4)]

    def run(compact):
        o = [x.clone() for x in olds]
        t = [x.clone() for x in tots]
        c = Ys if compact else [hip.broadcast_momenta(g, Y) for Y in Ys]
        spec = [(t[0], [0], "accumulate"), (t[1], [1], "accumulate"), (t[2], [2], "accumulate"),
                (t[3], [1], "accumulate,candidate")]
        r = hip.mix_fields_residual(list(zip(o, c)), 0.3, [-1, -1, -1], totals=spec,
                                    compact=[1, 1, 1] if compact else None, grid=g)
        return o, t, r

    o1, t1, r1 = run(True)
    o2, t2, r2 = run(False)
    for a, b in zip(o1 + t1, o2 + t2):
        assert torch.equal(a, b)
    assert r1 == r2
