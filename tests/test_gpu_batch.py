"""Batched GF solve over uneven instances (own N_E, N_k, N_b per instance):
every instance equals the oracle (FP64 bar, 1e-12) and the single-instance
entry point, across tile classes and with ragged energy tiles."""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _instance(N, NE, Nk, seed):
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=seed)
    inj = fx.syn.injection_family(w, seed=seed + 100)
    rng = np.random.default_rng(seed)
    scat = list(fx.syn.diag_fields_host(w))
    d = 0.005 * (rng.normal(size=scat[0].shape) + 1j * rng.normal(size=scat[0].shape))
    scat[0] = np.asfortranarray(scat[0] + d)  # dense Sigma^R
    return w, list(fx.syn.diag_fields_host(inj)), scat


def test_batch_uneven_vs_oracle_and_single(hip, orc):
    import torch
    shapes = [(3, 24, 2), (5, 33, 3), (9, 48, 2), (12, 25, 3), (17, 40, 2), (24, 31, 3), (31, 24, 2), (40, 17, 2)]
    insts = [_instance(N, NE, Nk, 300 + i) for i, (N, NE, Nk) in enumerate(shapes)]
    grids = [hip.Grid(w["eps"], w["wE"], w["wk"], w["N"]) for w, _, _ in insts]
    batch = hip.Batch(grids)
    hs = [hip.h_to_device(w["h"]) for w, _, _ in insts]
    fams = [[tuple(hip.to_device(x) for x in plus), tuple(hip.to_device(x) for x in scat)]
            for _, plus, scat in insts]
    # targets between the raw number and number + holes exercise both branches
    targets = []
    for g, h, f in zip(grids, hs, fams):
        _, info = hip.current_green_production(g, h, f, 1.0, spectral=False)
        targets.append(info["number"] * (0.9 if len(targets) % 2 == 0 else 1.02))
    states, infos = hip.batch_current_green(batch, hs, fams, targets)
    for i, ((w, plus, scat), g) in enumerate(zip(insts, grids)):
        ref = orc.current_green(w["eps"], w["h"], [plus[0], scat[0]], [plus[1], scat[1]], [plus[2], scat[2]],
                                w["wE"], w["wk"], 2, targets[i])
        for k in ("GR", "GL", "GG"):
            err = float(rel(hip.to_host(getattr(states[i], k)), ref[k]))
            assert err < TOL, (i, k)
        for k in ("lam", "number", "holes"):
            assert abs(infos[i][k] - ref[k]) <= TOL * abs(ref[k]), (i, k)
        single, sinfo = hip.current_green_production(g, hs[i], fams[i], targets[i], spectral=False)
        for k in ("GR", "GL", "GG"):
            err = float(torch.max(torch.abs(getattr(single, k) - getattr(states[i], k))))
            assert err <= 1e-15 * float(torch.max(torch.abs(getattr(single, k)))), (i, k)


def test_batch_medium_uneven_vs_single(hip):
    """Sizes of the uneven sweep (N in 12..48, N_E up to 1024) at N_k = 4:
    the batch equals the single-instance calls."""
    import torch
    rng = np.random.default_rng(77)
    shapes = [(int(rng.integers(12, 49)), int(rng.choice([256, 512, 777, 1024])), 4) for _ in range(10)]
    insts = [_instance(N, NE, Nk, 500 + i) for i, (N, NE, Nk) in enumerate(shapes)]
    grids = [hip.Grid(w["eps"], w["wE"], w["wk"], w["N"]) for w, _, _ in insts]
    batch = hip.Batch(grids)
    hs = [hip.h_to_device(w["h"]) for w, _, _ in insts]
    fams = [[tuple(hip.to_device(x) for x in plus), tuple(hip.to_device(x) for x in scat)]
            for _, plus, scat in insts]
    targets = [1.0] * len(insts)
    states, infos = hip.batch_current_green(batch, hs, fams, targets)
    for i, g in enumerate(grids):
        single, sinfo = hip.current_green_production(g, hs[i], fams[i], targets[i], spectral=False)
        assert abs(sinfo["lam"] - infos[i]["lam"]) <= 1e-14 * abs(sinfo["lam"])
        for k in ("GR", "GL", "GG"):
            err = float(torch.max(torch.abs(getattr(single, k) - getattr(states[i], k))))
            assert err <= 1e-14 * float(torch.max(torch.abs(getattr(single, k)))), (shapes[i], k)


def test_batch_argument_errors(hip):
    from qclhip import ArgumentError
    w1, p1, s1 = _instance(4, 16, 2, 1)
    w2, p2, s2 = _instance(6, 20, 2, 2)
    grids = [hip.Grid(w["eps"], w["wE"], w["wk"], w["N"]) for w in (w1, w2)]
    batch = hip.Batch(grids)
    hs = [hip.h_to_device(w["h"]) for w in (w1, w2)]
    f1 = [tuple(hip.to_device(x) for x in p1), tuple(hip.to_device(x) for x in s1)]
    f2 = [tuple(hip.to_device(x) for x in p2)]
    with pytest.raises(ArgumentError):
        hip.batch_current_green(batch, hs, [f1, f2], [1.0, 1.0])
    f2 = [tuple(hip.to_device(x) for x in p2), tuple(hip.to_device(x) for x in s2)]
    with pytest.raises(ArgumentError):
        hip.batch_current_green(batch, hs, [f1, f2], [1.0, -1.0])
