"""The fused FP64 Dyson + Keldysh kernel (QCLHIP_ENGINE_FUSED).

It runs the arithmetic of k_dyson and k_keldysh in one launch with G^R kept
in LDS, so every output must equal the two-launch vector engine bit for
bit: G^R, G<, G>, the block traces, the scale, A and the family totals of
current_green, on partial tiles and on the plane-strided layout; singular
blocks are reported with the same index; the oracle closes the loop.
"""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _case(hip, N, NE=45, Nk=3):
    rng = np.random.default_rng(500 + N)
    w = fx.syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=N)
    SR, SL, SG = fx.syn.diag_fields_host(w)
    d = [0.01 * (rng.normal(size=SR.shape) + 1j * rng.normal(size=SR.shape)) for _ in range(3)]
    sig = [hip.to_device(np.asfortranarray(x + y)) for x, y in zip((SR, SL, SG), d)]
    return w, sig


def _run(hip, g, w, sig):
    import torch
    NE, Nk = w["NE"], w["Nk"]
    traces = torch.empty(4 * NE * Nk, dtype=torch.float64, device="cuda")
    scale = torch.empty(NE * Nk, dtype=torch.float64, device="cuda")
    out = hip.dyson_keldysh(g, hip.h_to_device(w["h"]), *sig, traces=traces, scale=scale, check_status=True)
    return list(out) + [traces, scale]


@pytest.mark.parametrize("N", [1, 5, 8, 13, 24, 31, 40, 48])
def test_fused_equals_vector_engine(hip, N):
    import torch
    w, sig = _case(hip, N)
    ref = _run(hip, hip.Grid(w["eps"], w["wE"], w["wk"], N).set_engine("vector"), w, sig)
    got = _run(hip, hip.Grid(w["eps"], w["wE"], w["wk"], N).set_engine("fused"), w, sig)
    for name, a, b in zip(("GR", "GL", "GG", "traces", "scale"), ref, got):
        assert torch.equal(a, b), name


@pytest.mark.parametrize("N", [9, 24])
def test_fused_current_green_bitwise(hip, N):
    """current_green with three families, A and the family totals."""
    import torch
    w = fx.syn.gf_workload(N=N, NE=40, Nk=3)
    inj = fx.syn.injection_family(w)
    fams = [fx.syn.diag_fields_device(inj), fx.syn.diag_fields_device(w)]
    fams.append(tuple(0.37 * x for x in fams[1]))
    h = hip.h_to_device(w["h"])
    res = {}
    for engine in ("vector", "fused"):
        g = hip.Grid(w["eps"], w["wE"], w["wk"], N).set_engine(engine)
        tot = tuple(torch.empty(hip.field_shape(w["NE"], w["Nk"], N), dtype=torch.complex128, device="cuda")
                    for _ in range(3))
        st, info = hip.current_green_production(g, h, fams, 1.0, totals=tot)
        res[engine] = (st, info, tot)
    (s1, i1, t1), (s2, i2, t2) = res["vector"], res["fused"]
    for k in ("GR", "GL", "GG", "A"):
        assert torch.equal(getattr(s1, k), getattr(s2, k)), k
    assert i1 == i2
    for a, b in zip(t1, t2):
        assert torch.equal(a, b)


@pytest.mark.parametrize("N", [5, 24])
def test_fused_strided_equals_contiguous(hip, N):
    import torch
    NE, Nk = 21, 3
    w, sig = _case(hip, N, NE=NE, Nk=Nk)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N).set_engine("fused")
    h = hip.h_to_device(w["h"])
    GR, GL, GG = hip.dyson_keldysh(g, h, *sig, check_status=True)
    plane = NE * Nk + 37
    fields = [hip.strided_field(g, plane) for _ in range(6)]
    for f, x in zip(fields[:3], sig):
        f.copy_(x)
    for f in fields[3:]:
        f._base[:, NE * Nk:] = 1.5
    hip.dyson_keldysh_strided(g, plane, h, *fields)
    g.ctx.synchronize()
    for got, ref in zip(fields[3:], (GR, GL, GG)):
        assert torch.equal(got, ref)
        assert bool((got._base[:, NE * Nk:] == 1.5).all())


def test_fused_singular_block(hip):
    from qclhip import SingularException
    eps = np.array([0.5, 1.0, 2.0])
    g = hip.Grid(eps, np.ones(3), np.ones(2), 2).set_engine("fused")
    h = np.zeros((2, 2, 2), np.complex128)
    h[1] = np.eye(2)  # D = (e - 1) I vanishes at (e=1, m=1): one-based block 5
    z = hip.to_device(np.zeros((3, 2, 2, 2), np.complex128, order="F"))
    with pytest.raises(SingularException) as err:
        hip.dyson_keldysh(g, hip.h_to_device(h), z, z, z, check_status=True)
    assert err.value.block == 5


def test_fused_vs_oracle(hip, orc):
    N = 24
    w, sig = _case(hip, N, NE=37, Nk=2)
    SR, SL, SG = (hip.to_host(x) for x in sig)
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], SR)
    GL_o, GG_o = orc.keldysh(GR_o, SL), orc.keldysh(GR_o, SG)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N).set_engine("fused")
    GR, GL, GG, _, _ = _run(hip, g, w, sig)
    for got, ref in ((GR, GR_o), (GL, GL_o), (GG, GG_o)):
        assert rel(hip.to_host(got), ref) < 1e-12
