"""FP64 matrix-core Keldysh engine and the FP32 + FP64-refinement GF solve.

The matrix engine must agree with the oracle to the FP64 bar (1e-12). The
mixed-precision Dyson step is held to: G^R within 1e-9 of the oracle (two
Newton steps, guard 1e-10), G</G> within 1e-9 with FP64 products and 2e-6
with FP32 products, ill-conditioned blocks re-solved in FP64 (oracle parity
at FP64 level), singular blocks reported as in FP64, and - with FP64
products - the current functional within 1e-8 relative of the FP64 path on
a driven (non-equilibrium) workload.
"""
import numpy as np
import pytest

from conftest import rel
import scba_fixture as fx

pytestmark = pytest.mark.gpu

TOL = 1e-12
GR_TOL_MIXED = 1e-9
KELDYSH_TOL = {"fp32_refine": 1e-9, "fp32_products": 2e-6}
CURRENT_TOL = 1e-8

SIZES = [1, 3, 8, 12, 17, 24, 31, 32, 40, 48]


@pytest.fixture(scope="module")
def hip():
    import torch
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from qclhip import api
    api.context(0)
    return api


def _dense_case(Nb, NE=24, Nk=3):
    rng = np.random.default_rng(200 + Nb)
    w = fx.syn.gf_workload(N=Nb, NE=NE, Nk=Nk, seed=Nb)
    SR, SL, SG = fx.syn.diag_fields_host(w)
    d = [np.asfortranarray(0.01 * (rng.normal(size=SR.shape) + 1j * rng.normal(size=SR.shape))) for _ in range(3)]
    return w, np.asfortranarray(SR + d[0]), np.asfortranarray(SL + d[1]), np.asfortranarray(SG + d[2])


def _run_dk(hip, g, w, SR, SL, SG):
    import torch
    NE, Nk = w["NE"], w["Nk"]
    traces = torch.empty(2 * NE * Nk * 2, dtype=torch.float64, device="cuda")
    GR, GL, GG = hip.dyson_keldysh(g, hip.h_to_device(w["h"]), hip.to_device(SR), hip.to_device(SL),
                                   hip.to_device(SG), traces=traces, check_status=True)
    tr = traces.cpu().numpy().view(np.complex128).reshape(2, Nk, NE)
    return hip.to_host(GR), hip.to_host(GL), hip.to_host(GG), tr


@pytest.mark.parametrize("Nb", SIZES)
def test_matrix_engine_fp64(hip, orc, Nb):
    w, SR, SL, SG = _dense_case(Nb)
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], SR)
    GL_o, GG_o = orc.keldysh(GR_o, SL), orc.keldysh(GR_o, SG)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], Nb).set_engine("matrix")
    GR, GL, GG, tr = _run_dk(hip, g, w, SR, SL, SG)
    for got, ref in ((GR, GR_o), (GL, GL_o), (GG, GG_o), (tr[0].T, np.einsum("emaa->em", GL_o)),
                     (tr[1].T, np.einsum("emaa->em", GG_o))):
        err = float(rel(got, ref))
        assert err < TOL


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_products"])
@pytest.mark.parametrize("Nb", SIZES)
def test_mixed_dyson_keldysh(hip, orc, Nb, precision):
    w, SR, SL, SG = _dense_case(Nb)
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], SR)
    GL_o, GG_o = orc.keldysh(GR_o, SL), orc.keldysh(GR_o, SG)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], Nb).set_precision(precision, refine_steps=2)
    GR, GL, GG, tr = _run_dk(hip, g, w, SR, SL, SG)
    err = float(rel(GR, GR_o))
    assert err < GR_TOL_MIXED
    for got, ref in ((GL, GL_o), (GG, GG_o), (tr[0].T, np.einsum("emaa->em", GL_o)),
                     (tr[1].T, np.einsum("emaa->em", GG_o))):
        err = float(rel(got, ref))
        assert err < KELDYSH_TOL[precision]


def test_mixed_without_refinement_is_fp32(hip, orc):
    """refine_steps=0 leaves the FP32 inverse: error well above the refined one,
    so the refinement steps are what brings G^R to 1e-9."""
    w, SR, SL, SG = _dense_case(24)
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], SR)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 24).set_precision("fp32_refine", refine_steps=0)
    GR, _ = hip.retarded_green_production(g, hip.h_to_device(w["h"]), hip.to_device(SR))
    err0 = float(rel(hip.to_host(GR), GR_o))
    assert 1e-8 < err0 < 1e-3
    g.set_precision("fp32_refine", refine_steps=1, refine_tol=1.0)
    GR, _ = hip.retarded_green_production(g, hip.h_to_device(w["h"]), hip.to_device(SR))
    err1 = float(rel(hip.to_host(GR), GR_o))
    assert err1 < 1e-2 * err0


def test_mixed_ill_conditioned_falls_back(hip, orc):
    """Blocks with cond ~ 1e6 exceed the refinement guard (FP32 elimination
    error ~ cond * 6e-8) and are re-solved in FP64: they match the oracle to
    ~cond * 1e-16; the well-conditioned blocks stay on the mixed path."""
    rng = np.random.default_rng(11)
    N, NE, Nk = 12, 16, 2
    Q, _ = np.linalg.qr(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    eps = np.linspace(-1.5, 1.5, NE)
    lam = np.linspace(-1.05, 1.05, N)
    lam[5] = eps[7]  # an eigenvalue on a grid energy
    h = np.zeros((Nk, N, N), np.complex128, order="F")
    for m in range(Nk):
        h[m] = Q @ np.diag(lam) @ Q.conj().T
    SR = np.zeros((NE, Nk, N, N), np.complex128, order="F")
    for a in range(N):
        SR[:, :, a, a] = -0.05j
        SR[7, :, a, a] = -2.5e-6j  # D(e=7) has cond ~ 1e6
    g = hip.Grid(eps, np.ones(NE), np.ones(Nk), N).set_precision("fp32_refine", refine_steps=2)
    GR, _ = hip.retarded_green_production(g, hip.h_to_device(h), hip.to_device(SR))
    GR = hip.to_host(GR)
    GR_o, _ = orc.retarded_green(eps, h, SR)
    err = float(rel(GR[7], GR_o[7]))
    assert err < 1e-8
    assert g.refine_fallbacks() == Nk  # exactly the e=7 blocks
    others = [e for e in range(NE) if e != 7]
    err = float(rel(GR[others], GR_o[others]))
    assert err < GR_TOL_MIXED


def test_mixed_singular_block_raises(hip, orc):
    from qclhip import SingularException
    eps = np.array([0.5, 1.0, 2.0])
    g = hip.Grid(eps, np.ones(3), np.ones(2), 2).set_precision("fp32_refine")
    h = np.zeros((2, 2, 2), np.complex128)
    h[1] = np.eye(2)
    with pytest.raises(SingularException) as err:
        hip.retarded_green_production(g, hip.h_to_device(h), None)
    assert err.value.block == 5


@pytest.mark.parametrize("precision", ["fp32_refine", "fp32_products"])
def test_mixed_current_green_vs_oracle(hip, orc, precision):
    prob = fx.make_problem(N=5, NE=48, Nk=3)
    full, state, green = fx.oracle_state(prob)
    fams_o = ([state["embR"]] + state["scatR"], [state["embL"]] + state["scatL"], [state["embG"]] + state["scatG"])
    ref = orc.current_green(prob["eps"], prob["h"], *fams_o, prob["wE"], prob["wk"], 2, full["target"])
    g = hip.Grid(prob["eps"], prob["wE"], prob["wk"], prob["Nb"]).set_precision(precision)
    fams = [tuple(hip.to_device(x) for x in (R, L, G)) for R, L, G in zip(*fams_o)]
    st, info = hip.current_green_production(g, hip.h_to_device(prob["h"]), fams, full["target"])
    err = float(rel(hip.to_host(st.GR), ref["GR"]))
    assert err < GR_TOL_MIXED
    tol = KELDYSH_TOL[precision]
    for k in ("GL", "GG"):
        err = float(rel(hip.to_host(getattr(st, k)), ref[k]))
        assert err < tol, k
    for k in ("lam", "number", "holes"):
        assert abs(info[k] - ref[k]) <= tol * abs(ref[k]), k


def _driven_current(hip, w, inj, precision):
    g = hip.Grid(w["eps"], w["wE"], w["wk"], w["N"])
    if precision != "fp64":
        g.set_precision(precision)
    plus = fx.syn.diag_fields_device(inj)
    scat = fx.syn.diag_fields_device(w)
    st, info = hip.current_green_production(g, hip.h_to_device(w["h"]), [plus, scat], target=1.0)
    J = hip.boundary_flux_bar(g, plus[1], plus[2], st.GL, st.GG)
    return J, st, info


def test_mixed_current_error_driven(hip, orc):
    """Relative current error against the FP64 path on the synthetic workload
    (N=24, NE=512, Nk=16) driven by a reservoir at mu+50 meV: FP32 + refinement
    meets 1e-8; FP32 Keldysh products stay at the FP32 level (< 1e-6)."""
    w = fx.syn.gf_workload(N=24, NE=512, Nk=16)
    inj = fx.syn.injection_family(w)
    J64, st64, info64 = _driven_current(hip, w, inj, "fp64")
    assert abs(J64) > 1e-6
    J32, st32, info32 = _driven_current(hip, w, inj, "fp32_refine")
    assert abs(J32 - J64) <= CURRENT_TOL * abs(J64)
    assert abs(info32["lam"] - info64["lam"]) <= CURRENT_TOL * abs(info64["lam"])
    err = float(rel(st32.GR.cpu().numpy(), st64.GR.cpu().numpy()))
    assert err < GR_TOL_MIXED
    del st32
    Jp, stp, infop = _driven_current(hip, w, inj, "fp32_products")
    assert abs(Jp - J64) <= 1e-6 * abs(J64)


def test_autotune_selects_engine_and_keeps_parity(hip, orc):
    w, SR, SL, SG = _dense_case(24, NE=64, Nk=4)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], 24)
    t = g.autotune()
    assert t["vector_ms"] > 0 and t["matrix_ms"] > 0 and t["fused_ms"] > 0
    assert t["engine"] in ("vector", "matrix", "fused")
    g2 = hip.Grid(w["eps"], w["wE"], w["wk"], 24)
    assert g2.autotune() == t  # cached per shape in the context
    GR_o, _ = orc.retarded_green(w["eps"], w["h"], SR)
    GL_o = orc.keldysh(GR_o, SL)
    GR, GL, GG, tr = _run_dk(hip, g, w, SR, SL, SG)
    err = float(rel(GL, GL_o))
    assert err < TOL


@pytest.mark.parametrize("N", [8, 24, 32])
def test_tile_shapes_bitwise_and_autotuned(hip, N):
    """Every tile shape of the FP64 vector kernels (Dyson 4 or 8 energies
    per workgroup, Keldysh 2 or 4) gives the same bits; the autotuner's
    record lists the timed shapes and its choice is one of them."""
    import torch
    from qclhip import ArgumentError
    w, SR, SL, SG = _dense_case(N, NE=45, Nk=3)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N)
    ref = _run_dk(hip, g, w, SR, SL, SG)
    ran = 0
    for wd in (4, 8):
        for wk in (2, 4):
            try:
                g.set_tiles(wd, wk)
            except ArgumentError:  # not compiled or over the LDS of a CU at this N
                continue
            out = _run_dk(hip, g, w, SR, SL, SG)
            ran += 1
            for a, b in zip(ref, out):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (wd, wk)
    assert ran >= 1
    t = hip.Grid(w["eps"], w["wE"], w["wk"], N).autotune()
    assert t["dyson_waves"] in (4, 8) and t["keldysh_waves"] in (2, 4)
    assert t["dyson_w4_ms"] > 0 and t["keldysh_w2_ms"] > 0


@pytest.mark.parametrize("N", [25, 32, 37, 40, 44, 48])
def test_keldysh_split_bitwise(hip, N):
    """The vector Keldysh kernel with a block's rows split over S waves
    gives the one-wave results bit for bit: G<, G>, the block traces (the
    same summation tree) and, through current_green, A and the totals."""
    import torch
    from qclhip import ArgumentError
    w, SR, SL, SG = _dense_case(N, NE=45, Nk=3)
    g = hip.Grid(w["eps"], w["wE"], w["wk"], N)
    ref = _run_dk(hip, g, w, SR, SL, SG)
    # physical families (the realness gates of current_green hold)
    w2 = fx.syn.gf_workload(N=N, NE=40, Nk=3)
    g2 = hip.Grid(w2["eps"], w2["wE"], w2["wk"], N)
    fams = [fx.syn.diag_fields_device(fx.syn.injection_family(w2)), fx.syn.diag_fields_device(w2)]
    h = hip.h_to_device(w2["h"])

    def cg():
        tot = tuple(torch.empty(hip.field_shape(w2["NE"], w2["Nk"], N), dtype=torch.complex128, device="cuda")
                    for _ in range(3))
        st, info = hip.current_green_production(g2, h, fams, 1.0, totals=tot)
        return [st.GR, st.GL, st.GG, st.A, *tot], info

    cref, iref = cg()
    ran = []
    for ks in (2, 3, 4, 5, 6):
        try:
            g.set_keldysh_split(ks)
        except ArgumentError:  # not compiled for this basis size
            continue
        g2.set_keldysh_split(ks)
        ran.append(ks)
        out = _run_dk(hip, g, w, SR, SL, SG)
        for a, b in zip(ref, out):
            assert np.array_equal(np.asarray(a), np.asarray(b)), ks
        for wk in (1, 2, 4):  # energies per workgroup of the split kernel
            try:
                g.set_tiles(0, wk)
            except ArgumentError:
                continue
            out = _run_dk(hip, g, w, SR, SL, SG)
            for a, b in zip(ref, out):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (ks, wk)
        g.set_tiles(0, 0)
        cout, iout = cg()
        for a, b in zip(cref, cout):
            assert torch.equal(a, b), ks
        assert iout == iref
    assert ran
    g.set_keldysh_split(0)
    with pytest.raises(ArgumentError):
        g.set_keldysh_split(7)
    t = hip.Grid(w["eps"], w["wE"], w["wk"], N).autotune()
    assert t["keldysh_split"] in [1] + ran and t["keldysh_split_ms"] > 0
