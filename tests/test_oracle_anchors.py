"""Pins of the CPU oracle against the reference's own known answers.

Each test restates a case of the reference test suite (cited by file:line)
or a mathematical identity that a plausible mistake would break (a dropped
term, a wrong sign, a transposed operand). Tolerances are the reference's.
"""
import math

import numpy as np
import pytest

from conftest import rel

EPS = np.finfo(float).eps


def scalar_problem(eps, level, broadening):
    NE = len(eps)
    h = np.array([[[level]]], np.complex128, order="F")
    sigR = np.full((NE, 1, 1, 1), -0.5j * broadening, np.complex128, order="F")
    return h, sigR


def test_scalar_lorentzian(orc):
    # physics/scalar_dyson_and_spectral_identity.jl:5-14
    eps = np.linspace(-1.0, 1.0, 2001)
    h, sigR = scalar_problem(eps, 0.17, 0.08)
    GR, scale = orc.retarded_green(eps, h, sigR)
    A = orc.spectral(GR)
    exact = 0.08 / ((eps - 0.17) ** 2 + 0.04 ** 2)
    assert rel(A[:, 0, 0, 0].real, exact) < 2e-13
    assert np.max(np.abs(A.imag)) == 0.0
    # scale = |D| for a 1x1 block
    assert rel(scale[:, 0], np.abs(eps - 0.17 + 0.04j)) < 1e-15


def test_rational_two_by_two_inverse(orc):
    # physics/direct_static_retarded.jl:18-22: H=[0 1+i; 1-i 2], poles 1 +- sqrt(3)
    NE, Nk = 9, 2
    eps = np.linspace(-0.7, 1.3, NE)
    H = np.array([[0, 1 + 1j], [1 - 1j, 2]])
    h = np.zeros((Nk, 2, 2), np.complex128, order="F")
    sigR = np.zeros((NE, Nk, 2, 2), np.complex128, order="F")
    exact = np.zeros_like(sigR)
    for m in range(Nk):
        mm = m + 1
        h[m] = H + 0.17 * mm * np.eye(2)
        sigR[:, m, 0, 0] = sigR[:, m, 1, 1] = -(0.2 + 0.1 * mm) * 1j
        for e in range(NE):
            z = eps[e] - 0.17 * mm + (0.2 + 0.1 * mm) * 1j
            exact[e, m] = np.array([[z - 2, 1 + 1j], [1 - 1j, z]]) / (z * z - 2 * z - 2)
    GR, _ = orc.retarded_green(eps, h, sigR)
    assert rel(GR, exact) < 1e-14


@pytest.mark.parametrize("n", range(1, 10))
def test_small_inverse_backward_error(orc, n):
    # numerics/production_hotpaths.jl:7-50
    rng = np.random.default_rng(1082 + n)
    random = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    q, _ = np.linalg.qr(random)
    cases = [random + 0.1 * np.eye(n),
             q @ np.diag([1e-10] if n == 1 else np.linspace(1e-10, 1, n)) @ q.conj().T + 1e-12j * np.eye(n)]
    if n > 1:
        p = np.eye(n, dtype=np.complex128)
        p[[0, n - 1]] = p[[n - 1, 0]]
        cases.append(p)
    for a0 in cases:
        for s in (1e-150, 1.0, 1e150):
            a = a0 * s
            out = orc.invert_block(a)
            ref = np.linalg.inv(a)
            den = np.linalg.norm(a, 2) * np.linalg.norm(out, 2) + math.sqrt(n)
            assert np.linalg.norm(a @ out - np.eye(n), 2) / den < 32 * n * EPS
            c = np.linalg.cond(a)
            assert rel(out, ref) < 64 * n * EPS * c
            if c < 1e4:
                assert rel(out, ref) < 2e-13


@pytest.mark.parametrize("n", range(1, 10))
def test_small_inverse_singular(orc, n):
    with pytest.raises(orc.OracleError):
        orc.invert_block(np.zeros((n, n)))
    if n > 1:
        d = np.eye(n, dtype=np.complex128)
        d[-1] = d[0]
        with pytest.raises(orc.OracleError):
            orc.invert_block(d)


def test_singular_block_is_reported(orc):
    eps = np.array([0.0, 1.0])
    h = np.zeros((1, 2, 2), np.complex128, order="F")
    h[0] = [[1.0, 0.0], [0.0, 1.0]]  # D = (e-1) I singular at e = 1
    sigR = np.zeros((2, 1, 2, 2), np.complex128, order="F")
    rc, _, _ = orc.retarded_green_status(eps, h, sigR)
    assert rc == 2  # one-based linear block e + NE*m + 1


def test_kms_fdt(orc):
    # physics/analytic_equilibrium_sum_rule_lo_balance_and_zero_current_anchors.jl:5-28
    eps = np.linspace(-0.4, 0.4, 4001)
    h, sigR = scalar_problem(eps, 0.03, 0.07)
    GR, _ = orc.retarded_green(eps, h, sigR)
    A = orc.spectral(GR)[:, 0, 0, 0]
    f = 1.0 / (np.exp((eps + 0.02) / 0.025) + 1.0)
    GL = 1j * f * A
    GG = (GR[:, 0, 0, 0] - np.conj(GR[:, 0, 0, 0])) + GL  # greater_green (greens.jl:83-91)
    assert rel(GG, -1j * (1 - f) * A) < 3e-13
    inside = (f > 1e-7) & (f < 1 - 1e-7)
    ratio = GG[inside].imag / GL[inside].imag
    assert rel(ratio, -np.exp((eps[inside] + 0.02) / 0.025)) < 2e-12


def test_finite_window_sum_rule(orc):
    # analytic_equilibrium...jl:30-49
    eps = np.linspace(-10.0, 10.0, 40001)
    h, sigR = scalar_problem(eps, 0.17, 0.08)
    GR, _ = orc.retarded_green(eps, h, sigR)
    A = orc.spectral(GR)[:, 0, 0, 0].real
    w = np.full(len(eps), eps[1] - eps[0])
    w[[0, -1]] /= 2
    num = np.dot(w, A) / (2 * math.pi)
    exact = (math.atan(2 * (eps[-1] - 0.17) / 0.08) - math.atan(2 * (eps[0] - 0.17) / 0.08)) / math.pi
    assert abs(num - exact) < 1e-5 and 0 < num < 1


def test_non_periodic_shift_by_hand(orc):
    # numerics/non_periodic_energy_shift.jl:4-15
    eps = np.arange(0.0, 1.0001, 0.25)
    x = eps.astype(np.complex128).reshape(5, 1, 1, 1, order="F")
    plus = orc.apply_shift(orc.shift_plan(eps, 0.25), x)[:, 0, 0, 0]
    minus = orc.apply_shift(orc.shift_plan(eps, -0.25), x)[:, 0, 0, 0]
    assert np.allclose(plus[:4], eps[1:]) and plus[-1] == 0
    assert np.allclose(minus[1:], eps[:4]) and minus[0] == 0
    dense = orc.apply_shift_dense(orc.shift_matrix(eps, 0.25), x)[:, 0, 0, 0]
    assert np.allclose(dense, plus, rtol=0, atol=1e-15)


@pytest.mark.parametrize("delta", [0.0, 0.37, -0.37, 1.5, -2.0, 0.1234567])
def test_sparse_shift_equals_dense(orc, delta):
    # numerics/production_sparse_energy_shifts.jl:5-26
    rng = np.random.default_rng(7)
    eps = np.cumsum(rng.uniform(0.05, 0.2, 31)) - 1.5
    X = rng.normal(size=(31, 2, 2, 2)) + 1j * rng.normal(size=(31, 2, 2, 2))
    X = np.asfortranarray(X)
    sparse = orc.apply_shift(orc.shift_plan(eps, delta), X)
    dense = orc.apply_shift_dense(orc.shift_matrix(eps, delta), X)
    assert rel(sparse, dense) < 4e-15


def test_direct_pv_gaussian(orc):
    # numerics/direct_principal_value_transform.jl:4-14
    eps = np.round(np.arange(-2.0, 2.0001, 0.05), 12)
    w = np.full(len(eps), 0.05)
    w[[0, -1]] /= 2
    G = np.exp(-eps ** 2).astype(np.complex128).reshape(-1, 1, 1, 1, order="F")
    L = orc.direct_hilbert(np.asfortranarray(G), eps, w)[:, 0, 0, 0]
    c = int(np.argmin(np.abs(eps)))
    assert abs(L[c]) < 1e-14
    assert abs(L[c + 5].real + L[c - 5].real) <= 1e-12 * abs(L[c + 5].real)


def test_own_fft_matches_numpy(orc):
    rng = np.random.default_rng(3)
    for L in (2, 8, 64, 4096):
        x = rng.normal(size=L) + 1j * rng.normal(size=L)
        assert rel(orc.fft(x), np.fft.fft(x)) < 1e-15 * max(1, math.log2(L))
        assert rel(orc.fft(x, inverse=True), np.fft.ifft(x)) < 1e-15 * max(1, math.log2(L))


def test_fft_hilbert_equals_direct(orc):
    # numerics/fft_linear_hilbert_map_equals_direct_principal_value.jl:5-57
    rng = np.random.default_rng(0xFF71)
    NE, Nk, Nb = 37, 2, 2
    eps = np.linspace(-1.3, 0.9, NE)
    w = np.full(NE, eps[1] - eps[0])
    w[[0, -1]] /= 2
    G = np.asfortranarray(rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb)))
    direct = orc.direct_hilbert(G, eps, w)
    fft, rr = orc.fft_hilbert(G, eps, w)
    assert rel(fft, direct) < 2e-13
    assert math.isfinite(rr) and rr < 1e-12
    SL = np.asfortranarray(rng.normal(size=G.shape) + 1j * rng.normal(size=G.shape))
    SG = np.asfortranarray(rng.normal(size=G.shape) + 1j * rng.normal(size=G.shape))
    SR, Gam, _ = orc.retarded_self_energy_fft(SL, SG, eps, w)
    Gd = 1j * (SG - SL)
    assert np.array_equal(Gam, Gd)
    assert rel(SR, orc.direct_hilbert(np.asfortranarray(Gd), eps, w) - 0.5j * Gd) < 2e-13


def test_fft_rejects_nonuniform_grid(orc):
    eps = np.array([0.0, 0.1, 0.25, 0.3])
    G = np.zeros((4, 1, 1, 1), np.complex128, order="F")
    with pytest.raises(orc.OracleError):
        orc.fft_hilbert(G, eps, np.ones(4))


def _random_K6(rng, Nk, Nb):
    return np.asfortranarray(rng.normal(size=(Nk, Nk, Nb, Nb, Nb, Nb)) +
                             1j * rng.normal(size=(Nk, Nk, Nb, Nb, Nb, Nb)))


def _K6_from_flat(F, Nk, Nb):
    K = np.zeros((Nk, Nk, Nb, Nb, Nb, Nb), np.complex128, order="F")
    for m in range(Nk):
        for mp in range(Nk):
            for a in range(Nb):
                for b in range(Nb):
                    for c in range(Nb):
                        for d in range(Nb):
                            K[m, mp, a, c, d, b] = F[m + Nk * (a + Nb * b), mp + Nk * (c + Nb * d)]
    return K


def test_contractions_equal_literal(orc):
    # numerics/production_contractions_equal_literal_six_index_map.jl:5-26 and
    # literal_blas_and_low_rank_kernel_operators.jl:5-54
    rng = np.random.default_rng(11)
    NE, Nk, Nb = 7, 3, 2
    wk = rng.uniform(0.1, 1.0, Nk)
    G = np.asfortranarray(rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb)))
    K = _random_K6(rng, Nk, Nb)
    lit = orc.contract(orc.make_kernel("literal", K), G, wk, 0.7)
    dense = orc.contract(orc.make_kernel("dense", orc.kernel_dense_from_K6(K)), G, wk, 0.7, energy_chunk=3)
    assert rel(dense, lit) < 2e-14
    assert not orc.is_momentum_independent(K)
    # momentum independent
    Kmi = np.asfortranarray(np.broadcast_to(K[0:1, 0:1], K.shape).copy(order="F"))
    assert orc.is_momentum_independent(Kmi)
    lit = orc.contract(orc.make_kernel("literal", Kmi), G, wk, 0.7)
    mi = orc.contract(orc.make_kernel("momentum_independent", orc.kernel_mi_from_K6(Kmi)), G, wk, 0.7)
    assert rel(mi, lit) < 2e-14
    # low rank
    Nc = Nk * Nb * Nb
    left = rng.normal(size=(Nc, 3)) + 1j * rng.normal(size=(Nc, 3))
    radj = rng.normal(size=(3, Nc)) + 1j * rng.normal(size=(3, Nc))
    Klr = _K6_from_flat(left @ radj, Nk, Nb)
    lit = orc.contract(orc.make_kernel("literal", Klr), G, wk, 0.7)
    lr = orc.contract(orc.make_kernel("low_rank", left, radj, rank=3), G, wk, 0.7)
    assert rel(lr, lit) < 2e-13
    # separable IFR
    J = 2
    L_ = rng.normal(size=(Nb * Nb, J)) + 1j * rng.normal(size=(Nb * Nb, J))
    R_ = rng.normal(size=(J, Nb * Nb)) + 1j * rng.normal(size=(J, Nb * Nb))
    rad = rng.normal(size=(Nk, Nk)) + 0j
    F = np.zeros((Nc, Nc), np.complex128)
    for m in range(Nk):
        for mp in range(Nk):
            for q in range(Nb * Nb):
                for qp in range(Nb * Nb):
                    F[m + Nk * q, mp + Nk * qp] = rad[m, mp] * (L_[q] @ R_[:, qp])
    lit = orc.contract(orc.make_kernel("literal", _K6_from_flat(F, Nk, Nb)), G, wk, 0.7)
    ifr = orc.contract(orc.make_kernel("separable_ifr", L_, R_, rad, rank=J), G, wk, 0.7)
    assert rel(ifr, lit) < 2e-13
    # alpha/beta accumulation and beta = 0 ignoring the old content
    S = np.full_like(G, np.nan)
    out = orc.contract(orc.make_kernel("dense", orc.kernel_dense_from_K6(K)), G, wk, 0.7, alpha=2.0, beta=0.0, S=S)
    assert np.all(np.isfinite(out))
    acc = orc.contract(orc.make_kernel("dense", orc.kernel_dense_from_K6(K)), G, wk, 0.7, alpha=1.5, beta=1.0,
                       S=out.copy(order="F"))
    assert rel(acc, 3.5 / 2.0 * out) < 2e-14


def test_covariance_oracle(orc):
    # physics/direct_static_retarded.jl:46-57 (Sigma_V V G V^dagger with a momentum coefficient)
    rng = np.random.default_rng(5)
    NE, Nk, Nb = 5, 3, 3
    wk = rng.uniform(0.2, 1.0, Nk)
    V = [rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb)) for _ in range(2)]
    K = np.zeros((Nk, Nk, Nb, Nb, Nb, Nb), np.complex128, order="F")
    for m in range(Nk):
        for mp in range(Nk):
            coef = (m + 1 + mp + 1) / (Nk + 1)
            for Vh in V:
                K[m, mp] += coef * np.einsum("ac,bd->acdb", Vh, Vh.conj())
    G = np.asfortranarray(rng.normal(size=(NE, Nk, Nb, Nb)) + 1j * rng.normal(size=(NE, Nk, Nb, Nb)))
    expect = np.zeros_like(G)
    for e in range(NE):
        for m in range(Nk):
            for mp in range(Nk):
                coef = (m + 1 + mp + 1) / (Nk + 1)
                for Vh in V:
                    expect[e, m] += coef * wk[mp] * (Vh @ G[e, mp] @ Vh.conj().T)
    lit = orc.contract(orc.make_kernel("literal", K), G, wk, 1.0)
    dense = orc.contract(orc.make_kernel("dense", orc.kernel_dense_from_K6(K)), G, wk, 1.0)
    assert rel(lit, expect) < 1e-14 and rel(dense, expect) < 1e-14


def test_semicircle_root_through_contraction(orc):
    # physics/direct_static_retarded.jl:158-190
    s = 0.3
    Nk, Nb = 2, 2
    eps = np.linspace(-1.0, 1.0, 9)
    wk = np.array([0.4, 0.9])
    K = np.zeros((Nk, Nk, Nb, Nb, Nb, Nb), np.complex128, order="F")
    for a in range(Nb):
        for b in range(Nb):
            K[:, :, a, a, b, b] = s ** 2 / wk.sum()
    GR = np.zeros((9, Nk, Nb, Nb), np.complex128, order="F")
    roots = []
    for e, x in enumerate(eps):
        z = x + 0.15j
        g = 2 / (z + np.sqrt(z - 2 * s) * np.sqrt(z + 2 * s))
        assert g.imag < 0 and abs((z - s ** 2 * g) * g - 1) < 2e-13
        roots.append(g)
        for a in range(Nb):
            GR[e, :, a, a] = g
    SR = orc.contract(orc.make_kernel("momentum_independent", orc.kernel_mi_from_K6(K)), GR, wk, 1.0)
    for e, g in enumerate(roots):
        for a in range(Nb):
            assert abs(SR[e, 0, a, a] - s ** 2 * g) <= 2e-13 * abs(s ** 2 * g)


def test_paired_normalization_cone(orc):
    # numerics/paired_occupation_normalization.jl:6-40
    Gn = np.array([0.999999, 0.1])
    Gp = 1 - Gn
    for target in (1.11, 0.7, 1.099999):
        lesser = np.asfortranarray((1j * Gn).reshape(2, 1, 1, 1))
        greater = np.asfortranarray((-1j * Gp).reshape(2, 1, 1, 1))
        a, c, lam = orc.paired_coefficients(Gn.sum(), Gp.sum(), target)
        orc.normalize_pair(lesser, greater, a, c)
        assert abs(np.sum((-1j * lesser).real) - target) <= 4 * EPS * target
        assert np.min((-1j * lesser).real) >= 0 and np.min((1j * greater).real) >= 0
        assert np.max(np.abs(1j * (greater - lesser).ravel() - (Gn + Gp))) <= 4 * EPS
        if target == 1.11:
            assert lam < 1 and abs(lam - 1) > 1e-8
    with pytest.raises(orc.OracleError):
        orc.paired_coefficients(1.0, 0.1, 1.2)
    with pytest.raises(orc.OracleError):
        orc.paired_coefficients(1.0, -0.1, 0.8)


def test_number_functional_realness_gate(orc):
    NE, Nk = 3, 2
    wE, wk = np.ones(NE), np.ones(Nk)
    G = np.zeros((NE, Nk, 1, 1), np.complex128, order="F")
    G[:] = 1j
    assert abs(orc.number_functional(G, wE, wk, 2) - NE * Nk * 2 / (2 * math.pi)) < 1e-14
    G[:] = 1.0 + 1j  # -i*g_s/(2 pi)*(1+i) has an imaginary part
    with pytest.raises(orc.OracleError):
        orc.number_functional(G, wE, wk, 2)


def test_equilibrium_current_vanishes(orc):
    # analytic_equilibrium...jl:68-138 (zero-bias zero current): with one Fermi
    # function in Sigma and G the collision current cancels term by term.
    rng = np.random.default_rng(9)
    NE, Nk, Nb = 21, 3, 3
    eps = np.linspace(-1, 1, NE)
    wE = np.full(NE, eps[1] - eps[0])
    wk = rng.uniform(0.1, 1, Nk)
    f = 1 / (np.exp(eps / 0.1) + 1)
    A = np.zeros((NE, Nk, Nb, Nb), np.complex128, order="F")
    Gam = np.zeros_like(A)
    for e in range(NE):
        for m in range(Nk):
            X = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
            Y = rng.normal(size=(Nb, Nb)) + 1j * rng.normal(size=(Nb, Nb))
            A[e, m] = X @ X.conj().T
            Gam[e, m] = Y @ Y.conj().T
    F = f[:, None, None, None]
    GL, GG = 1j * F * A, -1j * (1 - F) * A
    SL, SG = 1j * F * Gam, -1j * (1 - F) * Gam
    J = orc.boundary_flux_bar(eps, wE, wk, 2, np.asfortranarray(SL), np.asfortranarray(SG),
                              np.asfortranarray(GL), np.asfortranarray(GG))
    scale = abs(orc.boundary_flux_bar(eps, wE, wk, 2, np.asfortranarray(SL), np.asfortranarray(SG),
                                      np.asfortranarray(GL), np.asfortranarray(np.zeros_like(GG))))
    assert scale > 1 and abs(J) < 2e-13 * scale
    # nonequilibrium: a shifted Fermi function in Sigma gives a nonzero current
    f2 = 1 / (np.exp((eps - 0.3) / 0.1) + 1)[:, None, None, None]
    J2 = orc.boundary_flux_bar(eps, wE, wk, 2, np.asfortranarray(1j * f2 * Gam),
                               np.asfortranarray(-1j * (1 - f2) * Gam), np.asfortranarray(GL),
                               np.asfortranarray(GG))
    assert abs(J2) > 1e-3 * scale


def test_seed_reproduces_the_target_number(orc):
    # seed.jl:151-197 / threaded_production_seed...jl:46-53
    from qclhip import synthetic as syn
    w = syn.gf_workload(N=3, NE=64, Nk=4, seed=3)
    GR, _ = orc.retarded_green(w["eps"], w["h"], None, eta=0.02)
    q = orc.seed_weights(GR, w["wE"], w["wk"], 2)
    target = 0.5 * q.sum()
    mu = orc.seed_mu(w["eps"], q, w["kbt"], target)
    GL = orc.seed_lesser(GR, w["eps"], mu, w["kbt"])
    assert abs(orc.number_functional(GL, w["wE"], w["wk"], 2) - target) < 1e-12 * target


def test_two_l_transform_from_two_l_point_halves():
    """The identity behind the device's 2L round-off witness: for a column of
    length NE and the kernel of lags -(NE-1)..NE-1 (both shorter than L), the
    2L cyclic convolution at n < L equals (y_n + exp(i pi n/L) y'_n)/2, with y
    the L-point convolution and y' the L-point convolution of
    x_n exp(-i pi n/L) with k_n exp(-i pi n/L)."""
    rng = np.random.default_rng(0)
    NE = 37
    L = 1 << int(np.ceil(np.log2(2 * NE - 1)))
    x = rng.normal(size=NE) + 1j * rng.normal(size=NE)
    k = rng.normal(size=2 * NE - 1)
    xp, kp = np.zeros(L, complex), np.zeros(L, complex)
    xp[:NE], kp[:2 * NE - 1] = x, k
    y2 = np.fft.ifft(np.fft.fft(np.r_[xp, np.zeros(L)]) * np.fft.fft(np.r_[kp, np.zeros(L)]))[:L]
    ph = np.exp(-1j * np.pi * np.arange(L) / L)
    y = np.fft.ifft(np.fft.fft(xp) * np.fft.fft(kp))
    yo = np.fft.ifft(np.fft.fft(xp * ph) * np.fft.fft(kp * ph))
    assert np.max(np.abs(0.5 * (y + np.conj(ph) * yo) - y2)) < 1e-13 * np.max(np.abs(y2))
    # on the extracted range the L result itself equals the 2L one
    e = np.arange(NE) + NE - 1
    assert np.max(np.abs(y[e] - y2[e])) < 1e-13 * np.max(np.abs(y2))


def landauer_chain(NE=41, Nk=3, seed=5, fLR=None):
    """Three-site chain with leads on sites 1 and 3 (the closed form of
    physics/cavity_current_and_spectral_oracles.jl:5-36), on an energy x
    momentum grid, in a basis rotated by a random unitary U so that every
    self-energy and Green block is dense. Returns the fields, the plus family
    (the left lead alone) and the closed-form current
    Jbar = g_s/(2 pi) sum_{e,m} wE wk (-T_em (fL_e - fR_e)),
    T = gL gR |t1 t2 / det D|^2, det D = a b c - t1^2 c - t2^2 a with
    a = E - l1 + i gL/2, b = E - l2, c = E - l3 + i gR/2 (levels shifted by
    the momentum's kinetic energy). The per-energy term follows from G< and
    G> = G^R S^{<,>} G^A: Re tr(S>_+ G< - S<_+ G>) = gL gR |G^R_13|^2 (fR - fL)."""
    rng = np.random.default_rng(seed)
    levels = np.array([0.1, -0.05, 0.2])
    t1, t2, gL, gR = 0.25, 0.32, 0.07, 0.11
    eps = np.linspace(-0.6, 0.7, NE)
    wE = rng.uniform(0.5, 1.5, NE) * (eps[1] - eps[0])
    kin = np.array([0.0, 0.03, 0.11])[:Nk]
    wk = rng.uniform(0.2, 1.0, Nk)
    if fLR is None:
        fL = 1 / (np.exp((eps - 0.15) / 0.05) + 1)
        fR = 1 / (np.exp((eps + 0.05) / 0.05) + 1)
    else:
        fL, fR = np.full(NE, fLR[0]), np.full(NE, fLR[1])
    U, _ = np.linalg.qr(rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
    rot = lambda X: U @ X @ U.conj().T
    h = np.zeros((Nk, 3, 3), np.complex128, order="F")
    for m in range(Nk):
        h[m] = rot(np.diag(levels + kin[m]) + np.diag([t1, t2], 1) + np.diag([t1, t2], -1))
    SR, SL, SG, PL, PG = (np.zeros((NE, Nk, 3, 3), np.complex128, order="F") for _ in range(5))
    J = 0.0
    for e in range(NE):
        sl = np.diag([1j * gL * fL[e], 0, 1j * gR * fR[e]])
        sg = np.diag([-1j * gL * (1 - fL[e]), 0, -1j * gR * (1 - fR[e])])
        pl = np.diag([1j * gL * fL[e], 0, 0])
        pg = np.diag([-1j * gL * (1 - fL[e]), 0, 0])
        for m in range(Nk):
            SR[e, m] = rot(np.diag([-0.5j * gL, 0, -0.5j * gR]))
            SL[e, m], SG[e, m], PL[e, m], PG[e, m] = rot(sl), rot(sg), rot(pl), rot(pg)
            E = eps[e] - kin[m]
            a, b, c = E - levels[0] + 0.5j * gL, E - levels[1], E - levels[2] + 0.5j * gR
            T = gL * gR * abs(t1 * t2 / (a * b * c - t1 ** 2 * c - t2 ** 2 * a)) ** 2
            assert 0 <= T <= 1 + 2e-14
            J += wE[e] * wk[m] * (-T * (fL[e] - fR[e]))
    return dict(eps=eps, wE=wE, wk=wk, h=h, SR=SR, SL=SL, SG=SG, plusL=PL, plusG=PG, J=2 * J / (2 * math.pi))


def test_current_equals_landauer_transmission(orc):
    # physics/cavity_current_and_spectral_oracles.jl:5-36 through the production
    # current functional (production_iteration.jl:174-201) with g_s = 2: a sign
    # flip, a lost g_s/(2 pi) or a transposed trace operand (the rotated blocks
    # are dense) breaks the equality.
    c = landauer_chain()
    GR, _ = orc.retarded_green(c["eps"], c["h"], c["SR"])
    GL, GG = orc.keldysh(GR, c["SL"]), orc.keldysh(GR, c["SG"])
    J = orc.boundary_flux_bar(c["eps"], c["wE"], c["wk"], 2, c["plusL"], c["plusG"], GL, GG)
    assert abs(c["J"]) > 1e-3
    assert abs(J - c["J"]) <= 2e-13 * abs(c["J"])
    # equal reservoir occupations: zero current to rounding
    c0 = landauer_chain(fLR=(0.4, 0.4))
    GR0, _ = orc.retarded_green(c0["eps"], c0["h"], c0["SR"])
    J0 = orc.boundary_flux_bar(c0["eps"], c0["wE"], c0["wk"], 2, c0["plusL"], c0["plusG"],
                               orc.keldysh(GR0, c0["SL"]), orc.keldysh(GR0, c0["SG"]))
    assert abs(J0) < 3e-14


def dawson_columns(n=961, half=12.0):
    """Gaussian columns c e^{-((x-x0)/s)^2} on a uniform grid with trapezoid
    weights, and the closed form of the discrete principal-value map
    Lambda_i = sum_{j != i} w_j Gamma_j / (2 pi (x_i - x_j)):
    (1/(2 pi)) PV int e^{-u^2}/(x - u) du = D(x)/sqrt(pi) (Dawson's integral,
    independent of the width s), plus the omitted-node term h Gamma'(x_i)/(2 pi)
    of the trapezoid rule on the regular part; the remaining error is spectral."""
    from scipy.special import dawsn
    x = np.linspace(-half, half, n)
    hstep = x[1] - x[0]
    w = np.full(n, hstep)
    w[[0, -1]] /= 2
    params = [(1.0, 0.0, 1.0), (0.5 - 0.25j, 0.7, 1.3), (-2.0j, -1.1, 0.8)]
    G = np.zeros((n, 1, len(params), 1), np.complex128, order="F")
    exact = np.zeros_like(G)
    for k, (c, x0, s) in enumerate(params):
        u = (x - x0) / s
        G[:, 0, k, 0] = c * np.exp(-u ** 2)
        exact[:, 0, k, 0] = c * (dawsn(u) / math.sqrt(math.pi) + hstep * (-2 * u / s) * np.exp(-u ** 2) / (2 * math.pi))
    return x, w, G, exact


@pytest.mark.parametrize("n", [481, 961, 1921])
def test_hilbert_sign_and_scale_dawson(orc, n):
    # Sign and scale of the Hilbert kernel 1/(2 pi (eps_e - eps_e')) of
    # fft_hilbert_transform (production_operators.jl:735-809) and of the direct
    # sum (reference/kernels.jl:467-510): a flipped sign or a lost 2 pi misses
    # Dawson's integral by O(1) at every grid spacing.
    x, w, G, exact = dawson_columns(n)
    L, r = orc.fft_hilbert(G, x, w)
    D = orc.direct_hilbert(G, x, w)
    scale = np.max(np.abs(exact))
    assert np.max(np.abs(L - exact)) < 1e-13 * scale
    # the O(NE^2) direct sum accumulates rounding of order n eps
    assert np.max(np.abs(D - exact)) < 1e-15 * n * scale
    assert r < 1e-12
