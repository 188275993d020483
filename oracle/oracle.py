"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP backend is compared against.
Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module; the product package never does.

Fields are numpy complex128 arrays of shape (NE, Nk, Nb, Nb) in Fortran
(column-major) order, i.e. exactly the memory layout of the reference's
Array{ComplexF64,4}. All indices are 0-based.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_up = ct.POINTER(ct.c_ubyte)


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle returned {code}")
        self.code = code


class Kernel(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("rank", ct.c_int), ("a", _dp), ("b", _dp), ("c", _dp)]


class Problem(ct.Structure):
    _fields_ = [
        ("NE", ct.c_int), ("Nk", ct.c_int), ("Nb", ct.c_int), ("nmech", ct.c_int),
        ("gs", ct.c_int), ("energy_chunk", ct.c_int),
        ("eps", _dp), ("wE", _dp), ("wk", _dp), ("h", _dp), ("Tplus", _dp), ("Tminus", _dp),
        ("Ep", ct.c_double), ("hwLO", ct.c_double), ("NLO", ct.c_double),
        ("target", ct.c_double), ("alpha_mix", ct.c_double),
        ("kernels", ct.POINTER(Kernel)), ("qK", _dp), ("is_lo", _ip),
    ]


class State(ct.Structure):
    _fields_ = [
        ("scatR", ct.POINTER(_dp)), ("scatL", ct.POINTER(_dp)), ("scatG", ct.POINTER(_dp)),
        ("plusR", _dp), ("plusL", _dp), ("plusG", _dp),
        ("minusR", _dp), ("minusL", _dp), ("minusG", _dp),
        ("embR", _dp), ("embL", _dp), ("embG", _dp),
        ("GR", _dp), ("GL", _dp), ("GG", _dp), ("A", _dp), ("scale", _dp),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (make rebuilds it
    when a source is newer than the library)."""
    subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ct.CDLL(_LIB_PATH)
        _lib.orc_number_functional.restype = ct.c_double
        _lib.orc_boundary_flux_bar.restype = ct.c_double
        _lib.orc_fermi.restype = ct.c_double
    return _lib


def P(x):
    """Raw double pointer to a numpy array (complex arrays are interleaved)."""
    if x is None:
        return ct.cast(None, _dp)
    assert x.flags.f_contiguous or x.flags.c_contiguous
    return x.ctypes.data_as(_dp)


def field(NE, Nk, Nb, dtype=np.complex128):
    return np.zeros((NE, Nk, Nb, Nb), dtype=dtype, order="F")


def _check(rc, what):
    if rc != 0:
        raise OracleError(rc, what)


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def retarded_green(eps, h, sigR, eta=0.0):
    NE, (Nk, Nb) = len(eps), h.shape[:2]
    GR = field(NE, Nk, Nb)
    scale = np.zeros((NE, Nk), order="F")
    rc = lib().orc_retarded_green(NE, Nk, Nb, P(eps), P(h), P(sigR), ct.c_double(eta), P(GR), P(scale))
    _check(rc, "retarded_green")
    return GR, scale


def retarded_green_status(eps, h, sigR, eta=0.0):
    NE, Nk, Nb, _ = sigR.shape
    GR = field(NE, Nk, Nb)
    scale = np.zeros((NE, Nk), order="F")
    rc = lib().orc_retarded_green(NE, Nk, Nb, P(eps), P(h), P(sigR), ct.c_double(eta), P(GR), P(scale))
    return rc, GR, scale


def invert_block(A):
    X = np.asfortranarray(A.astype(np.complex128))
    rc = lib().orc_invert_block(X.shape[0], P(X))
    _check(rc, "invert_block")
    return X


def spectral(GR):
    NE, Nk, Nb, _ = GR.shape
    A = field(NE, Nk, Nb)
    lib().orc_spectral(NE, Nk, Nb, P(GR), P(A))
    return A


def keldysh(GR, S):
    NE, Nk, Nb, _ = GR.shape
    out = field(NE, Nk, Nb)
    lib().orc_keldysh(NE, Nk, Nb, P(GR), P(S), P(out))
    return out


def number_functional(G, wE, wk, gs):
    NE, Nk, Nb, _ = G.shape
    st = ct.c_int(0)
    v = lib().orc_number_functional(NE, Nk, Nb, P(G), P(wE), P(wk), gs, ct.byref(st))
    _check(st.value, "number_functional")
    return v


def paired_coefficients(number, holes, target):
    a, c, lam = ct.c_double(), ct.c_double(), ct.c_double()
    rc = lib().orc_paired_coefficients(ct.c_double(number), ct.c_double(holes), ct.c_double(target),
                                       ct.byref(a), ct.byref(c), ct.byref(lam))
    _check(rc, "paired_coefficients")
    return a.value, c.value, lam.value


def normalize_pair(GL, GG, a, c):
    lib().orc_normalize_pair(ct.c_longlong(GL.size), P(GL), P(GG), ct.c_double(a), ct.c_double(c))


def current_green(eps, h, famR, famL, famG, wE, wk, gs, target):
    """families: lists whose index 0 is the embedding total."""
    NE, Nk, Nb, _ = famR[0].shape
    nf = len(famR)
    arr = lambda fs: (_dp * nf)(*[P(f) for f in fs])
    GR, GL, GG, A = (field(NE, Nk, Nb) for _ in range(4))
    scale = np.zeros((NE, Nk), order="F")
    out3 = np.zeros(3)
    rc = lib().orc_current_green(NE, Nk, Nb, P(eps), P(h), nf, arr(famR), arr(famL), arr(famG),
                                 P(wE), P(wk), gs, ct.c_double(target), P(GR), P(GL), P(GG),
                                 P(A), P(scale), P(out3))
    _check(rc, "current_green")
    return dict(GR=GR, GL=GL, GG=GG, A=A, scale=scale, lam=out3[0], number=out3[1], holes=out3[2])


def current_green_scalar_lesser(eps, h, famR, famL, famG, wE, wk, gs, target):
    """The occupation_normalization = :scalar_lesser branch of
    _current_green_production (production_state.jl:54-61), composed from the
    oracle's Dyson, Keldysh product and number functional: G< = G<raw/lambda,
    G> = G^R - G^R^dagger + G< (production_operators.jl:1387-1415)."""
    SR = sum(famR[1:], famR[0].copy())
    SL = sum(famL[1:], famL[0].copy())
    GR, scale = retarded_green(eps, h, np.asfortranarray(SR))
    GL = keldysh(GR, np.asfortranarray(SL))
    number = number_functional(GL, wE, wk, gs)
    lam = number / target
    if not (np.isfinite(lam) and lam > 0):
        raise ValueError("kinetic eigenvalue lambda must be positive")
    GL = np.asfortranarray(GL / lam)
    A = spectral(GR)
    GG = np.asfortranarray(GR - np.conj(np.transpose(GR, (0, 1, 3, 2))) + GL)
    return dict(GR=GR, GL=GL, GG=GG, A=A, scale=scale, lam=lam, number=number)


def shift_plan(eps, delta):
    NE = len(eps)
    left = np.zeros(NE, np.int32)
    right = np.zeros(NE, np.int32)
    wr = np.zeros(NE)
    valid = np.zeros(NE, np.uint8)
    rc = lib().orc_shift_plan(NE, P(eps), ct.c_double(delta), left.ctypes.data_as(_ip),
                              right.ctypes.data_as(_ip), P(wr), valid.ctypes.data_as(_up))
    _check(rc, "shift_plan")
    return left, right, wr, valid


def apply_shift(plan, X):
    left, right, wr, valid = plan
    Y = np.zeros_like(X, order="F")
    NE = X.shape[0]
    rc = lib().orc_apply_shift(NE, ct.c_longlong(X.size // NE), left.ctypes.data_as(_ip),
                               right.ctypes.data_as(_ip), P(wr), valid.ctypes.data_as(_up), P(X), P(Y))
    _check(rc, "apply_shift")
    return Y


def shift_matrix(eps, delta):
    NE = len(eps)
    W = np.zeros((NE, NE), order="F")
    _check(lib().orc_shift_matrix(NE, P(eps), ct.c_double(delta), P(W)), "shift_matrix")
    return W


def apply_shift_dense(W, X):
    Y = np.zeros_like(X, order="F")
    NE = X.shape[0]
    lib().orc_apply_shift_dense(NE, ct.c_longlong(X.size // NE), P(W), P(X), P(Y))
    return Y


def is_momentum_independent(K6):
    Nk, _, Nb = K6.shape[:3]
    return bool(lib().orc_is_momentum_independent(Nk, Nb, P(K6)))


def kernel_dense_from_K6(K6):
    Nk, Nb = K6.shape[0], K6.shape[2]
    Nc = Nk * Nb * Nb
    F = np.zeros((Nc, Nc), np.complex128, order="F")
    lib().orc_kernel_dense_from_K6(Nk, Nb, P(K6), P(F))
    return F


def kernel_mi_from_K6(K6):
    Nk, Nb = K6.shape[0], K6.shape[2]
    B = np.zeros((Nb * Nb, Nb * Nb), np.complex128, order="F")
    lib().orc_kernel_mi_from_K6(Nk, Nb, P(K6), P(B))
    return B


KIND = {"dense": 0, "momentum_independent": 1, "separable_ifr": 2, "low_rank": 3, "literal": 4}


def make_kernel(kind, a, b=None, c=None, rank=0):
    keep = [np.asfortranarray(x) if x is not None else None for x in (a, b, c)]
    k = Kernel(KIND[kind], int(rank), P(keep[0]), P(keep[1]), P(keep[2]))
    k._keep = keep
    return k


def contract(kernel, G, wk, qK, energy_chunk=128, alpha=1.0, beta=0.0, S=None):
    NE, Nk, Nb, _ = G.shape
    if S is None:
        S = field(NE, Nk, Nb)
    rc = lib().orc_contract(NE, Nk, Nb, ct.byref(kernel), P(G), P(wk), ct.c_double(qK),
                            int(energy_chunk), ct.c_double(alpha), ct.c_double(beta), P(S))
    _check(rc, "contract")
    return S


def uniform_spacing(eps):
    d = ct.c_double()
    _check(lib().orc_uniform_spacing(len(eps), P(eps), ct.byref(d)), "uniform_spacing")
    return d.value


def fft(x, inverse=False):
    y = np.array(x, dtype=np.complex128)
    lib().orc_fft(ct.c_longlong(len(y)), P(y), int(inverse))
    return y


def fft_hilbert(Gamma, eps, wE, roundoff=True):
    NE = Gamma.shape[0]
    Lam = np.zeros_like(Gamma, order="F")
    r = ct.c_double(0.0)
    rc = lib().orc_fft_hilbert_checked(NE, ct.c_longlong(Gamma.size // NE), P(eps), P(wE), P(Gamma),
                                       P(Lam), ct.byref(r) if roundoff else None)
    _check(rc, "fft_hilbert")
    return (Lam, r.value) if roundoff else Lam


def direct_hilbert(Gamma, eps, wE, reverse=False):
    NE = Gamma.shape[0]
    Lam = np.zeros_like(Gamma, order="F")
    rc = lib().orc_direct_hilbert(NE, ct.c_longlong(Gamma.size // NE), P(eps), P(wE), P(Gamma),
                                  P(Lam), int(reverse))
    _check(rc, "direct_hilbert")
    return Lam


def retarded_self_energy_fft(SL, SG, eps, wE):
    NE = SL.shape[0]
    SR = np.zeros_like(SL, order="F")
    Gam = np.zeros_like(SL, order="F")
    r = ct.c_double(0.0)
    rc = lib().orc_retarded_self_energy_fft(NE, ct.c_longlong(SL.size // NE), P(eps), P(wE), P(SL),
                                            P(SG), P(SR), P(Gam), ct.byref(r))
    _check(rc, "retarded_self_energy_fft")
    return SR, Gam, r.value


def lo_contraction(kernel, GL, GG, wk, qK, NLO, plan_plus, plan_minus, energy_chunk=128):
    NE, Nk, Nb, _ = GL.shape
    SL, SG = field(NE, Nk, Nb), field(NE, Nk, Nb)
    pl, pr, pw, pv = plan_plus
    ml, mr, mw, mv = plan_minus
    ip = lambda a: a.ctypes.data_as(_ip)
    up = lambda a: a.ctypes.data_as(_up)
    rc = lib().orc_lo_contraction(NE, Nk, Nb, ct.byref(kernel), P(GL), P(GG), P(wk), ct.c_double(qK),
                                  ct.c_double(NLO), ip(pl), ip(pr), P(pw), up(pv), ip(ml), ip(mr),
                                  P(mw), up(mv), int(energy_chunk), P(SL), P(SG))
    _check(rc, "lo_contraction")
    return SL, SG


def embedding_component(T, G, plan):
    NE, Nk, Nb, _ = G.shape
    S = field(NE, Nk, Nb)
    left, right, wr, valid = plan
    Tf = np.asfortranarray(T, dtype=np.complex128)
    rc = lib().orc_embedding_component(NE, Nk, Nb, P(Tf), P(G), left.ctypes.data_as(_ip),
                                       right.ctypes.data_as(_ip), P(wr), valid.ctypes.data_as(_up), P(S))
    _check(rc, "embedding_component")
    return S


def mix(old, cand, alpha):
    lib().orc_mix(ct.c_longlong(old.size), P(old), P(cand), ct.c_double(alpha))


def boundary_flux_bar(eps, wE, wk, gs, plusL, plusG, GL, GG):
    NE, Nk, Nb, _ = GL.shape
    return lib().orc_boundary_flux_bar(NE, Nk, Nb, P(eps), P(wE), P(wk), gs, P(plusL), P(plusG),
                                       P(GL), P(GG))


def state_populations(wE, wk, gs, target, GL):
    NE, Nk, Nb, _ = GL.shape
    out = np.zeros(Nb)
    lib().orc_state_populations(NE, Nk, Nb, P(wE), P(wk), gs, ct.c_double(target), P(GL), P(out))
    return out


def fermi(e, mu, kBT):
    return lib().orc_fermi(ct.c_double(e), ct.c_double(mu), ct.c_double(kBT))


def seed_weights(GR, wE, wk, gs):
    NE, Nk, Nb, _ = GR.shape
    q = np.zeros(NE)
    _check(lib().orc_seed_weights(NE, Nk, Nb, P(GR), P(wE), P(wk), gs, P(q)), "seed_weights")
    return q


def seed_mu(eps, q, kBT, target):
    mu = ct.c_double()
    _check(lib().orc_seed_mu(len(eps), P(eps), P(q), ct.c_double(kBT), ct.c_double(target),
                             ct.byref(mu)), "seed_mu")
    return mu.value


def seed_lesser(GR, eps, mu, kBT):
    NE, Nk, Nb, _ = GR.shape
    GL = field(NE, Nk, Nb)
    lib().orc_seed_lesser(NE, Nk, Nb, P(GR), P(eps), ct.c_double(mu), ct.c_double(kBT), P(GL))
    return GL


class ScbaOracle:
    """Holds an SCBA problem and state in host memory and runs oracle iterations."""

    def __init__(self, prob: dict, state: dict):
        self.prob = prob
        self.state = state
        NE, Nk, Nb = prob["NE"], prob["Nk"], prob["Nb"]
        M = len(prob["kernels"])
        self._kern = (Kernel * max(M, 1))(*prob["kernels"])
        self._qK = np.array(prob["qK"], dtype=np.float64)
        self._islo = np.array(prob["is_lo"], dtype=np.int32)
        p = Problem()
        p.NE, p.Nk, p.Nb, p.nmech, p.gs = NE, Nk, Nb, M, prob["gs"]
        p.energy_chunk = prob.get("energy_chunk", 128)
        for name in ("eps", "wE", "wk", "h", "Tplus", "Tminus"):
            setattr(p, name, P(prob[name]))
        p.Ep, p.hwLO, p.NLO = prob["Ep"], prob["hwLO"], prob["NLO"]
        p.target, p.alpha_mix = prob["target"], prob["alpha_mix"]
        p.kernels = self._kern
        p.qK = P(self._qK)
        p.is_lo = self._islo.ctypes.data_as(_ip)
        self._p = p
        s = State()
        mk = lambda lst: (_dp * max(M, 1))(*[P(x) for x in lst])
        self._scat = [mk(state["scatR"]), mk(state["scatL"]), mk(state["scatG"])]
        s.scatR, s.scatL, s.scatG = (ct.cast(a, ct.POINTER(_dp)) for a in self._scat)
        for name in ("plusR", "plusL", "plusG", "minusR", "minusL", "minusG", "embR", "embL", "embG",
                     "GR", "GL", "GG", "A", "scale"):
            setattr(s, name, P(state[name]))
        self._s = s

    def iterate(self, measure: bool = False):
        """One iteration; with measure, the row also holds the fixed-point
        measurement (r_D, r_A, r_K, r_PSD, r_caus, r_Sigma, r_lambda)."""
        out = np.zeros(5)
        pops = np.zeros(self.prob["Nb"])
        if measure:
            meas = np.zeros(7)
            rc = lib().orc_scba_iteration_measured(ct.byref(self._p), ct.byref(self._s), P(out), P(pops), P(meas))
        else:
            rc = lib().orc_scba_iteration(ct.byref(self._p), ct.byref(self._s), P(out), P(pops))
        _check(rc, "scba_iteration")
        row = dict(lam=out[0], Jbar=out[1], roundoff=out[2], number=out[3], holes=out[4], populations=pops)
        if measure:
            row.update(zip(("r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma", "r_lambda"), meas))
        return row

    def state_fields(self):
        """The ordered families' fields (mechanisms R,<,>, plus, minus)."""
        st = self.state
        out = []
        for i in range(len(self.prob["kernels"])):
            out += [st["scatR"][i], st["scatL"][i], st["scatG"][i]]
        return out + [st["plusR"], st["plusL"], st["plusG"], st["minusR"], st["minusL"], st["minusG"]]

    def iterate_anderson(self, mixer):
        """One iteration with Anderson mixing instead of the linear one: the
        candidates are produced unmixed, then mixed by mixer (AndersonOracle)
        with alpha_mix, and the embedding sum is refreshed."""
        M = len(self.prob["kernels"])
        NE, Nk, Nb = self.prob["NE"], self.prob["Nk"], self.prob["Nb"]
        cand = [field(NE, Nk, Nb) for _ in range(3 * (M + 2))]
        out = np.zeros(5)
        pops = np.zeros(Nb)
        meas = np.zeros(7)
        rc = lib().orc_scba_iteration_candidates(ct.byref(self._p), ct.byref(self._s), P(out), P(pops), P(meas),
                                                 _ptrs(cand))
        _check(rc, "scba_iteration_candidates")
        info = mixer.mix(self.state_fields(), cand, self.prob["alpha_mix"], M + 2)
        st = self.state
        for k, c in zip(("R", "L", "G"), range(3)):
            st["emb" + k][...] = st["plus" + k] + st["minus" + k]
        row = dict(lam=out[0], Jbar=out[1], roundoff=out[2], number=out[3], holes=out[4], populations=pops)
        row.update(zip(("r_D", "r_A", "r_K", "r_PSD", "r_caus", "r_Sigma", "r_lambda"), meas))
        row["anderson"] = info
        return row

def hermitian_eigvals(H):
    H = np.asfortranarray(H, dtype=np.complex128)
    n = H.shape[0]
    w = np.zeros(n)
    lib().orc_hermitian_eigvals(n, P(H), P(w))
    return w


def residual_suite(eps, h, GR, GL, GG, A, SR, SL, SG, cand_SL):
    """production_residual_suite: dict r_D, r_A, r_K, r_PSD, r_caus."""
    NE, Nk, Nb, _ = GR.shape
    out = np.zeros(5)
    arrs = [np.asfortranarray(x, dtype=np.complex128) for x in (h, GR, GL, GG, A, SR, SL, SG, cand_SL)]
    rc = lib().orc_residual_suite(NE, Nk, Nb, P(np.asarray(eps, dtype=np.float64)), *[P(x) for x in arrs], P(out))
    _check(rc, "residual_suite")
    return dict(zip(("r_D", "r_A", "r_K", "r_PSD", "r_caus"), out))


def selfenergy_residual(old, cand):
    """production_selfenergy_residual; old, cand = (SR, SL, SG)."""
    o = [np.asfortranarray(x, dtype=np.complex128) for x in old]
    c = [np.asfortranarray(x, dtype=np.complex128) for x in cand]
    arr = lambda xs: (_dp * 3)(*[P(x) for x in xs])
    lib().orc_selfenergy_residual.restype = ct.c_double
    return lib().orc_selfenergy_residual(ct.c_longlong(o[0].size), arr(o), arr(c))


def electron_density(GL, wE, wk, chi, gs):
    """_electron_density_bar: nbar on the position grid (Nz,)."""
    NE, Nk, Nb, _ = GL.shape
    chi = np.asfortranarray(chi, dtype=np.complex128)
    out = np.zeros(chi.shape[0])
    rc = lib().orc_electron_density(NE, Nk, Nb, chi.shape[0], P(np.asfortranarray(GL)), P(np.asarray(wE, float)),
                                    P(np.asarray(wk, float)), P(chi), ct.c_double(gs), P(out))
    _check(rc, "electron_density")
    return out


def periodic_poisson(eps_r, wx, ND, lambda_P, nbar):
    """solve_periodic_poisson: (U, zeta, r_P, neutrality)."""
    N = len(nbar)
    arr = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    eps_r, wx, ND, nbar = arr(eps_r), arr(wx), arr(ND), arr(nbar)
    U = np.zeros(N)
    out = np.zeros(3)
    rc = lib().orc_periodic_poisson(N, P(eps_r), P(wx), P(ND), ct.c_double(lambda_P), P(nbar), P(U), P(out))
    _check(rc, "periodic_poisson")
    return U, out[0], out[1], out[2]


def project_hamiltonians(Phi, U, S, kin, Minv):
    """project_hamiltonians with the static part S = H0 - F (Z - z0) given:
    h (Nk, Nb, Nb)."""
    Phi = np.asfortranarray(Phi, dtype=np.complex128)
    Nz, Nb = Phi.shape
    kin = np.ascontiguousarray(kin, dtype=np.float64)
    h = np.zeros((len(kin), Nb, Nb), np.complex128, order="F")
    rc = lib().orc_project_hamiltonians(Nz, Nb, len(kin), P(Phi), P(np.ascontiguousarray(U, dtype=np.float64)),
                                        P(np.asfortranarray(S, dtype=np.complex128)), P(kin),
                                        P(np.asfortranarray(Minv, dtype=np.complex128)), P(h))
    _check(rc, "project_hamiltonians")
    return h


def mix_potential(U, cand, alpha):
    """U <- (1-alpha) U + alpha cand in place; returns max |cand - U_old|."""
    lib().orc_mix_potential.restype = ct.c_double
    return lib().orc_mix_potential(len(U), P(U), P(np.ascontiguousarray(cand, dtype=np.float64)),
                                   ct.c_double(alpha))


def _ptrs(arrs):
    return (_dp * len(arrs))(*[P(a) for a in arrs])


def anderson_inner(a, b):
    """_anderson_inner over equal-length lists of fields."""
    lib().orc_anderson_inner.restype = ct.c_double
    return lib().orc_anderson_inner(ct.c_longlong(a[0].size), len(a), _ptrs(a), _ptrs(b))


def anderson_combine(coef, damping, X, R, out):
    """out_f = sum_h c_h (X[h][f] + damping R[h][f]); X, R lists of entries."""
    nf = len(out)
    flat = lambda H: [f for entry in H for f in entry]
    lib().orc_anderson_combine(ct.c_longlong(out[0].size), nf, len(coef), P(np.asarray(coef, float)),
                               ct.c_double(damping), _ptrs(flat(X)), _ptrs(flat(R)), _ptrs(out))


def anderson_guard(families):
    """_anderson_guard over (R, L, G) families: (first_failure, code)."""
    NE, Nk, Nb, _ = families[0][0].shape
    first, code = ct.c_longlong(), ct.c_int()
    lib().orc_anderson_guard(NE, Nk, Nb, len(families), _ptrs([f[0] for f in families]),
                             _ptrs([f[1] for f in families]), _ptrs([f[2] for f in families]),
                             ct.byref(first), ct.byref(code))
    return first.value, code.value


def anderson_coefficients(gram, regularization):
    """_anderson_coefficients (production_mixing.jl:160-187)."""
    count = gram.shape[0]
    if count == 1:
        return np.ones(1)
    g = gram.copy()
    scale = max(np.max(np.abs(np.diag(g))), np.finfo(float).eps)
    g[np.diag_indices(count)] += regularization * scale
    lam, V = np.linalg.eigh(g / scale)
    cutoff = max(regularization, np.sqrt(np.finfo(float).eps))
    proj = V.T @ np.ones(count)
    weights = np.array([proj[i] / lam[i] if lam[i] > cutoff else 0.0 for i in range(count)])
    c = V @ weights
    den = np.sum(c)
    if not (np.isfinite(den) and abs(den) > np.finfo(float).eps):
        raise ValueError("Anderson history is rank deficient")
    c = c / den
    if not np.all(np.isfinite(c)):
        raise ValueError("Anderson coefficient solve produced non-finite values")
    return c


class AndersonOracle:
    """_anderson_mix! (production_mixing.jl:195-317) over numpy fields: the
    history of saved states and residuals, the Gram matrix, eviction and
    reset, the bounded coefficients, the combination and the guard with the
    convex fallback."""

    def __init__(self, depth=5, damping=1.0, regularization=1e-12):
        self.depth, self.damping, self.reg = depth, damping, regularization
        self.states, self.residuals = [], []
        self.gram = np.zeros((0, 0))

    def clear(self):
        self.states, self.residuals = [], []
        self.gram = np.zeros((0, 0))

    def append(self, state, cand):
        prior = None if not self.residuals else self.gram[-1, -1]
        if len(self.states) >= self.depth:
            self.states.pop(0)
            self.residuals.pop(0)
            self.gram = self.gram[1:, 1:]
        saved = [np.array(x, order="F") for x in state]
        resid = [np.asfortranarray(c - x) for c, x in zip(cand, state)]
        norm = anderson_inner(resid, resid)
        eps = np.finfo(float).eps
        reset = prior is not None and norm > (1 + 64 * eps) * max(prior, np.finfo(float).tiny)
        if reset:
            self.clear()
        count = len(self.states) + 1
        g = np.zeros((count, count))
        if count > 1:
            g[:-1, :-1] = self.gram
        for i in range(count - 1):
            g[i, -1] = g[-1, i] = anderson_inner(self.residuals[i], resid)
        g[-1, -1] = norm
        self.gram = g
        self.states.append(saved)
        self.residuals.append(resid)
        return reset

    def mix(self, state, cand, alpha, nfam):
        """state, cand: field lists (families x R,L,G); state updated in place."""
        reset = self.append(state, cand)
        reason = "residual_increase" if reset else "none"
        try:
            c = anderson_coefficients(self.gram, self.reg)
        except ValueError:
            reason = "coefficient_solve"
            c = np.r_[np.zeros(len(self.residuals) - 1), 1.0]
        if np.sum(np.abs(c)) > 10 or not np.all(np.isfinite(c)):
            reason = "coefficient_bound"
            c = np.r_[np.zeros(len(self.residuals) - 1), 1.0]
        damping = alpha * self.damping
        anderson_combine(c, damping, self.states, self.residuals, state)
        fams = [state[3 * f:3 * f + 3] for f in range(nfam)]
        first, code = anderson_guard(fams)
        accepted = first == 0 and np.count_nonzero(c) > 1
        if first != 0:
            x, r = self.states[-1], self.residuals[-1]
            for out, xf, rf in zip(state, x, r):
                out[...] = xf + damping * rf
            self.clear()
            reason = "causality_guard"
            if anderson_guard(fams)[0] != 0:
                raise ValueError("convex Anderson fallback is not admissible")
        return dict(coefficients=c, reset_reason=reason, accepted=accepted, first_failure=first, code=code)
