"""Seeded synthetic inputs in the solver's dimensionless units.

Units follow the reference's ScaleSystem (src/domain/scaling.jl:10-24):
energies in E0 = 0.1 eV, lengths in L0 = 10 nm, m_ref = 0.067 m0. Grids follow
build_grids (src/numerics/reference/grids.jl:10-47): uniform energies with
trapezoid weights, momenta on [0, k_max] with annular weights
wk_m = (b_{m+1}^2 - b_m^2)/(4 pi).

The batched Green's-function workload ("GF solves"): N basis states, NE
energies uniform on [-0.1, 0.3] eV, Nk momenta with parabolic dispersion
(m* = 0.067 m0), h_m = H0 + Cbar kappa_m^2 I with H0 Hermitian banded
(bandwidth 4, diagonal U[-0.1, 0.3] eV, couplings N(0, 5 meV)), Sigma^R
diagonal with Im = -U[0.5, 5] meV, Sigma< = i f Gamma, Sigma> = -i (1-f) Gamma
(Gamma = -2 Im Sigma^R, f Fermi at 77 K, mu = 0.1 eV).
"""
from __future__ import annotations

import math

import numpy as np

E0_EV = 0.1
L0_M = 10e-9
HBAR = 1.054571817e-34
M0 = 9.1093837015e-31
QE = 1.602176634e-19
KB_EV = 8.617333262e-5
# Kinetic coefficient hbar^2/(2 m_ref E0 L0^2) with m_ref = 0.067 m0 (scaling.jl:92-93).
CBAR = HBAR ** 2 / (2 * 0.067 * M0 * E0_EV * QE * L0_M ** 2)
SEED = 20261203


def energy_grid(NE, emin_eV=-0.1, emax_eV=0.3):
    eps = np.linspace(emin_eV / E0_EV, emax_eV / E0_EV, NE)
    de = (emax_eV - emin_eV) / E0_EV / (NE - 1)
    wE = np.full(NE, de)
    wE[[0, -1]] /= 2
    return eps, wE


def momentum_grid(Nk, kmax):
    kappa = np.linspace(0.0, kmax, Nk)
    b = np.zeros(Nk + 1)
    for m in range(1, Nk):
        b[m] = (kappa[m - 1] + kappa[m]) / 2
    b[-1] = kmax
    wk = np.array([(b[m + 1] ** 2 - b[m] ** 2) / (4 * math.pi) for m in range(Nk)])
    return kappa, wk


def kmax_for_energy(E_eV=0.3):
    """Momentum at which the parabolic kinetic energy reaches E_eV."""
    return math.sqrt((E_eV / E0_EV) / CBAR)


def kbt_scaled(T_K):
    return KB_EV * T_K / E0_EV


def banded_hamiltonian(rng, N, bandwidth=4, diag_eV=(-0.1, 0.3), coupling_eV=0.005):
    H = np.zeros((N, N), np.complex128)
    H[np.diag_indices(N)] = rng.uniform(diag_eV[0], diag_eV[1], N) / E0_EV
    s = coupling_eV / E0_EV / math.sqrt(2.0)
    for a in range(N):
        for b in range(a + 1, min(N, a + bandwidth + 1)):
            v = complex(rng.normal(0.0, s), rng.normal(0.0, s))
            H[a, b] = v
            H[b, a] = np.conj(v)
    return H


def gf_workload(N=24, NE=2048, Nk=64, seed=SEED, T_K=77.0, mu_eV=0.1):
    """Host description of the synthetic GF-solve batch.

    Returns dict with eps, wE, kappa, wk, h (Nk,N,N), and the diagonal
    self-energy data gamma (NE,Nk,N) and f (NE) from which
    Sigma^R = -i Gamma/2, Sigma< = i f Gamma, Sigma> = -i (1-f) Gamma."""
    rng = np.random.default_rng(seed)
    eps, wE = energy_grid(NE)
    kappa, wk = momentum_grid(Nk, kmax_for_energy(0.3))
    H0 = banded_hamiltonian(rng, N)
    h = np.zeros((Nk, N, N), np.complex128, order="F")
    for m in range(Nk):
        h[m] = H0 + CBAR * kappa[m] ** 2 * np.eye(N)
    gamma = 2.0 * rng.uniform(0.0005, 0.005, size=(NE, Nk, N)) / E0_EV
    kbt = kbt_scaled(T_K)
    x = (eps - mu_eV / E0_EV) / kbt
    with np.errstate(over="ignore"):
        f = np.where(x > 40, 0.0, np.where(x < -40, 1.0, 1.0 / (np.exp(np.clip(x, -700, 700)) + 1.0)))
    return dict(N=N, NE=NE, Nk=Nk, eps=eps, wE=wE, kappa=kappa, wk=wk, h=h, H0=H0, gamma=gamma, f=f,
                kbt=kbt, mu=mu_eV / E0_EV)


def diag_fields_host(w):
    """Dense host fields (NE,Nk,N,N) Fortran order of the workload (or of an
    injection_family dict): (SR, SL, SG)."""
    NE, Nk, N = w["NE"], w["Nk"], w["N"]
    SR = np.zeros((NE, Nk, N, N), np.complex128, order="F")
    SL = np.zeros_like(SR)
    SG = np.zeros_like(SR)
    g = w["gamma"]
    f = w["f"][:, None]
    for a in range(N):
        SR[:, :, a, a] = -0.5j * g[:, :, a]
        SL[:, :, a, a] = 1j * f * g[:, :, a]
        SG[:, :, a, a] = -1j * (1.0 - f) * g[:, :, a]
    return SR, SL, SG


def injection_family(w, dmu_eV=0.05, gamma_meV=(0.5, 2.0), seed=SEED + 1):
    """Diagonal family of a reservoir at mu + dmu (the `plus` embedding family of
    the current functional): dict(gamma (NE,Nk,N), f (NE)). Together with the
    workload's own family it drives a non-zero current."""
    rng = np.random.default_rng(seed)
    NE, Nk, N = w["NE"], w["Nk"], w["N"]
    gamma = 2.0 * rng.uniform(gamma_meV[0] * 1e-3, gamma_meV[1] * 1e-3, size=(NE, Nk, N)) / E0_EV
    x = (w["eps"] - w["mu"] - dmu_eV / E0_EV) / w["kbt"]
    with np.errstate(over="ignore"):
        f = np.where(x > 40, 0.0, np.where(x < -40, 1.0, 1.0 / (np.exp(np.clip(x, -700, 700)) + 1.0)))
    return dict(N=N, NE=NE, Nk=Nk, gamma=gamma, f=f)


def diag_fields_device(w, device="cuda"):
    """The same fields built directly in device memory (torch shape (N,N,Nk,NE)).
    `w` may also be an injection_family dict."""
    import torch
    NE, Nk, N = w["NE"], w["Nk"], w["N"]
    g = torch.from_numpy(np.ascontiguousarray(np.transpose(w["gamma"], (2, 1, 0)))).to(device)  # (N,Nk,NE)
    f = torch.from_numpy(w["f"]).to(device)
    out = []
    for coef in (lambda gg: -0.5j * gg, lambda gg: 1j * f * gg, lambda gg: -1j * (1.0 - f) * gg):
        F = torch.zeros((N, N, Nk, NE), dtype=torch.complex128, device=device)
        d = torch.diagonal(F, dim1=0, dim2=1)  # (Nk, NE, N) view
        d.copy_(coef(g.to(torch.complex128)).permute(1, 2, 0))
        out.append(F)
    return tuple(out)


def covariance_block(rng, N, nvert=3, scale=0.05):
    """Momentum-independent elastic kernel block[(a,b),(c,d)] = sum_h V_h[a,c] conj(V_h[b,d])
    with random Hermitian vertices (a positivity-preserving map; rows (a,b),
    columns (c,d), first index fastest)."""
    B = np.zeros((N * N, N * N), np.complex128)
    for _ in range(nvert):
        V = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        V = scale * (V + V.conj().T) / (2 * math.sqrt(N))
        B += np.einsum("ac,bd->abcd", V, V.conj()).reshape(N * N, N * N, order="F")
    return B


def separable_factors(rng, N, Nk, J=3, scale=0.05):
    """IFR-like separable kernel: left[(a,b),j] = s u_j[a] conj(u_j[b]),
    right[j,(c,d)] = conj(u_j[c]) u_j[d], radial[m,m'] smooth and positive."""
    left = np.zeros((N * N, J), np.complex128)
    right = np.zeros((J, N * N), np.complex128)
    for j in range(J):
        u = rng.normal(size=N) + 1j * rng.normal(size=N)
        u /= np.linalg.norm(u)
        for b in range(N):
            for a in range(N):
                left[a + N * b, j] = scale * u[a] * np.conj(u[b])
                right[j, a + N * b] = np.conj(u[a]) * u[b]
    x = np.arange(Nk)
    radial = np.exp(-((x[:, None] - x[None, :]) / max(1.0, Nk / 4)) ** 2).astype(np.complex128)
    return left, right, radial


def low_rank_factors(rng, N, Nk, r=4, scale=0.05):
    """Low-rank LO-like kernel Kflat = left * right_adjoint (Nc = Nk N^2) with
    left[(m,a,b),i] = s p_m u_i[a] conj(u_i[b]) and
    right_adjoint[i,(m',c,d)] = p_m' conj(u_i[c]) u_i[d] for real positive p,
    so K maps anti-Hermitian blocks to anti-Hermitian blocks."""
    Nc = Nk * N * N
    p = 1.0 + 0.5 * np.cos(np.linspace(0.0, math.pi, Nk))
    left = np.zeros((Nc, r), np.complex128)
    radj = np.zeros((r, Nc), np.complex128)
    for i in range(r):
        u = rng.normal(size=N) + 1j * rng.normal(size=N)
        u /= np.linalg.norm(u)
        P = np.outer(u, u.conj())  # P[a,b] = u_a conj(u_b)
        for b in range(N):
            for a in range(N):
                for m in range(Nk):
                    q = m + Nk * (a + N * b)
                    left[q, i] = scale * p[m] * P[a, b]
                    radj[i, q] = p[m] * np.conj(P[a, b])
    return left, radj


def dense_covariance_kernel(rng, N, Nk, nvert=2, scale=0.03):
    """Dense Kflat[(m,a,b),(m',c,d)] = R[m,m'] sum_h V_h[a,c] conj(V_h[b,d]) with a
    smooth positive R (a Hermiticity-preserving momentum-dependent map)."""
    B = covariance_block(rng, N, nvert=nvert, scale=scale)
    x = np.arange(Nk)
    R = np.exp(-((x[:, None] - x[None, :]) / max(1.0, Nk / 3)) ** 2)
    Nc = Nk * N * N
    F = np.zeros((Nc, Nc), np.complex128)
    for q in range(N * N):
        for qp in range(N * N):
            F[q * Nk:(q + 1) * Nk, qp * Nk:(qp + 1) * Nk] = R * B[q, qp]
    return F


# ---- one GaAs/Al(x)Ga(1-x)As period: the full-iteration workload ----------

EPS0 = 8.8541878128e-12
# Poisson coupling e^2/(eps0 E0 L0) of the scale system (scaling.jl:10-24)
LAMBDA_P = QE ** 2 / (EPS0 * E0_EV * QE * L0_M)


def period_profile(thick_nm, Ldom, Nz, x_al=0.15, Ep_eV=0.05):
    """Band edge, inverse mass and material fractions of a six-layer period
    (barrier first, alternating) on Nz cells of a fixed domain [0, Ldom]
    (solver units): each cell takes its barrier fraction (the overlap of the
    cell with the barrier layers), so the profile moves continuously with
    the thicknesses; cells beyond the last layer are well material.
    Returns (z, dz, V, inv_mass, barrier_fraction, doped_fraction, edges)."""
    dz = Ldom / Nz
    z = (np.arange(Nz) + 0.5) * dz
    edges = np.concatenate([[0.0], np.cumsum(thick_nm) * 1e-9 / L0_M])
    lo, hi = z - dz / 2, z + dz / 2
    overlap = lambda a, b: np.clip(np.minimum(hi, b) - np.maximum(lo, a), 0.0, None) / dz
    frac = sum(overlap(edges[j], edges[j + 1]) for j in (0, 2, 4))
    doped = overlap(edges[5], edges[6])
    dEc = 0.835 * x_al / E0_EV
    mb = (0.067 + 0.083 * x_al) / 0.067  # barrier mass in m_ref
    inv_mass = frac / mb + (1.0 - frac)
    V = frac * dEc - (Ep_eV / E0_EV) * z / Ldom
    return z, dz, V, inv_mass, frac, doped, edges


def period_fd_hamiltonian(V, inv_mass, dz):
    """Finite-difference BenDaniel-Duke Hamiltonian -d/dz (Cbar/m) d/dz + V
    with hard walls (1/m at half points: the mean of the neighbours')."""
    inv_half = 0.5 * (inv_mass[:-1] + inv_mass[1:])
    off = -CBAR * inv_half / (dz * dz)
    diag = V.copy()
    diag[:-1] += CBAR * inv_half / (dz * dz)
    diag[1:] += CBAR * inv_half / (dz * dz)
    diag[0] += CBAR * inv_mass[0] / (dz * dz)
    diag[-1] += CBAR * inv_mass[-1] / (dz * dz)
    return np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)


def period_h(w, thick_nm, U=None):
    """h_m (Nk, N, N) of the period with thicknesses thick_nm projected on the
    FIXED basis of w (its eigenstates at w's own thicknesses):
    h_m = Phi^H (H_fd + diag(U)) Phi + Cbar kappa_m^2 Phi^H diag(1/m) Phi,
    Phi = phi sqrt(dz) orthonormal (project_hamiltonians, basis.jl:288-308).
    At w's thicknesses and U = 0 this is w's h."""
    z, dz, V, inv_mass, _, _, _ = period_profile(thick_nm, w["Lp"], w["Nz"], w["x_al"], w["Ep"] * E0_EV)
    H = period_fd_hamiltonian(V + (0.0 if U is None else U), inv_mass, dz)
    P = w["phi"] * math.sqrt(dz)
    S = P.T @ H @ P
    Mi = (P * inv_mass[:, None]).T @ P
    h = np.zeros((w["Nk"], w["N"], w["N"]), np.complex128, order="F")
    for m in range(w["Nk"]):
        h[m] = S + w["kin"][m] * Mi
    return h


def two_well_period(N=32, NE=4096, Nk=128, Nz=512, x_al=0.15, Ep_eV=0.05, T_K=77.0, doping_cm3=3e16,
                    seed=SEED + 30):
    """A six-layer Al(x)Ga(1-x)As / GaAs period (barrier first, alternating)
    with seeded thicknesses U[2, 20] nm, on Nz grid points, in the solver's
    units. Material inputs are explicit (the reference's Layer takes the band
    edge and mass directly, types.jl:11-66): conduction-band offset
    0.835 x eV, masses 0.067 (GaAs) and 0.067 + 0.083 x m0, permittivities
    12.9 and 12.9 - 2.84 x; the bias drops Ep_eV over the period; the last
    well is doped at doping_cm3. Cells straddling an interface take their
    material fractions (period_profile).

    Basis: the N lowest eigenstates phi of the period's finite-difference
    BenDaniel-Duke Hamiltonian (hard walls); S = H0 = Phi^H H Phi (diagonal
    to rounding), the in-plane mass matrix Minv = Phi^H diag(m_ref/m) Phi and
    h_m = H0 + Cbar kappa_m^2 Minv (Phi = phi sqrt(dz)); the coupling across
    the period boundary T+_ab = -t phi_a(z_last) phi_b(z_first) dz (t the
    boundary hopping), T- = T+^H.

    Scattering kernels in the device representations, their shapes from the
    basis (magnitudes are calibrated later through qK):
      acoustic, momentum independent: sum_z dz V_z (x) conj(V_z), V_z[a,c] = phi_a phi_c;
      alloy, momentum independent: the same weighted by the barrier fraction;
      IFR, separable: one rank per interface z_j, u_j = phi(z_j), radial
        exp(-((kappa_m - kappa_m') Lambda / 2)^2), Lambda = 8 nm;
      LO, low rank: r projectors u_i u_i^H of smooth envelopes
        (u_i[a] = sum_z phi_a g_i dz), Froehlich-like weights 1/(1 + (kappa/q_i)^2);
      impurity, low rank: the same with envelopes exp(-|z - z_d|/l_i) about the doped layer.
    Returns a dict with the grids, h (Nk,N,N), Tplus, Tminus, the kernel
    factors, the Poisson inputs (eps_r, wx, ND, lambda_P, chi, Phi, S, kin,
    Minv, ND2D, Lp) and the layer thicknesses."""
    rng = np.random.default_rng(seed)
    thick_nm = rng.uniform(2.0, 20.0, 6)
    Lp = thick_nm.sum() * 1e-9 / L0_M
    z, dz, V, inv_mass, frac, doped, edges = period_profile(thick_nm, Lp, Nz, x_al, Ep_eV)
    H = period_fd_hamiltonian(V, inv_mass, dz)
    E, vec = np.linalg.eigh(H)
    phi = vec[:, :N] / math.sqrt(dz)  # continuum normalization: sum phi^2 dz = 1
    Ebasis = E[:N]
    P = phi * math.sqrt(dz)
    H0 = (P.T @ H @ P).astype(np.complex128)
    Minv = ((P * inv_mass[:, None]).T @ P).astype(np.complex128)
    eps, wE = energy_grid(NE)
    kappa, wk = momentum_grid(Nk, kmax_for_energy(0.3))
    kin = CBAR * kappa ** 2
    h = np.zeros((Nk, N, N), np.complex128, order="F")
    for m in range(Nk):
        h[m] = H0 + kin[m] * Minv
    t_edge = CBAR * inv_mass[0] / (dz * dz)
    Tplus = (-t_edge * dz * np.outer(phi[-1], phi[0])).astype(np.complex128)
    in_barrier = frac
    # kernels
    Vz = np.einsum("za,zc->zac", phi, phi).reshape(Nz, N * N)  # V_z[a,c], (a,c) with a fastest

    def mi_block(weight):
        # B[(a,b),(c,d)] = sum_z w_z V_z[a,c] V_z[b,d] (real basis)
        M = (Vz * weight[:, None]).T @ Vz  # M[(a,c),(b,d)]
        return np.ascontiguousarray(M.reshape(N, N, N, N).transpose(0, 2, 1, 3)).reshape(N * N, N * N).astype(
            np.complex128)

    acoustic = mi_block(np.full(Nz, dz))
    alloy = mi_block(dz * in_barrier * x_al * (1 - x_al))
    zint = edges[1:]  # six interfaces per period (the last is the period boundary)
    J = len(zint)
    idx = np.minimum((zint / dz).astype(int), Nz - 1)
    ifr_left = np.zeros((N * N, J), np.complex128)
    ifr_right = np.zeros((J, N * N), np.complex128)
    for j, iz in enumerate(idx):
        u = phi[iz]
        P = np.outer(u, u)  # P[a,b]
        ifr_left[:, j] = P.reshape(-1, order="F")
        ifr_right[j, :] = P.reshape(-1, order="F")
    Lam = 8e-9 / L0_M
    ifr_radial = np.exp(-(((kappa[:, None] - kappa[None, :]) * Lam / 2) ** 2)).astype(np.complex128)

    def low_rank(envelopes, weights):
        r = len(envelopes)
        Nc = Nk * N * N
        left = np.zeros((Nc, r), np.complex128)
        radj = np.zeros((r, Nc), np.complex128)
        for i, (g, wm) in enumerate(zip(envelopes, weights)):
            u = phi.T @ g * dz
            u /= max(np.linalg.norm(u), 1e-300)
            P = np.outer(u, u).reshape(-1, order="F")  # (a,b), a fastest
            col = (wm[:, None] * P[None, :]).reshape(-1, order="F")  # (m,(a,b)), m fastest
            left[:, i] = col
            radj[i, :] = np.conj(col)
        return left, radj

    zc = z / Lp
    lo_env = [np.exp(-((zc - c) / 0.18) ** 2) for c in (0.1, 0.3, 0.5, 0.7, 0.9, 0.5)]
    lo_env[-1] = np.ones(Nz)
    qs = [0.5, 1.0, 2.0, 1.0, 0.5, 3.0]
    lo_w = [1.0 / (1.0 + (kappa / q) ** 2) for q in qs]
    lo_left, lo_radj = low_rank(lo_env, lo_w)
    zd = 0.5 * (edges[5] + edges[6])
    imp_env = [np.exp(-np.abs(z - zd) / l) for l in (0.2, 0.5, 1.0, 2.0)]
    imp_w = [np.exp(-kappa * l) for l in (0.2, 0.5, 1.0, 2.0)]
    imp_left, imp_radj = low_rank(imp_env, imp_w)
    eps_r = 12.9 - 2.84 * x_al * frac
    ND = doped * doping_cm3 * 1e6 * L0_M ** 3
    return dict(N=N, NE=NE, Nk=Nk, Nz=Nz, eps=eps, wE=wE, kappa=kappa, wk=wk, kbt=kbt_scaled(T_K), h=h, H0=H0,
                Minv=Minv, kin=kin, Tplus=Tplus, Tminus=Tplus.conj().T.copy(), thick_nm=thick_nm, x_al=x_al,
                Ep=Ep_eV / E0_EV, Ebasis=Ebasis, phi=phi, dz=dz, z=z, Lp=Lp,
                kernels=dict(acoustic=("momentum_independent", acoustic, None, None, 0),
                             alloy=("momentum_independent", alloy, None, None, 0),
                             IFR=("separable_ifr", ifr_left, ifr_right, ifr_radial, J),
                             LO=("low_rank", lo_left, lo_radj, None, lo_left.shape[1]),
                             impurity=("low_rank", imp_left, imp_radj, None, imp_left.shape[1])),
                eps_r=eps_r, wx=np.full(Nz, dz), ND=ND, lambda_P=LAMBDA_P, ND2D=float(dz * ND.sum()),
                S=H0.copy())  # the bias is inside H0: h = S + Phi^H diag(U) Phi + kin Minv


# Scattering strength of each mechanism of the period workload: qK is chosen
# so that the mechanism's S< from the seed state peaks at this level (eV).
PERIOD_LEVELS_EV = dict(acoustic=1e-3, alloy=3e-4, IFR=1e-3, LO=2e-3, impurity=1e-3)
PERIOD_MECHANISMS = ("LO", "acoustic", "impurity", "IFR", "alloy")  # the reference's mechanism order
HW_LO_EV = 0.036


def calibrated_qk(contract_peak, w):
    """qK per mechanism from contract_peak(name) = max |K[G<_seed]| at qK = 1."""
    return {name: (PERIOD_LEVELS_EV[name] / E0_EV) / max(contract_peak(name), 1e-300) for name in PERIOD_MECHANISMS}


def period_scba_problem(w, precision="fp64", eta_seed=0.05, alpha_mix=0.3, verify_roundoff=True):
    """The SCBA problem of a two_well_period dict on the device: five
    mechanisms in the device representations with calibrated qK, the seed at
    the chemical potential that makes the number equal the sheet doping
    N_D^2D (neutrality; seed.jl:151-197), no stored A. Returns (problem,
    state, qK dict, mu)."""
    import torch

    from . import api, scba
    g = api.Grid(w["eps"], w["wE"], w["wk"], w["N"])
    g.set_precision(precision)
    h = api.h_to_device(w["h"])
    seed, mu = api.seed_green_production(g, h, eta_seed, w["kbt"], w["ND2D"], spectral=False)
    kernels = {name: api.ProductionKernel(g, kind, a, b, c, rank=r)
               for name, (kind, a, b, c, r) in w["kernels"].items()}

    def peak(name):
        S = api.production_static_contraction(kernels[name], seed.GL, 1.0)
        v = float(torch.abs(S).max().item())
        del S
        return v

    qK = calibrated_qk(peak, w)
    hw = HW_LO_EV / E0_EV
    mechs = [scba.Mechanism(name, kernels[name], qK[name], name == "LO") for name in PERIOD_MECHANISMS]
    p = scba.ScbaProblem(g, h, api.matrix_to_device(w["Tplus"]), api.matrix_to_device(w["Tminus"]), mechs,
                         w["ND2D"], Ep=w["Ep"], hwLO=hw, NLO=1.0 / math.expm1(hw / w["kbt"]), alpha_mix=alpha_mix,
                         verify_roundoff=verify_roundoff)
    s = scba.initial_state(p, seed)
    s.green.A = None
    return p, s, qK, mu
