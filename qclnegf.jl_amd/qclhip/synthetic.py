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
