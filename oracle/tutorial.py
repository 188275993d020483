"""Test infrastructure (not part of the product): a numpy restatement of the
reference's assembly of its tutorial problem, so that the device SCBA can be
run on the reference's own smallest configuration and checked against the
oracle. Only the tests load it.

Restated from QCLNEGF.jl v0.2.0:
  reference_parameters            src/domain/presets.jl:16-84
  tutorial_numerics               src/domain/presets.jl:144-158
  tutorial_options (alpha, gates) src/domain/presets.jl:194-228
  ScaleSystem, scaled_value, C    src/domain/scaling.jl:10-24,26-51,92-110; constants.jl:60
  build_grids, build_profiles     src/numerics/reference/grids.jl:10-92
  build_bdd_hamiltonian           src/numerics/reference/hamiltonian.jl:16-66
  build_localized_basis (:pzp)    src/numerics/reference/basis.jl:1-280 (window of 2P+1 periods,
                                  position operator diagonalized in the window's lowest
                                  eigenspace, central columns, Loewdin, _phase_fix!)
  project_hamiltonians            src/numerics/reference/basis.jl:288-308
  acoustic_kernel + KernelSet     src/numerics/reference/kernels.jl:239-270, 367-407
  seed temperature, target        src/numerics/optimized/seed.jl:151-197
Parity unpinned: no Julia exists here, so the equality of this assembly (basis,
h, T+-, the acoustic kernel) with the reference's numbers cannot be shown;
the assembly only supplies inputs of the reference's shape and physics to the
device-vs-oracle comparison, which is pinned separately.
"""
from __future__ import annotations

import math

import numpy as np

# CODATA 2018 (the values Unitful's PhysicalConstants.CODATA2018 carries)
HBAR = 1.054571817e-34
M0 = 9.1093837015e-31
QE = 1.602176634e-19
KB = 1.380649e-23
E0_EV = 0.1          # ScaleSystem defaults: E0 = 0.1 eV, L0 = 10 nm
L0_M = 10e-9
E0_J = E0_EV * QE
CBAR = HBAR ** 2 / (2 * M0) / (E0_J * L0_M ** 2)   # C0 = hbar^2/(2 m0) over E0 L0^2
J0_A_M2 = QE * E0_J / (HBAR * L0_M ** 2)            # current scale e E0/(hbar L0^2)


def reference_parameters(V_period_V=0.056, T_K=200.0):
    """The six layers of the reference structure and the parameters the
    acoustic-only tutorial uses (presets.jl:16-84)."""
    barrier = dict(Ec_eV=0.20775, mz=0.08775, mpar=0.08775)
    well = dict(Ec_eV=0.0, mz=0.067, mpar=0.067)
    layers = [(3.26, barrier, False), (7.99, well, False), (1.90, barrier, False), (8.40, well, False),
              (2.90, well, True), (5.16, well, False)]
    Lp_nm = sum(d for d, _, _ in layers)
    return dict(layers=layers, Lp_nm=Lp_nm, N_dop2D_m2=4.5e14, f_ion=1.0, F_V_per_m=V_period_V / (Lp_nm * 1e-9),
                T_K=T_K, Xi_eV=7.0, rho=5317.0, v_s=5240.0, gs=2)


def tutorial_numerics():
    """presets.jl:144-158."""
    return dict(N_z=48, N_b=3, P_basis=1, E_min=-0.10, E_max=0.28, N_E=49, k_max_nm=0.45, N_k=5, eta_seed_eV=2e-3)


def tutorial_options():
    """The SCBA part of tutorial_options (presets.jl:194-228): mixing and the
    fixed-point gates the device iteration measures."""
    return dict(alpha=0.35, max_scba=80,
                gates=dict(r_D=1e-10, r_A=1e-6, r_K=5e-4, r_Sigma=5e-4, r_lambda=5e-4, r_PSD=1e-7, r_caus=1e-7),
                r_roundoff=1e-6)


def build_grids(p, n):
    """grids.jl:10-47 (E_ref = 0)."""
    Lp = p["Lp_nm"] * 1e-9 / L0_M
    dx = Lp / n["N_z"]
    x = (np.arange(1, n["N_z"] + 1) - 0.5) * dx
    wx = np.full(n["N_z"], dx)
    emin, emax = n["E_min"] / E0_EV, n["E_max"] / E0_EV
    eps = np.linspace(emin, emax, n["N_E"])
    de = (emax - emin) / (n["N_E"] - 1)
    wE = np.full(n["N_E"], de)
    wE[[0, -1]] /= 2
    kmax = n["k_max_nm"] * 1e9 * L0_M
    kappa = np.linspace(0.0, kmax, n["N_k"])
    b = np.zeros(n["N_k"] + 1)
    for m in range(1, n["N_k"]):
        b[m] = (kappa[m - 1] + kappa[m]) / 2
    b[-1] = kmax
    wk = np.array([(b[m + 1] ** 2 - b[m] ** 2) / (4 * math.pi) for m in range(n["N_k"])])
    return dict(x=x, wx=wx, dx=dx, Lp=Lp, eps=eps, wE=wE, kappa=kappa, wk=wk)


def build_profiles(p, n, g):
    """grids.jl:49-92: half-open layers z_left <= z < z_right."""
    right = np.cumsum([d * 1e-9 / L0_M for d, _, _ in p["layers"]])
    idx = np.minimum(np.searchsorted(right, g["x"], side="right"), len(p["layers"]) - 1)
    Ec = np.array([p["layers"][i][1]["Ec_eV"] / E0_EV for i in idx])
    mz = np.array([p["layers"][i][1]["mz"] for i in idx])
    mpar = np.array([p["layers"][i][1]["mpar"] for i in idx])
    chiD = np.array([1.0 if p["layers"][i][2] else 0.0 for i in idx])
    ND2D = p["f_ion"] * p["N_dop2D_m2"] * L0_M ** 2
    ND = ND2D * chiD / float(np.dot(g["wx"], chiD))
    return dict(Ec=Ec, mz=mz, mpar=mpar, ND=ND, ND2D=ND2D)


def bdd_hamiltonian(Ec, mz, dx, U=None):
    """hamiltonian.jl:16-45, Dirichlet: face inverse mass 2/(m_l + m_r)."""
    N = len(Ec)
    U = np.zeros(N) if U is None else U
    t = CBAR * 2.0 / (mz[:-1] + mz[1:]) / dx ** 2
    H = np.zeros((N, N))
    H[np.arange(N - 1), np.arange(1, N)] = -t
    H[np.arange(1, N), np.arange(N - 1)] = -t
    left = np.r_[CBAR / (mz[0] * dx ** 2), t]
    right = np.r_[t, CBAR / (mz[-1] * dx ** 2)]
    H[np.diag_indices(N)] = Ec + U + left + right
    return H


def phase_fix(Phi):
    """_phase_fix! (basis.jl:1-9): the largest-magnitude entry of each column
    real and positive (first maximum on ties, as argmax)."""
    Phi = Phi.copy()
    for a in range(Phi.shape[1]):
        g = int(np.argmax(np.abs(Phi[:, a])))
        Phi[:, a] *= np.exp(-1j * np.angle(Phi[g, a]))
        if Phi[g, a].real < 0:
            Phi[:, a] *= -1
    return Phi


def build_localized_basis(p, n, g, prof):
    """build_localized_basis with localization = :pzp (basis.jl:101-245)."""
    Np = 2 * n["P_basis"] + 1
    Nz, Nb = n["N_z"], n["N_b"]
    Nsub = Np * Nb
    dx = g["dx"]
    Hwin = bdd_hamiltonian(np.tile(prof["Ec"], Np), np.tile(prof["mz"], Np), dx)
    E0grid = float(np.max(np.abs(np.diag(Hwin))))
    vals, vecs = np.linalg.eigh(Hwin / E0grid)
    C = vecs[:, :Nsub].astype(np.complex128)
    xwin = (np.arange(1, Np * Nz + 1) - 0.5) * dx
    Zsub = C.conj().T @ (xwin[:, None] * C)
    zv, zw = np.linalg.eigh((Zsub + Zsub.conj().T) / 2)
    order = np.argsort(zv, kind="stable")
    Phit = C @ zw[:, order]
    rows = slice(n["P_basis"] * Nz, (n["P_basis"] + 1) * Nz)
    cols = slice(n["P_basis"] * Nb, (n["P_basis"] + 1) * Nb)
    B = Phit[rows, cols]
    S = B.conj().T @ B
    S = (S + S.conj().T) / 2
    sv, sw = np.linalg.eigh(S)
    if sv.min() <= 0 or sv.min() <= 100 * np.finfo(float).eps * sv.max() or sv.max() / sv.min() > 1e8:
        raise ValueError("cell overlap is singular or ill-conditioned")
    Phi = phase_fix(B @ (sw @ np.diag(1 / np.sqrt(sv)) @ sw.conj().T))
    Hcell = Hwin[rows, rows]
    H0 = Phi.conj().T @ Hcell @ Phi
    Z = Phi.conj().T @ (g["x"][:, None] * Phi)
    Minv = Phi.conj().T @ ((1 / prof["mpar"])[:, None] * Phi)
    Hright = Hwin[rows, slice((n["P_basis"] + 1) * Nz, (n["P_basis"] + 2) * Nz)]
    Tplus = Phi.conj().T @ Hright @ Phi
    return dict(Phi=Phi, chi=Phi / math.sqrt(dx), H0=H0, Z=Z, Minv=Minv, Tplus=Tplus, Tminus=Tplus.conj().T.copy(),
                r_orth=float(np.linalg.norm(Phi.conj().T @ Phi - np.eye(Nb))))


def project_hamiltonians(basis, g, F, U=None, z0=0.0):
    """basis.jl:288-308 (parabolic): h_m = H0 - F (Z - z0) + Phi^H diag(U) Phi + C kappa_m^2 Minv."""
    Nb = basis["H0"].shape[0]
    Nz = basis["Phi"].shape[0]
    U = np.zeros(Nz) if U is None else U
    Uproj = basis["Phi"].conj().T @ (U[:, None] * basis["Phi"])
    static = basis["H0"] - F * (basis["Z"] - z0 * np.eye(Nb)) + Uproj
    h = np.zeros((len(g["kappa"]), Nb, Nb), np.complex128, order="F")
    for m, k in enumerate(g["kappa"]):
        h[m] = static + CBAR * k ** 2 * basis["Minv"]
    return h


def acoustic_kernel(p, g, basis):
    """acoustic_kernel (kernels.jl:239-270): the momentum-independent block
    K[a,c,d,b] = Xi^2 kBT/(rho v_s^2) sum_g wx conj(chi_a) chi_c conj(chi_d) chi_b / L0 / K0,
    K0 = E0^2 L0^2, then normalized by its largest magnitude qK
    (build_kernels, kernels.jl:367-407). Returns (K4[a,c,d,b], qK)."""
    chi = basis["chi"]
    Xi = p["Xi_eV"] * QE
    kBT = KB * p["T_K"]
    pref = Xi ** 2 * kBT / (p["rho"] * p["v_s"] ** 2)
    integ = np.einsum("g,ga,gc,gd,gb->acdb", g["wx"], chi.conj(), chi, chi.conj(), chi) / L0_M
    K = pref * integ / (E0_J ** 2 * L0_M ** 2)
    qK = float(np.max(np.abs(K)))
    return K / qK, qK


def momentum_independent_block(K4):
    """The device's momentum-independent block[(a,b),(c,d)] (first index
    fastest) of K[a,c,d,b]: Sigma_ab = sum_cd K[a,c,d,b] G_cd (kernels.jl:415-435)."""
    Nb = K4.shape[0]
    return np.ascontiguousarray(K4.transpose(0, 3, 1, 2)).reshape(Nb * Nb, Nb * Nb, order="F")


def tutorial_problem():
    """Everything the SCBA loop of the acoustic-only tutorial needs, in the
    solver's units (U = 0: the first inner loop of the outer Poisson loop)."""
    p, n, o = reference_parameters(), tutorial_numerics(), tutorial_options()
    g = build_grids(p, n)
    prof = build_profiles(p, n, g)
    basis = build_localized_basis(p, n, g, prof)
    F = p["F_V_per_m"] / (E0_EV / L0_M)
    h = project_hamiltonians(basis, g, F)
    K4, qK = acoustic_kernel(p, g, basis)
    kbt = KB * p["T_K"] / QE / E0_EV
    return dict(params=p, numerics=n, options=o, grids=g, profiles=prof, basis=basis, F=F, Ep=F * g["Lp"], h=h,
                acoustic=momentum_independent_block(K4), acoustic_K4=K4, qK=qK, kbt=kbt,
                eta_seed=n["eta_seed_eV"] / E0_EV, target=prof["ND2D"], J0_A_cm2=J0_A_M2 * 1e-4)
