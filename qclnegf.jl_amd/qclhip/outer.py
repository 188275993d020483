"""Host driver of the outer Poisson loop over the device operators.

Mirrors the Poisson step of solve_production (production_solver.jl:
1044-1072): after an accepted SCBA run, the electron density of its G<
(_electron_density_bar), the bordered periodic Poisson solve for the
candidate Hartree energy, the outer mixing U <- (1 - alpha_P) U + alpha_P
cand with r_U = max |cand - U|, the density change r_n, and the projected
Hamiltonians of the next SCBA run. Only scalars and the N_z-point density
come back to the host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import api


@dataclass
class OuterProblem:
    """Device copies of the problem data of the Poisson step.

    eps_r, wx, ND: (Nz,) float64; lambda_P: Poisson prefactor; chi, Phi:
    the basis on the position grid (api.basis_to_device, chi = Phi/sqrt(dx));
    S = H0 - F (Z - z0) and Minv (api.matrix_to_device); kin_m = C kappa_m^2;
    ND2D, Lp: the normalization of r_n (sheet doping, period)."""
    eps_r: torch.Tensor
    wx: torch.Tensor
    ND: torch.Tensor
    lambda_P: float
    chi: torch.Tensor
    Phi: torch.Tensor
    S: torch.Tensor
    kin: torch.Tensor
    Minv: torch.Tensor
    gs: float = 2.0
    alpha_P: float = 0.5
    ND2D: float = 1.0
    Lp: float = 1.0


def poisson_step(op: OuterProblem, grid: api.Grid, GL, U, nbar_previous=None):
    """One Poisson step from the G< of an accepted SCBA run; U (device, N_z)
    is mixed in place. Returns a row with the new h, the density and the
    outer residuals (r_P, r_U, r_n, neutrality, zeta)."""
    nbar = api.electron_density(grid, GL, op.chi, op.gs)
    cand, zeta, rP, neutrality = api.periodic_poisson(op.eps_r, op.wx, op.ND, op.lambda_P, nbar)
    rU = api.mix_potential(U, cand, op.alpha_P)
    h = api.project_hamiltonians(op.Phi, U, op.S, op.kin, op.Minv)
    n_host = nbar.cpu().numpy()
    if nbar_previous is None:
        rn = math.inf
    else:
        w = op.wx.cpu().numpy()
        rn = float(np.sqrt(np.sum(w * (n_host - nbar_previous) ** 2)) / (op.ND2D / math.sqrt(op.Lp)))
    return dict(h=h, nbar=n_host, zeta=zeta, r_P=rP, neutrality=neutrality, r_U=rU, r_n=rn)


def solve_outer(p, s, op: OuterProblem, U, max_poisson: int, scba_iterations: int):
    """Alternate SCBA runs (scba_iterations passes each) and Poisson steps,
    the Hamiltonian of each run projected from the mixed Hartree energy."""
    from . import scba
    rows = []
    nprev = None
    for _ in range(max_poisson):
        scba.solve_scba_production(p, s, scba_iterations)
        row = poisson_step(op, p.grid, s.green.GL, U, nprev)
        p.h = row["h"]
        nprev = row["nbar"]
        rows.append(row)
    return rows
