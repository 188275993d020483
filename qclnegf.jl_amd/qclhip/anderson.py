"""Anderson mixing of the SCBA self-energies over the device operators.

Mirrors _anderson_mix! (production_mixing.jl:246-317) with its bounded
history (_anderson_append!, :195-229, eviction before materialization and a
reset when the residual norm grows), the coefficient solve of the small
Gram matrix (_anderson_coefficients, :160-187: regularized, rank-revealed by
a symmetric eigen-decomposition, normalized to sum one; on the host), the
coefficient bound (sum |c| <= 10), the device combination and the device
causality guard with the convex fallback x + damping r.

The fields of one history entry follow _ordered_iteration_families
(:126-130): every mechanism in order, then plus, then minus; R, <, > each.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api
from ._lib import DomainError

_EPS = np.finfo(float).eps


def anderson_coefficients(gram: np.ndarray, regularization: float) -> np.ndarray:
    """_anderson_coefficients (production_mixing.jl:160-187); raises
    ValueError for a rank-deficient or non-finite solve."""
    count = gram.shape[0]
    if count < 1:
        raise ValueError("empty Anderson history")
    if count == 1:
        return np.ones(1)
    g = np.array(gram, dtype=np.float64)
    scale = max(float(np.max(np.abs(np.diag(g)))), _EPS)
    g[np.diag_indices(count)] += regularization * scale
    lam, vec = np.linalg.eigh(g / scale)
    cutoff = max(regularization, np.sqrt(_EPS))
    projected = vec.T @ np.ones(count)
    weights = np.where(lam > cutoff, projected / np.where(lam > cutoff, lam, 1.0), 0.0)
    coef = vec @ weights
    den = float(np.sum(coef))
    if not (np.isfinite(den) and abs(den) > _EPS):
        raise ValueError("Anderson history is rank deficient")
    coef = coef / den
    if not np.all(np.isfinite(coef)):
        raise ValueError("Anderson coefficient solve produced non-finite values")
    return coef


class AndersonMixer:
    """The Anderson workspace: device copies of up to `depth` states and
    residuals (3 fields per family each), the Gram matrix on the host, and
    free lists that recycle evicted entries."""

    def __init__(self, grid: api.Grid, depth: int = 5, damping: float = 1.0, regularization: float = 1e-12):
        if depth < 1 or depth > 16:
            raise ValueError("anderson_history_depth must lie in 1..16")
        if not (0.0 < damping <= 1.0):
            raise ValueError("anderson_damping must lie in (0,1]")
        if regularization < 0.0:
            raise ValueError("anderson_regularization cannot be negative")
        self.grid = grid
        self.depth, self.damping, self.regularization = depth, damping, regularization
        self.states, self.residuals = [], []
        self.free_states, self.free_residuals = [], []
        self.gram = np.zeros((0, 0))

    def clear(self):
        self.free_states += self.states
        self.free_residuals += self.residuals
        self.states, self.residuals = [], []
        self.gram = np.zeros((0, 0))

    def _entry(self, free, like):
        return free.pop() if free else [torch.empty_like(t) for t in like]

    def append(self, state, cand, reset_on_increase: bool = True) -> bool:
        prior = None if not self.residuals else float(self.gram[-1, -1])
        if len(self.states) >= self.depth:
            self.free_states.append(self.states.pop(0))
            self.free_residuals.append(self.residuals.pop(0))
            self.gram = self.gram[1:, 1:]
        saved = self._entry(self.free_states, state)
        resid = self._entry(self.free_residuals, state)
        dots = api.anderson_append(state, cand, saved, resid, self.residuals)
        # sum over the fields in family and component order
        inner = [sum(float(dots[f, h]) for f in range(dots.shape[0])) for h in range(dots.shape[1])]
        norm = inner[-1]
        reset = reset_on_increase and prior is not None and norm > (1 + 64 * _EPS) * max(prior, np.finfo(float).tiny)
        if reset:
            self.clear()
            inner = [norm]
        count = len(self.states) + 1
        g = np.zeros((count, count))
        if count > 1:
            g[:-1, :-1] = self.gram
            g[:-1, -1] = g[-1, :-1] = inner[:-1]
        g[-1, -1] = norm
        self.gram = g
        self.states.append(saved)
        self.residuals.append(resid)
        return reset

    def mix(self, state, cand, alpha: float, nfam: int, guard: bool = True, linear: bool = False) -> dict:
        """Replace the state fields (in place) by the accelerated update.
        guard=False skips the self-energy causality guard (for iterates that
        are not self-energies, such as the adjoint state). linear=True is for
        a linear fixed point (the adjoint): no restart when the residual
        grows and no bound on the coefficients, i.e. plain Anderson
        acceleration over the sliding window, a Krylov-type solver there."""
        reset = self.append(state, cand, reset_on_increase=not linear)
        reason = "residual_increase" if reset else "none"
        try:
            coef = anderson_coefficients(self.gram, self.regularization)
        except ValueError:
            reason = "coefficient_solve"
            coef = np.r_[np.zeros(len(self.residuals) - 1), 1.0]
        if (not linear and np.sum(np.abs(coef)) > 10) or not np.all(np.isfinite(coef)):
            reason = "coefficient_bound"
            coef = np.r_[np.zeros(len(self.residuals) - 1), 1.0]
        depth = len(self.states)
        damping = alpha * self.damping
        api.anderson_combine(coef, damping, self.states, self.residuals, state)
        fams = [tuple(state[3 * f:3 * f + 3]) for f in range(nfam)]
        first, code = api.anderson_guard(self.grid, fams) if guard else (0, 0)
        accepted = first == 0 and int(np.count_nonzero(coef)) > 1
        if first != 0:
            api.anderson_combine([1.0], damping, [self.states[-1]], [self.residuals[-1]], state)
            self.clear()
            reason = "causality_guard"
            if api.anderson_guard(self.grid, fams)[0] != 0:
                raise DomainError("convex Anderson fallback is not admissible; inspect raw self-energies")
        return dict(history_depth=depth, reset_reason=reason, coefficients_l1=float(np.sum(np.abs(coef))),
                    coefficients=coef, accepted=accepted, guard_first_failure=first, guard_failure_code=code)
