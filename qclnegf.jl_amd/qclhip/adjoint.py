"""Fixed-point adjoint of the production SCBA loop (no reference counterpart:
algorithm_registry.jl:226-249 lists the adjoint as missing).

At a converged state s* = F(s*) (F: current_green, then the scattering and
embedding candidates of production_solver.jl:435-844) the current functional
J = Jbar(s*.plus, G(s*)) (production_iteration.jl:189-201) has the gradient

    dJ/dh = dJ/dh|_G + (dF/dh)^H mu,   mu = (dF/ds)^H mu + dJ/ds,

and mu is found by the damped iteration mu <- (1-alpha) mu + alpha[(dF/ds)^H mu
+ dJ/ds] with the forward's mixing alpha, restricted to the tangent space of
the state (S< and S> blocks anti-Hermitian; S^R free). There it converges at
the forward's linear rate (the iteration matrix is the transpose of the
forward's on that space). Every
reverse operator runs on the device through the C-ABI (qclhip_*_adjoint,
qclhip_shift_transpose, qclhip_kernel_adjoint).

Gradient convention: Zbar = dJ/dRe Z + i dJ/dIm Z (dJ = Re sum conj(Zbar) dZ).
"""
from __future__ import annotations

import numpy as np

from . import api
from .scba import Family, ScbaProblem, ScbaState, new_family, scba_iteration, ScbaWorkspace


def converge(p: ScbaProblem, s: ScbaState, tol: float = 1e-15, max_iterations: int = 5000, min_iterations: int = 5):
    """Iterate the forward loop until |dJ| <= tol |J| (J the boundary flux)."""
    ws = ScbaWorkspace(p)
    prev = None
    it = 0
    for it in range(1, max_iterations + 1):
        row = scba_iteration(p, s, ws)
        J = row["Jbar"]
        if prev is not None and it >= min_iterations and abs(J - prev) <= tol * abs(J):
            break
        prev = J
    return it


class _Families:
    """The adjoint state mu, one Family of gradients per SCBA family.

    With shared=True the minus family and every mechanism's family are one
    Family object, and plus holds that object's R: under the linear
    iteration every family receives the same new value (sig; plus: sig +
    dJ/dplus in L and G) from the same zero start, so these stay equal to
    the last bit and one copy carries them all."""

    def __init__(self, p: ScbaProblem, shared: bool = False):
        g = p.grid
        self.shared = shared
        if shared:
            common = new_family(g)
            self.plus = Family(common.R, api.zeros_field(g), api.zeros_field(g))
            self.minus = common
            self.scat = {m.name: common for m in p.mechanisms}
        else:
            self.plus = new_family(g)
            self.minus = new_family(g)
            self.scat = {m.name: new_family(g) for m in p.mechanisms}

    def all(self):
        return [self.plus, self.minus] + list(self.scat.values())


class ScbaAdjoint:
    """Reverse operators of one problem: adjoint kernels built once."""

    def __init__(self, p: ScbaProblem):
        self.p = p
        self.kadj = {m.name: m.kernel.adjoint() for m in p.mechanisms}

    # -- reverse of the candidates: mu (families) -> (GRbar, GLbar, GGbar)
    def candidates_vjp(self, mu: _Families, out):
        p, g = self.p, self.p.grid
        GRb, GLb, GGb = out
        tmpL, tmpG, Z = api.empty_field(g), api.empty_field(g), api.empty_field(g)
        for k, m in enumerate(p.mechanisms):
            f = mu.scat[m.name]
            # S<_bar, S>_bar of the mechanism's candidate: its L, G gradients
            # plus the reverse of S^R = H[Gamma] - i Gamma/2 (once for a
            # shared family: the same input gives the same fields)
            if k == 0 or not mu.shared:
                api.copy_field(f.L, tmpL)
                api.copy_field(f.G, tmpG)
                api.retarded_self_energy_adjoint(p.hilbert, f.R, tmpL, tmpG)
            K = self.kadj[m.name]
            if m.is_lo:
                # S< = qK[(N+1) S+ K[G<] + N S- K[G<]],  S> = qK[(N+1) S- K[G>] + N S+ K[G>]]
                api.shift_transpose(p.plusLO, tmpL, Z, alpha=p.NLO + 1.0)
                api.shift_transpose(p.minusLO, tmpL, Z, alpha=p.NLO, accumulate=True)
                api.production_static_contraction(K, Z, m.qK, beta=1.0, S=GLb)
                api.shift_transpose(p.minusLO, tmpG, Z, alpha=p.NLO + 1.0)
                api.shift_transpose(p.plusLO, tmpG, Z, alpha=p.NLO, accumulate=True)
                api.production_static_contraction(K, Z, m.qK, beta=1.0, S=GGb)
            else:
                api.production_static_contraction(K, tmpL, m.qK, beta=1.0, S=GLb)
                api.production_static_contraction(K, tmpG, m.qK, beta=1.0, S=GGb)
        for T, plan, f in ((p.Tplus, p.plusE, mu.plus), (p.Tminus, p.minusE, mu.minus)):
            api.embedding_adjoint(g, T, plan, f.R, GRb)
            api.embedding_adjoint(g, T, plan, f.L, GLb)
            api.embedding_adjoint(g, T, plan, f.G, GGb)

    def gradient(self, s: ScbaState, tol: float = 1e-13, max_iterations: int = 5000, alpha: float | None = None,
                 acceleration: str = "linear", depth: int = 5, checkpoints=(), share_families: bool = True):
        """dJ/dh (device tensor (Nb,Nb,Nk), the reference layout of h) at the
        converged state s, plus the adjoint state and iteration count.
        acceleration: "linear" (the damped iteration at the forward's rate)
        or "anderson" (Anderson acceleration of the same linear fixed point:
        the forward's mixer without its self-energy guard, restarts or
        coefficient bound). checkpoints: iteration counts at which the
        gradient of the current adjoint state is also formed; they are
        returned as meta["checkpoints"] = [(iteration, dJ/dh), ...].
        share_families: carry the families that the linear iteration keeps
        equal in one copy (the same fields; the change measure sums the
        same terms in another grouping). Anderson acceleration mixes every
        family on its own."""
        p, g = self.p, self.p.grid
        alpha = p.alpha_mix if alpha is None else alpha
        if acceleration not in ("linear", "anderson"):
            raise ValueError("acceleration must be 'linear' or 'anderson'")
        mixer = None
        if acceleration == "anderson":
            from .anderson import AndersonMixer
            mixer = AndersonMixer(g, depth=depth)
        fams = [s.plus.parts(), s.minus.parts()] + [s.scattering[m.name].parts() for m in p.mechanisms]
        green, info = api.current_green_production(g, p.h, fams, p.target, spectral=False)
        # dJ/dG (constant) and the direct dJ/dplus.
        JL, JG = api.zeros_field(g), api.zeros_field(g)
        dplusL, dplusG = api.zeros_field(g), api.zeros_field(g)
        api.boundary_flux_adjoint(g, 1.0, s.plus.L, s.plus.G, green.GL, green.GG, dplusL, dplusG, JL, JG)
        api.antihermitian_part(g, dplusL)
        api.antihermitian_part(g, dplusG)
        mu = _Families(p, shared=share_families and mixer is None)
        GRb, GLb, GGb = api.empty_field(g), api.empty_field(g), api.empty_field(g)
        sig = (api.empty_field(g), api.empty_field(g), api.empty_field(g))
        plus_new = (api.empty_field(g), api.empty_field(g))  # sig + the direct dJ/dplus
        it, change = 0, float("inf")
        history, snaps = [], []

        def dh_of_mu():
            # dJ/dh of the current adjoint state (dJ/dh|_G + (dF/dh)^H mu)
            api.fill_zero(GRb)
            api.copy_field(JL, GLb)
            api.copy_field(JG, GGb)
            self.candidates_vjp(mu, (GRb, GLb, GGb))
            _, hb = api.current_green_adjoint(g, fams, p.target, info, green, GLb, GGb, GRbar=GRb, want_hbar=True,
                                              out=sig)
            return hb

        for it in range(1, max_iterations + 1):
            api.fill_zero(GRb)
            api.copy_field(JL, GLb)
            api.copy_field(JG, GGb)
            self.candidates_vjp(mu, (GRb, GLb, GGb))
            api.current_green_adjoint(g, fams, p.target, info, green, GLb, GGb, GRbar=GRb, want_hbar=False,
                                      out=sig)
            # Project S< and S> onto the anti-Hermitian blocks: the forward
            # state lives there, and outside it the linearized map keeps a
            # neutral mode (eigenvalue 1) that the forward never excites.
            api.antihermitian_part(g, sig[1])
            api.antihermitian_part(g, sig[2])
            api.sum_fields([sig[1], dplusL], plus_new[0])
            api.sum_fields([sig[2], dplusG], plus_new[1])
            # every family's new value is sig (plus: sig + dJ/dplus); mix and
            # measure the change ||new - mu|| / ||new|| in one pass
            if mu.shared:
                # R is carried M + 2 times, the common L and G M + 1 times
                M = len(p.mechanisms)
                nR, dR = api.mix_fields_change([(mu.minus.R, sig[0])], alpha)
                nC, dC = api.mix_fields_change([(mu.minus.L, sig[1]), (mu.minus.G, sig[2])], alpha)
                nP, dP = api.mix_fields_change([(mu.plus.L, plus_new[0]), (mu.plus.G, plus_new[1])], alpha)
                num = (M + 2) * nR + (M + 1) * nC + nP
                den = (M + 2) * dR + (M + 1) * dC + dP
            else:
                pairs = []
                for f in mu.all():
                    newL, newG = (plus_new if f is mu.plus else (sig[1], sig[2]))
                    pairs += [(f.R, sig[0]), (f.L, newL), (f.G, newG)]
                if mixer is None:
                    num, den = api.mix_fields_change(pairs, alpha)
                else:
                    num, den = api.mix_fields_change(pairs, 0.0)  # measure only
                    mixer.mix([o for o, _ in pairs], [c for _, c in pairs], alpha, len(pairs) // 3, guard=False,
                              linear=True)
            change = (num / max(den, 1e-300)) ** 0.5
            if it == 1 or it % 100 == 0:
                history.append((it, change))
            if it in checkpoints:
                snaps.append((it, dh_of_mu().clone()))
            if change <= tol:
                break
        return dh_of_mu(), dict(iterations=it, change=change, info=info, history=history, checkpoints=snaps)


def thickness_gradient(hbar, assemble, thick, step: float = 1e-3):
    """dJ/dL_i of layer thicknesses by the chain rule through h:
    dJ/dL_i = Re sum_m sum_ab conj(hbar_m[a,b]) dh_m[a,b]/dL_i, with dh/dL_i
    from central differences of the host assembly h(L) (assemble: thickness
    vector -> h of shape (Nk, Nb, Nb), e.g. synthetic.period_h on a fixed
    basis). Covers what J sees through h; inputs the assembly keeps fixed
    (basis, kernels, T+-) do not move. hbar: the device gradient of
    ScbaAdjoint.gradient (reference layout (Nb, Nb, Nk))."""
    hb = hbar.cpu().numpy().transpose(2, 1, 0)  # (Nk, Nb, Nb)
    thick = np.asarray(thick, dtype=np.float64)
    out = np.zeros(len(thick))
    for i in range(len(thick)):
        d = np.zeros(len(thick))
        d[i] = step
        dh = (assemble(thick + d) - assemble(thick - d)) / (2.0 * step)
        out[i] = float(np.real(np.sum(np.conj(hb) * dh)))
    return out

