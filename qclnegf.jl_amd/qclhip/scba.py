"""Host driver of the production SCBA loop over the device operators.

Mirrors solve_scba_production (src/numerics/optimized/production_solver.jl:
21-846) with linear or Anderson mixing: per iteration the phases :dyson,
:candidate_scattering, :candidate_embedding, :observables and :mixing run on
the device in that order; only scalars (lambda, J, round-off, populations)
come back to the host.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import api


@dataclass
class Mechanism:
    """One enabled scattering mechanism (problem.kernels.enabled order)."""
    name: str
    kernel: api.ProductionKernel
    qK: float
    is_lo: bool = False


def _compact_mechanism(m) -> bool:
    """A mechanism whose self-energy is the same at every momentum: a
    momentum-independent kernel outside the LO map."""
    return not m.is_lo and m.kernel.kind == "momentum_independent"


def _mi_candidate(p, mech, green, fam=None, parts=(0, 1, 2)):
    """Candidate of a momentum-independent mechanism: S< and S> contracted to
    the one column set that every momentum holds and S^R transformed once on
    it: bit for bit the values of the full-field route (production_operators.jl:
    535-595, 842-884), whose transform repeats each column for every
    momentum. Returns the round-off
    witness (None without verification). The columns stay in
    p.mi_compact[mech.name] (R, L, G) for a mixing pass that reads them for
    every momentum; with fam, the components in parts are broadcast into it."""
    g = p.grid
    YR, YL, YG = p.mi_compact[mech.name]
    api.momentum_independent_contraction_compact(mech.kernel, green.GL, mech.qK, Y=YL)
    api.momentum_independent_contraction_compact(mech.kernel, green.GG, mech.qK, Y=YG)
    r = None
    if p.verify_roundoff:
        _, _, r = api.retarded_self_energy_fft(p.hilbert_compact, YL, YG, return_roundoff=True, SR=YR)
    else:
        api.retarded_self_energy_fft(p.hilbert_compact, YL, YG, SR=YR)
    if fam is not None:
        for c in parts:
            api.broadcast_momenta(g, (YR, YL, YG)[c], fam.parts()[c])
    return r


@dataclass
class ScbaProblem:
    grid: api.Grid
    h: torch.Tensor                 # device (Nb,Nb,Nk) = reference (Nk,Nb,Nb)
    Tplus: torch.Tensor             # device Nb x Nb (column-major)
    Tminus: torch.Tensor
    mechanisms: list
    target: float                   # N_D^2D in solver units
    Ep: float                       # bias drop per period
    hwLO: float = 0.36
    NLO: float = 0.0                # thermal LO occupation 1/expm1(hw/kT)
    alpha_mix: float = 0.3
    verify_roundoff: bool = True
    measure: bool = True            # the fixed-point measurement of every iteration
    mixing: str = "linear"          # "linear" or "anderson" (production_solver.jl:808-832)
    anderson_depth: int = 5
    anderson_damping: float = 1.0
    anderson_regularization: float = 1e-12

    def __post_init__(self):
        g = self.grid
        self.plusE = api.EnergyShiftPlan(g, +self.Ep)
        self.minusE = api.EnergyShiftPlan(g, -self.Ep)
        self.plusLO = api.EnergyShiftPlan(g, +self.hwLO)
        self.minusLO = api.EnergyShiftPlan(g, -self.hwLO)
        self.hilbert = api.FFTHilbertPlan(g, g.ncols, verify_roundoff=self.verify_roundoff)
        # Momentum-independent mechanisms: one momentum's columns, their
        # transform (each column standing for Nk in the witness) and buffers.
        self.hilbert_compact, self.mi_compact = None, None
        if any(_compact_mechanism(m) for m in self.mechanisms):
            self.hilbert_compact = api.FFTHilbertPlan(g, g.Nb * g.Nb, verify_roundoff=self.verify_roundoff)
            self.hilbert_compact.set_multiplicity(g.Nk)
            self.mi_compact = {m.name: tuple(api.compact_field(g) for _ in range(3))
                               for m in self.mechanisms if _compact_mechanism(m)}


@dataclass
class Family:
    R: torch.Tensor
    L: torch.Tensor
    G: torch.Tensor

    def parts(self):
        return (self.R, self.L, self.G)


def new_family(grid):
    return Family(api.zeros_field(grid), api.zeros_field(grid), api.zeros_field(grid))


@dataclass
class ScbaState:
    """SCBA state. The embedding total plus + minus is never stored: the Dyson
    kernels sum the families in their loads in the reference order
    ((plus + minus) + mechanism_1 + ...), and `embedding()` forms it on request."""
    green: api.GreenState
    scattering: dict
    plus: Family
    minus: Family
    history: list = field(default_factory=list)

    def embedding(self) -> Family:
        g = self.plus.R
        out = Family(torch.empty_like(g), torch.empty_like(g), torch.empty_like(g))
        for a, b, c in zip(self.plus.parts(), self.minus.parts(), out.parts()):
            api.sum_fields([a, b], c)
        return out


def _candidate_scattering(p: ScbaProblem, green: api.GreenState, out: dict, compact_rg: bool = False):
    """_scattering_candidate_production (production_operators.jl:1258-1335).
    Returns the largest round-off witness (a list of Deferred slots in
    deferred-results mode). compact_rg: a momentum-independent mechanism's
    S^R and S> stay one-momentum columns (p.mi_compact), for a mixing pass
    that reads them so; only its S< (read by the residual suite) is
    broadcast."""
    roundoff = [] if p.grid.ctx.deferring else 0.0
    for mech in p.mechanisms:
        fam = out[mech.name]
        if _compact_mechanism(mech):
            r = _mi_candidate(p, mech, green, fam, (1,) if compact_rg else (0, 1, 2))
            if r is not None:
                roundoff = roundoff + [r] if isinstance(roundoff, list) else max(roundoff, r)
            continue
        if mech.is_lo:
            api.production_lo_contraction(mech.kernel, green.GL, green.GG, mech.qK, p.NLO, p.plusLO,
                                          p.minusLO, SL=fam.L, SG=fam.G)
        else:
            api.production_static_contraction(mech.kernel, green.GL, mech.qK, S=fam.L)
            api.production_static_contraction(mech.kernel, green.GG, mech.qK, S=fam.G)
        if p.verify_roundoff:
            _, _, r = api.retarded_self_energy_fft(p.hilbert, fam.L, fam.G, return_roundoff=True, SR=fam.R)
            roundoff = roundoff + [r] if isinstance(roundoff, list) else max(roundoff, r)
        else:
            api.retarded_self_energy_fft(p.hilbert, fam.L, fam.G, SR=fam.R)
    return roundoff


def _candidate_embedding(p: ScbaProblem, green: api.GreenState, plus: Family, minus: Family):
    """_embedding_self_energy_production (production_operators.jl:924-941)."""
    g = p.grid
    for X, pk, mk in ((green.GR, plus.R, minus.R), (green.GL, plus.L, minus.L), (green.GG, plus.G, minus.G)):
        api.embedding_component_production(g, p.Tplus, X, p.plusE, S=pk)
        api.embedding_component_production(g, p.Tminus, X, p.minusE, S=mk)


def initial_state(p: ScbaProblem, green: api.GreenState) -> ScbaState:
    """Fresh candidates from a seed state become the first self-energies
    (production_solver.jl:117-141)."""
    g = p.grid
    scat = {m.name: new_family(g) for m in p.mechanisms}
    _candidate_scattering(p, green, scat)
    plus, minus = new_family(g), new_family(g)
    _candidate_embedding(p, green, plus, minus)
    return ScbaState(green, scat, plus, minus)


class ScbaWorkspace:
    """Candidate buffers reused across iterations."""

    def __init__(self, p: ScbaProblem):
        g = p.grid
        self.scat = {m.name: new_family(g) for m in p.mechanisms}
        self.plus = new_family(g)
        self.minus = new_family(g)
        # the total self-energy of the iteration: written by the GF solve
        # when it sums the families, or by the previous linear mixing pass
        # (which mixes every family anyway); the residual suite reads it.
        # totals_valid: it holds the sum of the current state's families, so
        # the GF solve loads it as one family. A caller that changes the
        # state's families between iterations clears the flag.
        self.totals = new_family(g) if p.measure else None
        self.totals_valid = False
        # raw-product screen words of the GF solve, read by the residual suite
        self.raw_screen = api.screen_buffer(g) if p.measure else None
        self.anderson = None
        if p.mixing == "anderson":
            from .anderson import AndersonMixer
            self.anderson = AndersonMixer(g, p.anderson_depth, p.anderson_damping, p.anderson_regularization)
        elif p.mixing != "linear":
            raise ValueError("mixing must be 'linear' or 'anderson'")


MEASURE_KEYS = ("r_D", "r_A", "r_K", "r_PSD", "r_caus")


def _screen_buffer(g, ws):
    """The screen words the GF solve writes for the residual suite, where its
    Keldysh products run in the FP64 vector kernels (else None: the suite
    forms the raw products itself)."""
    if ws.raw_screen is None or g.engine != "vector" or g.precision == "fp32_products" or g.Nb > 32:
        return None
    return ws.raw_screen


def scba_iteration(p: ScbaProblem, s: ScbaState, ws: ScbaWorkspace, timings: dict | None = None):
    """One pass of the production loop (production_solver.jl:435-844): Dyson,
    candidates, the fixed-point measurement (production_iteration.jl:102-166:
    residual suite before the observables, the self-energy residuals fused
    into the mixing pass, which reads the same unmixed pairs), observables,
    linear mixing."""
    g = p.grid

    def stamp():
        # with timings, each phase ends when its device work does
        if timings is not None and not g.ctx.deferring:
            g.ctx.synchronize()
        return time.perf_counter()

    t0 = stamp()
    # The Green state keeps no A: the residual suite forms A = i(G^R - G^R^dagger)
    # from its G^R tile (bitwise the stored spectral function), and nothing
    # else in the iteration reads it (api.spectral_function gives it on demand).
    if s.green is not None:
        s.green.A = None
    rs = _screen_buffer(g, ws) if p.measure else None
    if ws.totals is not None and ws.totals_valid:
        green, info = api.current_green_production(g, p.h, [ws.totals.parts()], p.target, state=s.green,
                                                   spectral=False, raw_screen=rs)
    else:
        fams = [s.plus.parts(), s.minus.parts()] + [s.scattering[m.name].parts() for m in p.mechanisms]
        green, info = api.current_green_production(g, p.h, fams, p.target, state=s.green, spectral=False,
                                                   totals=ws.totals.parts() if ws.totals is not None else None,
                                                   raw_screen=rs)
        ws.totals_valid = ws.totals is not None
    s.green = green
    t1 = stamp()
    # linear mixing with totals reads a momentum-independent mechanism's S^R
    # and S> candidates as one-momentum columns
    compact_rg = ws.anderson is None and p.measure
    roundoff = _candidate_scattering(p, green, ws.scat, compact_rg)
    t2 = stamp()
    _candidate_embedding(p, green, ws.plus, ws.minus)
    t3 = stamp()
    suite = None
    if p.measure:
        # the unmixed candidate total S<, summed by the suite as it loads the
        # candidates, in the reference's order ((plus + minus) + mechanism 1 + ...)
        suite = api.residual_suite(g, p.h, green, [ws.totals.parts()],
                                   [ws.plus.L, ws.minus.L] + [ws.scat[m.name].L for m in p.mechanisms],
                                   raw_screen=rs)
    t4 = stamp()
    Jbar = api.boundary_flux_bar(g, s.plus.L, s.plus.G, green.GL, green.GG)
    pops = api.state_populations(g, green.GL, p.target)
    t5 = stamp()
    pairs, compact = [], []
    for m in p.mechanisms:
        cand = ws.scat[m.name].parts()
        cflags = (0, 0, 0)
        if compact_rg and _compact_mechanism(m):
            YR, _, YG = p.mi_compact[m.name]
            cand, cflags = (YR, cand[1], YG), (1, 0, 1)
        pairs += list(zip(s.scattering[m.name].parts(), cand))
        compact += cflags
    base = len(pairs)
    compact += [0] * 6
    for fam, cand in ((s.plus, ws.plus), (s.minus, ws.minus)):
        pairs += list(zip(fam.parts(), cand.parts()))
    sigma = None
    anderson_info = None
    # plus X pairs with minus X: their sum is the embedding term
    partner = [-1] * base + [base + 3, base + 4, base + 5, base, base + 1, base + 2]
    if ws.anderson is not None:
        if g.ctx.deferring:
            raise ValueError("Anderson mixing synchronizes with the host and cannot be deferred or captured")
        if p.measure:
            sigma = api.mix_fields_residual(pairs, 0.0, partner)  # residuals only, before the update
        anderson_info = ws.anderson.mix([o for o, _ in pairs], [c for _, c in pairs], p.alpha_mix,
                                        len(p.mechanisms) + 2)
        ws.totals_valid = False
    elif p.measure:
        # the next solve's totals, per component in the family order plus,
        # minus, mechanisms (pairs: mechanisms first, then plus, then minus)
        M = len(p.mechanisms)
        members = [[base + c, base + 3 + c] + [3 * k + c for k in range(M)] for c in range(3)]
        sigma = api.mix_fields_residual(pairs, p.alpha_mix, partner,
                                        totals=list(zip(ws.totals.parts(), members)),
                                        compact=compact if any(compact) else None, grid=g)
    else:
        api.mix_fields(pairs, p.alpha_mix)
    if g.ctx.deferring:  # captured or deferred: results are slots of the result buffer
        return dict(info=info, Jbar=Jbar, roundoff=roundoff, populations=pops, suite=suite, sigma=sigma)
    g.ctx.synchronize()
    t6 = time.perf_counter()
    if timings is not None:
        for k, v in (("dyson", t1 - t0), ("candidate_scattering", t2 - t1), ("candidate_embedding", t3 - t2),
                     ("residuals", t4 - t3), ("observables", t5 - t4), ("mixing", t6 - t5)):
            timings[k] = timings.get(k, 0.0) + v
    row = dict(lam=info["lam"], number=info["number"], holes=info["holes"], Jbar=Jbar, roundoff=roundoff,
               populations=pops)
    if p.measure:
        row.update(suite)
        row["r_Sigma"] = max(sigma)
        row["r_lambda"] = abs(info["lam"] - 1.0)
    if anderson_info is not None:
        row["anderson"] = anderson_info
    s.history.append(row)
    return row


class LeanWorkspace:
    """Buffers of the memory-lean iteration: three candidate fields (one
    family at a time), the running candidate S< sum of the residual suite and
    the family totals. With M mechanisms the iteration then holds
    3 (G^R, G<, G>) + 3(M + 2) (state) + 3 (totals) + 3 + 1 fields instead of
    the reference's 3M + 15 candidate + 4 green + 3M + 9 stored ones
    (production_cache.jl:570-584): 31 fields at M = 5."""

    def __init__(self, p: ScbaProblem):
        g = p.grid
        if p.mixing != "linear":
            raise ValueError("the lean iteration mixes linearly, one family at a time")
        self.tmp = [api.empty_field(g) for _ in range(3)]
        self.csum = api.empty_field(g)
        self.totals = new_family(g)
        self.totals_valid = False
        self.raw_screen = api.screen_buffer(g)


def _lean_tail(p: ScbaProblem, s: ScbaState, ws: LeanWorkspace, green, info, stamp=None, raw_screen=None):
    """The lean iteration after its GF solve: the state residuals and
    observables, then each family's candidate built, measured and mixed at
    once (embedding plus/minus one component at a time, then every
    mechanism), then r_K. Values are host numbers, or Deferred slots in
    deferred-results mode."""
    g = p.grid
    stamp = stamp or (lambda: 0.0)
    suite = api.residual_state(g, p.h, green, [ws.totals.parts()], raw_screen=raw_screen)
    Jbar = api.boundary_flux_bar(g, s.plus.L, s.plus.G, green.GL, green.GG)
    pops = api.state_populations(g, green.GL, p.target)
    t2 = stamp()
    ratios = []
    totals = ws.totals.parts()
    tmp = ws.tmp
    for c, X in enumerate((green.GR, green.GL, green.GG)):  # embedding, one component at a time
        api.embedding_component_production(g, p.Tplus, X, p.plusE, S=tmp[0])
        api.embedding_component_production(g, p.Tminus, X, p.minusE, S=tmp[1])
        tot = [(totals[c], [0, 1])]
        if c == 1:
            tot.append((ws.csum, [0, 1], "candidate"))
        pairs = [(s.plus.parts()[c], tmp[0]), (s.minus.parts()[c], tmp[1])]
        ratios.append(api.mix_fields_residual(pairs, p.alpha_mix, [1, 0], totals=tot))
    t3 = stamp()
    roundoff = []
    cand = Family(tmp[0], tmp[1], tmp[2])
    for mech in p.mechanisms:
        compact = _compact_mechanism(mech)
        if compact:
            r = _mi_candidate(p, mech, green)  # mixed from its one-momentum columns
        else:
            if mech.is_lo:
                api.production_lo_contraction(mech.kernel, green.GL, green.GG, mech.qK, p.NLO, p.plusLO,
                                              p.minusLO, SL=cand.L, SG=cand.G)
            else:
                api.production_static_contraction(mech.kernel, green.GL, mech.qK, S=cand.L)
                api.production_static_contraction(mech.kernel, green.GG, mech.qK, S=cand.G)
            r = None
            if p.verify_roundoff:
                _, _, r = api.retarded_self_energy_fft(p.hilbert, cand.L, cand.G, return_roundoff=True,
                                                       SR=cand.R)
            else:
                api.retarded_self_energy_fft(p.hilbert, cand.L, cand.G, SR=cand.R)
        if r is not None:
            roundoff.append(r)
        fam = s.scattering[mech.name]
        pairs = list(zip(fam.parts(), p.mi_compact[mech.name] if compact else cand.parts()))
        tot = [(totals[0], [0], "accumulate"), (totals[1], [1], "accumulate"), (totals[2], [2], "accumulate"),
               (ws.csum, [1], "accumulate,candidate")]
        ratios.append(api.mix_fields_residual(pairs, p.alpha_mix, [-1, -1, -1], totals=tot,
                                              compact=[1, 1, 1] if compact else None, grid=g))
    t4 = stamp()
    r_K = api.residual_keldysh(g, green, [ws.csum])
    ws.totals_valid = True
    return dict(info=info, suite=suite, Jbar=Jbar, populations=pops, ratios=ratios, roundoff=roundoff, r_K=r_K,
                stamps=(t2, t3, t4))


def _lean_row(res, host=None) -> dict:
    """The iteration row from _lean_tail's values (host reads the deferred
    ones from the copied result buffer)."""
    rd = lambda x: x.read(host) if isinstance(x, api.Deferred) else x
    info = res["info"]
    if isinstance(info, api.Deferred):
        v = info.read(host)
        info = dict(lam=float(v[0]), number=float(v[1]), holes=float(v[2]), a=float(v[3]), c=float(v[4]))
    suite = res["suite"]
    if isinstance(suite, api.Deferred):
        suite = dict(zip(MEASURE_KEYS, (float(x) for x in suite.read(host))))
    r_K = res["r_K"]
    r_K = float(r_K.read(host)[2]) if isinstance(r_K, api.Deferred) else r_K
    ratios = [float(x) for r in res["ratios"] for x in np.atleast_1d(rd(r))]
    row = dict(lam=info["lam"], number=info["number"], holes=info["holes"], Jbar=float(rd(res["Jbar"])),
               roundoff=max([float(rd(r)) for r in res["roundoff"]], default=0.0),
               populations=np.asarray(rd(res["populations"]), dtype=np.float64))
    row.update(suite)
    row["r_K"] = r_K
    row["r_Sigma"] = max(ratios)
    row["r_lambda"] = abs(info["lam"] - 1.0)
    return row


def scba_iteration_lean(p: ScbaProblem, s: ScbaState, ws: LeanWorkspace, timings: dict | None = None):
    """One pass of the production loop (production_solver.jl:435-844) that
    never holds every candidate: the fixed-point measurement of the state
    (every residual but r_K) and the observables run right after the GF
    solve; then each family's candidate is built (embedding plus/minus one
    component at a time, then every mechanism), measured and mixed at once
    into the state, into the totals of the next solve and into the candidate
    S< sum; r_K runs last on that sum. Every field and scalar equals the
    standard iteration's bit for bit (same kernels, same folds); the Green
    state keeps no A (formed from G^R where the residuals need it). In
    deferred-results mode it returns the _lean_tail values (for capture)."""
    g = p.grid
    if not p.measure:
        raise ValueError("the lean iteration always measures")

    def stamp():
        if timings is not None and not g.ctx.deferring:
            g.ctx.synchronize()
        return time.perf_counter()

    t0 = stamp()
    if s.green is not None:
        s.green.A = None  # not kept
    rs = _screen_buffer(g, ws)
    if ws.totals_valid:
        green, info = api.current_green_production(g, p.h, [ws.totals.parts()], p.target, state=s.green,
                                                   spectral=False, raw_screen=rs)
    else:
        fams = [s.plus.parts(), s.minus.parts()] + [s.scattering[m.name].parts() for m in p.mechanisms]
        green, info = api.current_green_production(g, p.h, fams, p.target, state=s.green, spectral=False,
                                                   totals=ws.totals.parts(), raw_screen=rs)
    s.green = green
    t1 = stamp()
    res = _lean_tail(p, s, ws, green, info, stamp, raw_screen=rs)
    if g.ctx.deferring:
        return res
    g.ctx.synchronize()
    t5 = time.perf_counter()
    if timings is not None:
        t2, t3, t4 = res["stamps"]
        for k, v in (("dyson", t1 - t0), ("residuals_state_observables", t2 - t1),
                     ("candidate_embedding_mix", t3 - t2), ("candidate_scattering_mix", t4 - t3),
                     ("residual_keldysh", t5 - t4)):
            timings[k] = timings.get(k, 0.0) + v
    row = _lean_row(res)
    s.history.append(row)
    return row


class LeanBatch:
    """A wave of independent instances (bias points or structures of uneven
    N_b and N_E) iterated together: ONE batched GF solve over all of them
    (qclhip_batch_current_green: one persistent launch per kernel and tile
    class, or the per-instance launches when the batch autotuner finds them
    faster), then each instance's lean tail. capture() records a whole
    wave-iteration as one HIP graph (deferred results), replay() runs it.
    No counterpart in the reference, which runs one solve_production per
    instance (src/QCLNEGF.jl:982-983). Every instance must have run one lean
    iteration first (its totals and Green state exist)."""

    def __init__(self, problems, states, workspaces):
        self.ps, self.ss, self.wss = list(problems), list(states), list(workspaces)
        if not all(ws.totals_valid for ws in self.wss):
            raise ValueError("run one lean iteration per instance before batching")
        self.ctx = self.ps[0].grid.ctx
        self.batch = api.Batch([p.grid for p in self.ps])
        self.graph = None

    def _gf_args(self):
        return ([p.h for p in self.ps], [[ws.totals.parts()] for ws in self.wss], [p.target for p in self.ps],
                [s.green for s in self.ss])

    def autotune(self):
        hs, fams, tg, greens = self._gf_args()
        return api.batch_autotune(self.batch, hs, fams, tg, greens)

    def _iterate(self):
        hs, fams, tg, greens = self._gf_args()
        _, infos = api.batch_current_green(self.batch, hs, fams, tg, states=greens)
        out = []
        for i, (p, s, ws) in enumerate(zip(self.ps, self.ss, self.wss)):
            info = api.Deferred(infos.slot + 5 * i, 5) if isinstance(infos, api.Deferred) else infos[i]
            out.append(_lean_tail(p, s, ws, s.green, info))
        return out

    def iterate(self):
        """One eager batched iteration; returns the instances' rows."""
        res = self._iterate()
        self.ctx.synchronize()
        rows = [_lean_row(r) for r in res]
        for s, r in zip(self.ss, rows):
            s.history.append(r)
        return rows

    def capture(self):
        """Runs one eager batched iteration (stages the instance tables and
        sizes the scratch), then captures the next as one graph."""
        self.iterate()
        nres = sum(5 + 5 + 1 + p.grid.Nb + 9 + 4 * len(p.mechanisms) + 5 for p in self.ps)
        self.buffer = torch.zeros(nres + 8, dtype=torch.float64, device=self.ctx.device)
        self.stream = torch.cuda.Stream(self.ctx.device)
        torch.cuda.synchronize()
        base = torch.cuda.current_stream(self.ctx.device)
        with torch.cuda.stream(self.stream):
            self.ctx.set_stream(self.stream)
            try:
                self.ctx.defer(self.buffer)
                self.ctx.graph_begin()
                try:
                    self.slots = self._iterate()
                finally:
                    self.graph = self.ctx.graph_end()
                    self.ctx.defer(None)
            finally:
                self.ctx.set_stream(base)
        return self

    def replay(self):
        base = torch.cuda.current_stream(self.ctx.device)
        self.stream.wait_stream(base)
        self.ctx.set_stream(self.stream)
        try:
            self.graph.launch()
            self.ctx.synchronize()
        finally:
            self.ctx.set_stream(base)
        base.wait_stream(self.stream)
        host = self.buffer.cpu().numpy()
        rows = [_lean_row(r, host) for r in self.slots]
        for s, r in zip(self.ss, rows):
            s.history.append(r)
        return rows


def solve_scba_production(p: ScbaProblem, s: ScbaState, max_iterations: int, tol_population: float = 0.0):
    """Run up to max_iterations passes; stops early when the relative
    population change falls below tol_population (when > 0)."""
    ws = ScbaWorkspace(p)
    prev = None
    for _ in range(max_iterations):
        row = scba_iteration(p, s, ws)
        if prev is not None and tol_population > 0:
            change = np.linalg.norm(row["populations"] - prev) / (np.linalg.norm(row["populations"]) + 1e-14)
            row["r_population"] = change
            if change < tol_population:
                break
        prev = row["populations"]
    return s


class IterationGraph:
    """One production iteration captured as a HIP graph and replayed: for the
    small, launch-bound problems (tens of kernels of a few microseconds per
    iteration). Construction runs one eager iteration (it sizes the scratch
    and allocates the Green state), then captures the next iteration on a
    side stream in deferred-results mode; capture itself runs nothing.
    replay() runs one iteration and returns its row."""

    def __init__(self, p: ScbaProblem, s: ScbaState, ws: ScbaWorkspace):
        self.p, self.s, self.ws = p, s, ws
        ctx = p.grid.ctx
        scba_iteration(p, s, ws)
        nres = 5 + 1 + p.grid.Nb + len(p.mechanisms) + 5 + 3 * len(p.mechanisms) + 9 + 8
        self.buffer = torch.zeros(nres, dtype=torch.float64, device=ctx.device)
        self.stream = torch.cuda.Stream(ctx.device)
        torch.cuda.synchronize()
        base = torch.cuda.current_stream(ctx.device)
        with torch.cuda.stream(self.stream):
            ctx.set_stream(self.stream)
            try:
                ctx.defer(self.buffer)
                ctx.graph_begin()
                try:
                    self.slots = scba_iteration(p, s, ws)
                finally:
                    self.graph = ctx.graph_end()
                    ctx.defer(None)
            finally:
                ctx.set_stream(base)

    def replay(self):
        ctx = self.p.grid.ctx
        base = torch.cuda.current_stream(ctx.device)
        self.stream.wait_stream(base)
        ctx.set_stream(self.stream)
        try:
            self.graph.launch()
            ctx.synchronize()
        finally:
            ctx.set_stream(base)
        base.wait_stream(self.stream)
        host = self.buffer.cpu().numpy()
        info = self.slots["info"].read(host)
        ro = self.slots["roundoff"]
        row = dict(lam=info[0], number=info[1], holes=info[2], Jbar=self.slots["Jbar"].read(host),
                   roundoff=max([r.read(host) for r in ro]) if isinstance(ro, list) and ro else 0.0,
                   populations=self.slots["populations"].read(host))
        if self.slots["suite"] is not None:
            row.update(zip(MEASURE_KEYS, self.slots["suite"].read(host)))
            row["r_Sigma"] = max(self.slots["sigma"].read(host))
            row["r_lambda"] = abs(info[0] - 1.0)
        self.s.history.append(row)
        return row
