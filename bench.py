#!/usr/bin/env python3
"""Benchmark of the MI355X SCBA operator backend (one GPU).

A step is one GF-solve pass (A1+A2 of the production loop: Dyson inversion
G^R = (E - h - Sigma^R)^-1, Keldysh products G^R Sigma^{<,>} G^A, number
functionals and the paired normalization) over the synthetic batch
N=24, NE=2048, Nk=64 (131,072 blocks) in FP64, inputs resident in HBM.
value = GF solves/s. The same line carries the SCBA iteration rate of a
full production iteration at the same shape, the roofline of the fused
Dyson/Keldysh kernel (HIP-event timing of that kernel alone) and the CPU
oracle's rate on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "qclnegf.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "GF solves/s + SCBA iter/s, FP32+refine vs FP64 at ≤1e-8 rel. current error"
# HBM bytes per launch pair (k_dyson + k_keldysh) from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes, written by the profile summary below (None if absent).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gf_solve_traffic.txt")


def read_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_MATRIX_TFLOPS = 157.3  # dense FP32 MFMA: the ceiling of the mixed variant's products
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


VARIANTS = (("fp64", "vector"), ("fp64", "matrix"), ("fp32_refine", "vector"), ("fp32_refine", "matrix"),
            ("fp32_products", "matrix"))
CURRENT_TOL = 1e-8


def gf_bench(steps, warmup, variants=VARIANTS, autotune=True):
    """GF-solve step (current_green: Dyson, Keldysh, number functionals,
    paired normalization) on the synthetic batch driven by an injecting
    family, for each (precision, engine) variant: step time, the Dyson/Keldysh
    launches alone (HIP events on their stream), the current functional and
    its error against the FP64 vector path."""
    import numpy as np
    import torch

    from qclhip import api, synthetic as syn

    w = syn.gf_workload()
    N, NE, Nk = w["N"], w["NE"], w["Nk"]
    nblk = NE * Nk
    h = api.h_to_device(w["h"])
    plus = syn.diag_fields_device(syn.injection_family(w))
    scat = syn.diag_fields_device(w)
    fams = [plus, scat]
    SR, SL, SG = scat
    stream = torch.cuda.current_stream()
    out = {}
    for precision, engine in variants:
        name = f"{precision}_{engine}" if precision != "fp32_products" else precision
        log(f"  variant {name}")
        g = api.Grid(w["eps"], w["wE"], w["wk"], N)
        tiles = g.autotune() if autotune else None  # tile shapes of the FP64 vector kernels (cached per shape)
        g.set_engine(engine).set_precision(precision)
        st, info = api.current_green_production(g, h, fams, 1.0, spectral=True)
        target = 0.9 * info["number"]  # exercises the paired map (a = 0.9)

        def step():
            # the full output of _current_green_production: G^R, G<, G>, A
            api.current_green_production(g, h, fams, target, state=st, spectral=True, return_scalars=False)

        for _ in range(warmup):
            step()
        g.ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        g.ctx.synchronize()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        # Witnesses of the timed state: N[G<] equals the target; the current.
        num = api.number_functional(g, st.GL)
        J = api.boundary_flux_bar(g, plus[1], plus[2], st.GL, st.GG)
        fallback = g.refine_fallbacks() / nblk
        traces = torch.empty(4 * nblk, dtype=torch.float64, device="cuda")
        reps = max(5, steps)
        GR2, GL2, GG2 = (torch.empty_like(st.GR) for _ in range(3))
        api.dyson_keldysh(g, h, SR, SL, SG, GR=GR2, GL=GL2, GG=GG2, traces=traces)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            api.dyson_keldysh(g, h, SR, SL, SG, GR=GR2, GL=GL2, GG=GG2, traces=traces)
        e1.record(stream)
        torch.cuda.synchronize()
        g.ctx.synchronize()
        del GR2, GL2, GG2, traces
        out[name] = dict(precision=precision, engine=engine, step_s=dt, kernel_ms=e0.elapsed_time(e1) / reps,
                         tiles=None if tiles is None else dict(dyson_waves=tiles["dyson_waves"],
                                                               keldysh_waves=tiles["keldysh_waves"]),
                         number_check=abs(num - target) / target, J=J, solves_per_s=nblk / dt,
                         fp64_fallback_fraction=fallback)
        del st, g
        torch.cuda.empty_cache()
    # The same FP64 vector launch pair on fields whose planes carry a 256-byte
    # gap (plane stride NE*Nk + 16 elements, off the power of two of this size).
    padded = None
    if "fp64_vector" in out:
        g = api.Grid(w["eps"], w["wE"], w["wk"], N)
        pad = 16
        plane = nblk + pad
        fields = [api.strided_field(g, plane) for _ in range(6)]
        for f, x in zip(fields[:3], (SR, SL, SG)):
            f.copy_(x)
        traces = torch.empty(4 * nblk, dtype=torch.float64, device="cuda")
        api.dyson_keldysh_strided(g, plane, h, *fields, traces=traces)
        reps = max(5, steps)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            api.dyson_keldysh_strided(g, plane, h, *fields, traces=traces)
        e1.record(stream)
        torch.cuda.synchronize()
        g.ctx.synchronize()
        padded = dict(plane_pad_bytes=pad * 16, kernel_ms=e0.elapsed_time(e1) / reps)
        log(f"  padded planes: Dyson+Keldysh {padded['kernel_ms']:.3f} ms")
        del fields, traces, g
        torch.cuda.empty_cache()
    tune = api.Grid(w["eps"], w["wE"], w["wk"], N).autotune() if autotune else None
    log(f"  autotune: {tune}")
    ref = out["fp64_vector"]["J"] if "fp64_vector" in out else None
    for v in out.values():
        v["current_rel_err"] = abs(v["J"] - ref) / abs(ref) if ref else None
    flops = nblk * 40.0 * N ** 3          # 8N^3 inverse + 2 x 16N^3 Keldysh
    bytes_ = nblk * 16.0 * N * N * 6      # read S^R, S<, S>; write G^R, G<, G>
    return dict(N=N, NE=NE, Nk=Nk, nblk=nblk, flops=flops, bytes=bytes_, variants=out, autotune=tune,
                padded=padded)


def pair_only(reps):
    """The FP64 vector Dyson + Keldysh launch pair alone on the GF batch (one
    family), reps times, for kernel counters."""
    import torch

    from qclhip import api, synthetic as syn

    w = syn.gf_workload()
    g = api.Grid(w["eps"], w["wE"], w["wk"], w["N"])
    h = api.h_to_device(w["h"])
    SR, SL, SG = syn.diag_fields_device(w)
    GR, GL, GG = (torch.empty_like(SR) for _ in range(3))
    traces = torch.empty(4 * w["NE"] * w["Nk"], dtype=torch.float64, device="cuda")
    for _ in range(reps):
        api.dyson_keldysh(g, h, SR, SL, SG, GR=GR, GL=GL, GG=GG, traces=traces)
    g.ctx.synchronize()
    torch.cuda.synchronize()
    return dict(reps=reps, N=w["N"], NE=w["NE"], Nk=w["Nk"])


def residuals_only(reps):
    """The residual suite (every term) and its state pass (every term but
    r_K) on the state of one SCBA iteration of the benchmark problem, reps
    times each, for kernel counters."""
    import torch

    from qclhip import api, scba

    p, s = scba_problem()
    ws = scba.ScbaWorkspace(p)
    scba.scba_iteration(p, s, ws)
    g = p.grid
    cands = [ws.plus.L, ws.minus.L] + [ws.scat[m.name].L for m in p.mechanisms]
    out = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = api.residual_suite(g, p.h, s.green, [ws.totals.parts()], cands)
        t1 = time.perf_counter()
        api.residual_state(g, p.h, s.green, [ws.totals.parts()])
        t2 = time.perf_counter()
        out.append(dict(suite_ms=1e3 * (t1 - t0), state_ms=1e3 * (t2 - t1), suite=r))
    return out


def scba_problem(N=24, NE=2048, Nk=64, strength=1.0, coupling=1.0, lo_strength=1.0):
    """The SCBA benchmark problem: seeded synthetic Hamiltonian, three
    mechanisms (momentum-independent, separable IFR, low-rank thermal LO),
    bulk embedding, seed state at a fixed chemical potential. strength
    scales every mechanism's kernel factor (lo_strength the LO one once
    more) and coupling the embedding's T."""
    import numpy as np

    from qclhip import api, scba, synthetic as syn

    w = syn.gf_workload(N=N, NE=NE, Nk=Nk)
    rng = np.random.default_rng(syn.SEED + 7)
    g = api.Grid(w["eps"], w["wE"], w["wk"], N)
    mechs = [
        scba.Mechanism("acoustic", api.ProductionKernel(g, "momentum_independent",
                                                        syn.covariance_block(rng, N, 2, 0.04)), strength),
    ]
    l, r_, rad = syn.separable_factors(rng, N, Nk, J=3, scale=0.02)
    mechs.append(scba.Mechanism("IFR", api.ProductionKernel(g, "separable_ifr", l, r_, rad, rank=3), strength))
    lo_l, lo_r = syn.low_rank_factors(rng, N, Nk, r=4, scale=0.02)
    mechs.append(scba.Mechanism("LO", api.ProductionKernel(g, "low_rank", lo_l, lo_r, rank=4), strength * lo_strength,
                                 True))
    T = 0.05 * coupling * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    h = api.h_to_device(w["h"])
    kbt = w["kbt"]
    hw = 0.36
    # Seed at a fixed chemical potential; its number is the target.
    GR0, _ = api.retarded_green_production(g, h, None, 0.05)
    q = api.seed_number_weights(g, GR0)
    del GR0
    target = float(np.sum(q * api.fermi(w["eps"], w["mu"], kbt)))
    seed_green, mu = api.seed_green_production(g, h, 0.05, kbt, target)
    p = scba.ScbaProblem(g, h, api.matrix_to_device(T), api.matrix_to_device(T.conj().T), mechs, target,
                         Ep=0.5, hwLO=hw, NLO=1.0 / math.expm1(hw / kbt), alpha_mix=0.3)
    s = scba.initial_state(p, seed_green)
    return p, s


def scba_bench(iters, warmup, adjoint=True):
    from qclhip import scba

    p, s = scba_problem()
    g = p.grid
    ws = scba.ScbaWorkspace(p)
    for _ in range(warmup):
        scba.scba_iteration(p, s, ws)
    # Per-iteration times; the median iteration is reported (an occasional
    # slow iteration of this box is visible in the mean and the list).
    per_iter, per_phase = [], []
    for _ in range(iters):
        timings = {}
        t0 = time.perf_counter()
        row = scba.scba_iteration(p, s, ws, timings)
        per_iter.append(time.perf_counter() - t0)
        per_phase.append(timings)
    med = sorted(range(iters), key=lambda i: per_iter[i])[(iters - 1) // 2]
    dt = per_iter[med]
    out = dict(iter_s=dt, phases_ms={k: 1e3 * v for k, v in per_phase[med].items()}, last=row,
               iter_ms_each=[1e3 * x for x in per_iter], iter_ms_mean=1e3 * sum(per_iter) / iters)
    if adjoint:
        # Fixed-point adjoint iterations at the same size (the gradient dJ/dh):
        # the cost of one reverse sweep, from the difference of 1 and 4 steps.
        from qclhip import adjoint as adj
        import torch
        A = adj.ScbaAdjoint(p)
        A.gradient(s, tol=0.0, max_iterations=1)
        times = []
        for k in (1, 4):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            A.gradient(s, tol=0.0, max_iterations=k)
            g.ctx.synchronize()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        out["adjoint_iter_ms"] = (times[1] - times[0]) / 3 * 1e3
        out["adjoint_gradient_1iter_ms"] = times[0] * 1e3
    out["outer"] = outer_bench(p, s)
    return out


# Algorithmic HBM traffic of each phase of the lean iteration, in fields
# (one field = 16 N^2 NE Nk bytes): GF solve reads the three totals and writes
# G^R, G<, G>; the state residual pass reads G^R, S^R, S<, S> (Gamma), G<,
# G>, S<, S> (raw products) and the observables read S<_+, S>_+, G<, G>;
# each embedding component reads its field twice (plus, minus) and writes
# two candidates, then mixing reads 2 old + 2 candidates and writes 2 old
# and the total (and the candidate S< sum for S<); each mechanism reads G<
# and G> and writes S<, S> (contraction), reads S<, S> and writes S^R
# (Hilbert), then mixing reads 3 old + 3 candidates, writes 3 old, reads and
# writes 3 totals and the candidate sum; r_K reads G^R, the sum and G<.
def lean_phase_fields(M):
    return {"dyson": 6, "residuals_state_observables": 12, "candidate_embedding_mix": 34,
            "candidate_scattering_mix": 24 * M, "residual_keldysh": 3}


def scba_full_bench(iters=2, warmup=1):
    """The full SCBA iteration of a six-layer GaAs/Al0.15Ga0.85As period at
    N=32, NE=4096, Nk=128 (524,288 blocks; one field 8.59 GB): five
    mechanisms (LO thermal low rank, acoustic and alloy momentum independent,
    ionized impurity low rank, IFR separable), the embedding, FFT Hilbert at
    L=8192 with the 2L witness, the state residuals, r_K and r_Sigma every
    iteration, the memory-lean order (31 fields resident); the GF solve in
    FP32 with FP64 refinement, mixing, normalization, observables and
    Poisson in FP64. Reports iter/s, per-phase ms and HBM rates, the
    full-size invariants and the FP32-vs-FP64 error of every observable."""
    import numpy as np
    import torch

    from qclhip import api, outer, scba, synthetic as syn

    torch.cuda.empty_cache()
    free0, total = torch.cuda.mem_get_info()
    t0 = time.perf_counter()
    w = syn.two_well_period()
    p, s, qK, mu = syn.period_scba_problem(w, precision="fp32_refine")
    ws = scba.LeanWorkspace(p)
    g = p.grid
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    resident = (total - torch.cuda.mem_get_info()[0]) / 2**30
    log(f"  period problem ready in {setup_s:.1f} s, mu {mu:.6f}, qK {qK}, "
        f"device memory in use {resident:.1f} GiB of {total / 2**30:.1f}")
    for _ in range(warmup):
        row = scba.scba_iteration_lean(p, s, ws)
        log(f"  warm-up iteration: lam {row['lam']:.9f} J {row['Jbar']:.9e} r_Sigma {row['r_Sigma']:.3e}")
    rows, per_iter, per_phase = [], [], []
    for _ in range(iters):
        timings = {}
        t0 = time.perf_counter()
        rows.append(scba.scba_iteration_lean(p, s, ws, timings))
        per_iter.append(time.perf_counter() - t0)
        per_phase.append(timings)
        log(f"  iteration: {per_iter[-1] * 1e3:.1f} ms, lam {rows[-1]['lam']:.9f} J {rows[-1]['Jbar']:.9e} "
            f"r_Sigma {rows[-1]['r_Sigma']:.3e}")
    med = sorted(range(iters), key=lambda i: per_iter[i])[(iters - 1) // 2]
    dt = per_iter[med]
    timings = per_phase[med]
    # One more iteration with the GF solve in FP64 (the other phases are FP64 already).
    g.set_precision("fp64")
    t0 = time.perf_counter()
    scba.scba_iteration_lean(p, s, ws)
    dt64 = time.perf_counter() - t0
    g.set_precision("fp32_refine")
    log(f"  iteration with the FP64 GF solve: {dt64 * 1e3:.1f} ms")
    field = 16.0 * w["N"] ** 2 * w["NE"] * w["Nk"]
    phases = {}
    for k, nf in lean_phase_fields(len(p.mechanisms)).items():
        ms = 1e3 * timings[k]
        phases[k] = dict(ms=ms, fields=nf, GBs=nf * field / (ms * 1e-3) / 1e9,
                         hbm_frac=nf * field / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS)
    nblk = w["NE"] * w["Nk"]
    phases["dyson"]["tflops"] = nblk * 40.0 * w["N"] ** 3 / (phases["dyson"]["ms"] * 1e-3) / 1e12
    last = rows[-1]
    number = api.number_functional(g, s.green.GL)
    # FP32+refine against FP64 on the same state (the totals of the last mix)
    chi = api.basis_to_device(w["phi"])
    obs = {}
    for prec in ("fp64", "fp32_refine"):
        g.set_precision(prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        green, info = api.current_green_production(g, p.h, [ws.totals.parts()], p.target, state=s.green,
                                                   spectral=False)
        torch.cuda.synchronize()
        gf_ms = (time.perf_counter() - t1) * 1e3
        J = api.boundary_flux_bar(g, s.plus.L, s.plus.G, green.GL, green.GG)
        pops = api.state_populations(g, green.GL, p.target)
        dens = api.electron_density(g, green.GL, chi, 2.0).cpu().numpy()
        obs[prec] = dict(J=J, lam=info["lam"], pops=pops, dens=dens, gf_ms=gf_ms,
                         fallbacks=g.refine_fallbacks() if prec != "fp64" else 0)
    a, b = obs["fp64"], obs["fp32_refine"]
    errors = dict(J=abs(b["J"] - a["J"]) / abs(a["J"]), lam=abs(b["lam"] - a["lam"]) / abs(a["lam"]),
                  populations=float(np.linalg.norm(b["pops"] - a["pops"]) / np.linalg.norm(a["pops"])),
                  density=float(np.linalg.norm(b["dens"] - a["dens"]) / np.linalg.norm(a["dens"])))
    # Poisson step at N_z = 512 from the FP32+refine G<
    dev = lambda x: torch.as_tensor(x, device="cuda", dtype=torch.float64)
    op = outer.OuterProblem(dev(w["eps_r"]), dev(w["wx"]), dev(w["ND"]), w["lambda_P"], chi,
                            api.basis_to_device(w["phi"] * math.sqrt(w["dz"])), api.matrix_to_device(w["S"]),
                            dev(w["kin"]), api.matrix_to_device(w["Minv"]), ND2D=w["ND2D"], Lp=w["Lp"])
    U = torch.zeros(w["Nz"], dtype=torch.float64, device="cuda")
    outer.poisson_step(op, g, s.green.GL, U)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prow = outer.poisson_step(op, g, s.green.GL, U)
    torch.cuda.synchronize()
    poisson_ms = (time.perf_counter() - t1) * 1e3
    out = dict(iter_s=dt, iter_per_s=1.0 / dt, phases=phases, setup_s=setup_s,
               iter_ms_each=[1e3 * x for x in per_iter], iter_ms_fp64_gf=1e3 * dt64,
               memory_GiB=dict(before=(total - free0) / 2**30, resident=resident, device=total / 2**30,
                               field=field / 2**30), fields_resident=31,
               invariants=dict(number_rel=abs(number - p.target) / p.target, roundoff_witness=last["roundoff"],
                               r_D=last["r_D"], r_caus=last["r_caus"], r_PSD=last["r_PSD"]),
               fp32_vs_fp64=errors, gf_ms=dict(fp64=a["gf_ms"], fp32_refine=b["gf_ms"]),
               fp64_fallbacks=int(b["fallbacks"]),
               last={k: v for k, v in last.items() if k != "populations"},
               poisson=dict(ms=poisson_ms, r_P=prow.get("r_P"), neutrality=prow.get("neutrality")),
               qK=qK, mu=mu, thicknesses_nm=[float(x) for x in w["thick_nm"]])
    del p, s, ws, op, chi
    torch.cuda.empty_cache()
    return out


def _clone_state(s):
    """A device copy of an SCBA state (the Green functions and every family)."""
    from qclhip import api, scba

    cl = lambda f: scba.Family(f.R.clone(), f.L.clone(), f.G.clone())
    g = s.green
    green = api.GreenState(g.GR.clone(), g.GL.clone(), g.GG.clone(), None if g.A is None else g.A.clone(),
                           g.scale.clone())
    return scba.ScbaState(green, {k: cl(v) for k, v in s.scattering.items()}, cl(s.plus), cl(s.minus))


def _converge_logged(p, s, tol, max_iterations, label, chunk=100):
    """adjoint.converge in chunks of `chunk` iterations with a progress line
    per chunk; returns the iterations done."""
    from qclhip import adjoint as adj

    done = 0
    while done < max_iterations:
        n = min(chunk, max_iterations - done)
        it = adj.converge(p, s, tol=tol, max_iterations=n)
        done += it
        J = s.history[-1]["Jbar"]
        log(f"    {label}: {done} iterations, J {J:.15e}, last relative change "
            f"{abs(J - s.history[-2]['Jbar']) / abs(J):.2e}")
        if it < n:
            break
    return done


def gradient_bench(forward_tol=1e-12, adjoint_tol=1e-9, max_forward=1500, max_adjoint=600, Nz=512,
                   checkpoints=(100, 200, 400), fd_check=False, lo_strength=32.0):
    """The differentiation workload at N=24, NE=2048, Nk=64 (the SCBA
    benchmark problem): the forward loop converged (relative change of J
    below forward_tol, at most max_forward iterations), then the fixed-point
    adjoint dJ/dh (until the relative change of the adjoint state is below
    adjoint_tol or max_adjoint iterations; the implicit adjoint needs only
    the converged state, so no iterate is checkpointed), carried to dJ/dU on
    an Nz-point potential grid (project_hamiltonians' chain rule with a
    synthetic orthonormal basis) and to six layer thicknesses shaping that
    potential (central differences of the host assembly). Both fixed points
    iterate undamped (alpha = 1) and without the per-iteration measurement
    (neither moves the fixed point); dJ/dU is also formed at the adjoint's
    checkpoints and its relative change between them is reported. With
    fd_check, dJ/dU along a random potential direction is compared with the
    central difference of the converged current (two more forward solves)."""
    import dataclasses

    import numpy as np
    import torch

    from qclhip import adjoint as adj
    from qclhip import api

    # The LO mechanism 32 times stronger than in the SCBA benchmark: with the
    # benchmark's weak inelastic scattering the energy distribution relaxes
    # through a slowly oscillating mode (the current still moves by 1e-8 per
    # iteration after 1200 undamped iterations), with this one the fixed
    # point is reached to 1e-14 within about 800.
    p, s = scba_problem(lo_strength=lo_strength)
    p.alpha_mix = 1.0  # the fixed point does not depend on the damping
    p.measure = False  # neither the fixed-point measurement nor the witness moves the fixed point
    p.verify_roundoff = False
    g = p.grid
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = _converge_logged(p, s, forward_tol, max_forward, "forward", chunk=200)
    g.ctx.synchronize()
    t_fwd = time.perf_counter() - t0
    J = s.history[-1]["Jbar"]
    dJ_last = abs(J - s.history[-2]["Jbar"]) / abs(J)
    log(f"  forward: {fwd} iterations in {t_fwd:.1f} s, J {J:.12e}, last relative change {dJ_last:.2e}")
    rng = np.random.default_rng(91)
    Q, _ = np.linalg.qr(rng.normal(size=(Nz, g.Nb)) + 1j * rng.normal(size=(Nz, g.Nb)))
    h_host = api.h_to_host(p.h)
    fd = None
    if fd_check:
        # central differences of the converged current along a random potential
        # direction dU at two step sizes (max |dU| 0.2 and 0.1 meV) and their
        # Richardson extrapolation; each side restarts from the converged state
        # and runs fd_iterations undamped iterations (the per-iteration change
        # of J alone can stop early on the problem's slow mode)
        fd_iterations = 600
        dU1 = rng.normal(size=Nz)
        dU1 /= np.max(np.abs(dU1))
        t0 = time.perf_counter()
        cds = {}
        for step in (2e-4, 1e-4):
            dV = (Q.conj().T * (step * dU1)[None, :]) @ Q
            Js = []
            for sgn in (1.0, -1.0):
                p2 = dataclasses.replace(p, h=api.h_to_device(h_host + sgn * dV[None, :, :]))
                s2 = _clone_state(s)
                _converge_logged(p2, s2, 0.0, fd_iterations, f"forward at U {'+' if sgn > 0 else '-'} {step:g} dU",
                                 chunk=200)
                Js.append(s2.history[-1]["Jbar"])
                del p2, s2
                torch.cuda.empty_cache()
            cds[step] = (Js[0] - Js[1]) / (2.0 * step)
        rich = (4.0 * cds[1e-4] - cds[2e-4]) / 3.0
        fd = dict(steps_eV=[2e-4, 1e-4], central_differences=[cds[2e-4], cds[1e-4]], richardson=rich,
                  forward_iterations_each=fd_iterations, seconds=time.perf_counter() - t0)
        log(f"  finite differences (per unit step): {cds[2e-4]:.12e}, {cds[1e-4]:.12e}; Richardson {rich:.12e}")
    A = adj.ScbaAdjoint(p)
    t0 = time.perf_counter()
    hbar, meta = A.gradient(s, tol=adjoint_tol, max_iterations=max_adjoint, acceleration="linear", alpha=1.0,
                            checkpoints=checkpoints)
    g.ctx.synchronize()
    t_adj = time.perf_counter() - t0
    log(f"  adjoint: {meta['iterations']} iterations in {t_adj:.1f} s, change {meta['change']:.2e}")
    t0 = time.perf_counter()
    gU = api.potential_gradient(hbar, Q)
    t_U = time.perf_counter() - t0
    steps = []
    prev = None
    for it_c, hb in list(meta["checkpoints"]) + [(meta["iterations"], hbar)]:
        gc = api.potential_gradient(hb, Q)
        if prev is not None:
            steps.append(dict(from_iteration=prev[0], to_iteration=it_c,
                              dJdU_rel_change=float(np.linalg.norm(gc - prev[1]) / np.linalg.norm(gc))))
        prev = (it_c, gc)
    if fd is not None:
        fd["adjoint"] = float(np.dot(gU, dU1))
        fd["rel_diff"] = [abs(fd["adjoint"] - c) / abs(c) for c in fd["central_differences"]]
        fd["rel_diff_richardson"] = abs(fd["adjoint"] - fd["richardson"]) / abs(fd["richardson"])
        log(f"  dJ/dU . dU {fd['adjoint']:.12e}: rel. diff {fd['rel_diff']} to the central differences, "
            f"{fd['rel_diff_richardson']:.2e} to their Richardson extrapolation")
    z = np.linspace(0.0, 6.0, Nz)
    L0 = np.array([0.9, 1.3, 0.8, 1.1, 0.7, 1.2])
    edges = lambda L: np.concatenate([[0.0], np.cumsum(L)])
    sig = lambda x: 1.0 / (1.0 + np.exp(-x / 0.05))

    def assemble(L):
        e = edges(L)
        U = sum(0.08 * (sig(z - e[j]) - sig(z - e[j + 1])) for j in (0, 2, 4))
        V = (Q.conj().T * U[None, :]) @ Q
        return h_host + V[None, :, :]

    t0 = time.perf_counter()
    gL = adj.thickness_gradient(hbar, assemble, L0, step=1e-4)
    t_L = time.perf_counter() - t0
    hb = hbar.cpu().numpy()
    out = dict(forward_iterations=fwd, forward_s=t_fwd, forward_last_rel_change=dJ_last, J=J,
               adjoint_iterations=meta["iterations"], adjoint_s=t_adj, adjoint_change=meta["change"],
               adjoint_ms_per_iter=1e3 * t_adj / max(meta["iterations"], 1), dJdU_convergence=steps,
               hbar_norm=float(np.linalg.norm(hb)), dJdU_norm=float(np.linalg.norm(gU)), dJdU_s=t_U,
               dJdU_sample=[float(x) for x in gU[::64]], dJdL=[float(x) for x in gL], dJdL_s=t_L,
               config=f"N=24 NE=2048 Nk=64 SCBA benchmark problem (3 mechanisms + embedding) with the LO "
                      f"mechanism {lo_strength:g} times stronger, forward and adjoint fixed points iterated undamped; "
                      f"potential on Nz={Nz} points (synthetic orthonormal basis), six layer thicknesses "
                      "shaping a barrier profile on that grid; fixed-point adjoint (no checkpoints needed)")
    if fd is not None:
        out["fd_check"] = fd
    return out


def dense_contraction_bench(N=24, NE=2048, Nk=64, reps=2):
    """The dense kernel representation (Kflat, Nc x Nc with Nc = Nk N^2 =
    36,864: 21.7 GB at N=24, NE=2048, Nk=64) on the hand-written FP64
    matrix-core GEMM: S = qK K[w G] over all energies, 8 NE Nc^2 flops per
    contraction, HIP events on the context's stream. Kflat is random
    (built on the device); four columns of S are checked against an
    independent product."""
    import numpy as np
    import torch

    from qclhip import api, synthetic as syn

    torch.cuda.empty_cache()
    w = syn.gf_workload(N=N, NE=NE, Nk=Nk)
    g = api.Grid(w["eps"], w["wE"], w["wk"], N)
    Nc = Nk * N * N
    gen = torch.Generator(device="cuda").manual_seed(11)
    Kt = torch.randn((Nc, Nc), dtype=torch.complex128, device="cuda", generator=gen) * (1.0 / Nc)
    K = api.ProductionKernel.dense_from_device(g, Kt)
    G = torch.randn(api.field_shape(NE, Nk, N), dtype=torch.complex128, device="cuda", generator=gen)
    S = api.production_static_contraction(K, G, 1.0)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        api.production_static_contraction(K, G, 1.0, S=S)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = 8.0 * NE * Nc * Nc
    # check: S2[:, q] = sum_q' G2[:, q'] Kflat[q, q'] w(q'), G2 = G as NE x Nc
    G2 = G.reshape(Nc, NE)
    wq = torch.as_tensor(np.tile(w["wk"], N * N), device="cuda", dtype=torch.float64)
    rng = np.random.default_rng(5)
    err = 0.0
    S2 = S.reshape(Nc, NE)
    for q in rng.integers(0, Nc, 4):
        ref = (Kt[:, q] * wq) @ G2  # sum over q' of Kflat[q, q'] w(q') G2[:, q']
        err = max(err, float(torch.linalg.norm(S2[q] - ref) / torch.linalg.norm(ref)))
    del Kt, K, G, S, S2, G2
    torch.cuda.empty_cache()
    return dict(ms=ms, tflops=flops / (ms * 1e-3) / 1e12, frac_fp64=flops / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                kflat_GB=16.0 * Nc * Nc / 1e9, sampled_rel_err=err,
                config=f"N={N} NE={NE} Nk={Nk} (Nc={Nc}), k_zgemm_nt: 64x64 tiles, 16-deep stages, four FP64 "
                       "16x16x4 MFMAs per complex step")


def long_axis_hilbert_bench(NE=30721, Nk=32, N=10, reps=3):
    """S^R = H[Gamma] - i Gamma/2 with the 2L round-off witness on an energy
    axis of NE = 30721 nodes (L = 65536, the four-step transform through
    HBM) over Nk N^2 columns; algorithmic bytes: read S< and S>, write S^R.
    Beside it, both transform paths at L = 16384 (NE = 8192), the longest
    length the one-workgroup path takes."""
    import torch

    from qclhip import api

    def run(NE_, path):
        torch.cuda.empty_cache()
        eps = torch.linspace(-0.3, 0.6, NE_, dtype=torch.float64).numpy()
        wE = eps.copy()
        wE[:] = eps[1] - eps[0]
        wE[[0, -1]] /= 2
        g = api.Grid(eps, wE, torch.ones(Nk, dtype=torch.float64).numpy(), N)
        plan = api.FFTHilbertPlan(g, verify_roundoff=True)
        plan.set_path(path)
        gen = torch.Generator(device="cuda").manual_seed(NE_)
        SL = torch.randn(api.field_shape(NE_, Nk, N), dtype=torch.complex128, device="cuda", generator=gen)
        SG = torch.randn(api.field_shape(NE_, Nk, N), dtype=torch.complex128, device="cuda", generator=gen)
        SR = torch.empty_like(SL)
        api.retarded_self_energy_fft(plan, SL, SG, return_roundoff=True, SR=SR)
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            _, _, r = api.retarded_self_energy_fft(plan, SL, SG, return_roundoff=True, SR=SR)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 3.0 * SL.numel() * 16
        del SL, SG, SR, plan
        torch.cuda.empty_cache()
        return dict(ms=ms, GBs=nbytes / (ms * 1e-3) / 1e9, frac_hbm=nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                    witness=r, field_GB=nbytes / 3 / 1e9)

    out = dict(config=f"NE={NE} (L=65536) Nk={Nk} N={N}: {Nk * N * N} columns, S^R with the 2L witness; "
                      f"k_fs_columns<16> + k_fs_rows<4096> + k_fs_output<16>",
               long_axis=run(NE, "hbm"))
    out["L16384"] = {path: run(8192, path) for path in ("lds", "hbm")}
    return out


def outer_bench(p, s, Nz=512, reps=5):
    """The Poisson step after an SCBA run at the same size: density of G<
    (N_E N_k N_b^2 values), the bordered Poisson LU solve (N_z + 1 = 513),
    potential mixing and the projection of N_k Hamiltonians; synthetic
    orthonormal basis on N_z points."""
    import numpy as np
    import torch
    from qclhip import api, outer
    g = p.grid
    rng = np.random.default_rng(5)
    Nb, dx = g.Nb, 0.1
    Q, _ = np.linalg.qr(rng.normal(size=(Nz, Nb)) + 1j * rng.normal(size=(Nz, Nb)))
    ND = np.where(np.arange(Nz) % 64 < 8, 2.0, 0.0)
    dev = lambda x: torch.as_tensor(x, device="cuda", dtype=torch.float64)
    op = outer.OuterProblem(dev(rng.uniform(10, 13, Nz)), dev(np.full(Nz, dx)), dev(ND), 1.0,
                            api.basis_to_device(Q / np.sqrt(dx)), api.basis_to_device(Q),
                            api.matrix_to_device(np.diag(np.linspace(0, 0.4, Nb)) + 0j),
                            dev(np.linspace(0, 1, g.Nk)), api.matrix_to_device(np.eye(Nb) + 0j),
                            ND2D=float(dx * ND.sum()), Lp=Nz * dx)
    U = torch.zeros(Nz, dtype=torch.float64, device="cuda")
    outer.poisson_step(op, g, s.green.GL, U)
    parts = {"density": 0.0, "poisson": 0.0, "mix_project": 0.0}
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nbar = api.electron_density(g, s.green.GL, op.chi, op.gs)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cand, zeta, rP, neut = api.periodic_poisson(op.eps_r, op.wx, op.ND, op.lambda_P, nbar)
        t2 = time.perf_counter()
        api.mix_potential(U, cand, op.alpha_P)
        api.project_hamiltonians(op.Phi, U, op.S, op.kin, op.Minv)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        parts["density"] += (t1 - t0) / reps * 1e3
        parts["poisson"] += (t2 - t1) / reps * 1e3
        parts["mix_project"] += (t3 - t2) / reps * 1e3
    parts["step_ms"] = sum(parts.values())
    parts["config"] = f"Nz={Nz} (bordered system {Nz + 1}), Nb={Nb}, Nk={g.Nk}, G< of NE={g.NE}"
    return parts


def graph_bench(iters=200):
    """Small SCBA problem (the tutorial's size: N=3, NE=49, Nk=5), launch-bound:
    eager iterations against replays of one captured iteration."""
    import torch

    from qclhip import scba

    out = {}
    for name, (N, NE, Nk) in (("tutorial_N3_NE49_Nk5", (3, 49, 5)), ("small_N8_NE256_Nk8", (8, 256, 8))):
        p, s = scba_problem(N, NE, Nk)
        ws = scba.ScbaWorkspace(p)
        for _ in range(3):
            scba.scba_iteration(p, s, ws)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            scba.scba_iteration(p, s, ws)
        te = (time.perf_counter() - t0) / iters
        graph = scba.IterationGraph(p, s, ws)
        graph.replay()
        t0 = time.perf_counter()
        for _ in range(iters):
            graph.replay()
        tg = (time.perf_counter() - t0) / iters
        out[name] = dict(eager_ms_per_iter=te * 1e3, graph_ms_per_iter=tg * 1e3, speedup=te / tg)
    return out


def _instance_problem(N, NE, Nk, seed):
    """One instance of the uneven sweep: the synthetic banded Hamiltonian of
    the GF workload, two mechanisms (momentum independent; low-rank thermal
    LO), the bulk embedding, seed at the workload's mu, lean workspace."""
    import numpy as np

    from qclhip import api, scba, synthetic as syn

    w = syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=seed)
    rng = np.random.default_rng(seed + 7)
    g = api.Grid(w["eps"], w["wE"], w["wk"], N)
    mechs = [scba.Mechanism("acoustic", api.ProductionKernel(g, "momentum_independent",
                                                             syn.covariance_block(rng, N, 2, 0.04)), 1.0)]
    lo_l, lo_r = syn.low_rank_factors(rng, N, Nk, r=4, scale=0.02)
    mechs.append(scba.Mechanism("LO", api.ProductionKernel(g, "low_rank", lo_l, lo_r, rank=4), 1.0, True))
    T = 0.05 * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    h = api.h_to_device(w["h"])
    kbt, hw = w["kbt"], 0.36
    GR0, _ = api.retarded_green_production(g, h, None, 0.05)
    q = api.seed_number_weights(g, GR0)
    del GR0
    target = float(np.sum(q * api.fermi(w["eps"], w["mu"], kbt)))
    seed_green, _ = api.seed_green_production(g, h, 0.05, kbt, target, spectral=False)
    p = scba.ScbaProblem(g, h, api.matrix_to_device(T), api.matrix_to_device(T.conj().T), mechs, target, Ep=0.5,
                         hwLO=hw, NLO=1.0 / math.expm1(hw / kbt), alpha_mix=0.3)
    s = scba.initial_state(p, seed_green)
    return p, s, scba.LeanWorkspace(p)


def batch_scba_bench(ninst=64, Nk=64, N_range=(12, 48), NE_range=(1024, 4096), replays=2, wave_GiB=200.0):
    """The uneven sweep as batched SCBA iterations: ninst instances (N, NE
    uniform in the ranges, Nk momenta, seeded), streamed in waves whose lean
    state fits wave_GiB of HBM (22 fields per instance: G^R, G<, G>, four
    families, totals, three candidate fields and the candidate S< sum). Per
    wave: every instance's first lean iteration, the batch autotuner (launch
    mode of the batched GF solve), one wave-iteration captured as a HIP
    graph, `replays` timed replays, and the same instances iterated one at a
    time (eager lean iterations) for comparison. Batched results equal the
    one-at-a-time ones bit for bit (tests/test_gpu_batch_scba.py)."""
    import gc

    import numpy as np
    import torch

    from qclhip import api, scba, synthetic as syn

    rng = np.random.default_rng(syn.SEED + 64)
    shapes = [(int(rng.integers(N_range[0], N_range[1] + 1)), int(rng.integers(NE_range[0], NE_range[1] + 1)))
              for _ in range(ninst)]
    fields = 22
    waves, cur, cur_b = [], [], 0.0
    for i, (N, NE) in enumerate(shapes):
        b = fields * 16.0 * N * N * NE * Nk / 2**30
        if cur and cur_b + b > wave_GiB:
            waves.append(cur)
            cur, cur_b = [], 0.0
        cur.append(i)
        cur_b += b
    waves.append(cur)
    def run_wave(wv, members):
        insts = [_instance_problem(shapes[i][0], shapes[i][1], Nk, syn.SEED + 100 + i) for i in members]
        ps, ss, ws = [x[0] for x in insts], [x[1] for x in insts], [x[2] for x in insts]
        for p, s, w in zip(ps, ss, ws):
            scba.scba_iteration_lean(p, s, w)
        lb = scba.LeanBatch(ps, ss, ws)
        if len(members) > 1:
            tune = lb.autotune()
        else:  # one instance: its own launches (register prefetch), nothing to share
            api.batch_set_mode(lb.batch, "per_instance")
            tune = dict(mode="per_instance")
        lb.capture()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(replays):
            lb.replay()
        tb = (time.perf_counter() - t0) / replays
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for p, s, w in zip(ps, ss, ws):
            scba.scba_iteration_lean(p, s, w)
        ts = time.perf_counter() - t0
        return tb, ts, dict(wave=wv, instances=len(members), **tune)

    t_batch = t_single = 0.0
    solves = 0
    tunes = []
    t_all = time.perf_counter()
    for wv, members in enumerate(waves):
        tb, ts, tune = run_wave(wv, members)
        gc.collect()
        torch.cuda.empty_cache()
        t_batch += tb
        t_single += ts
        tunes.append(tune)
        solves += sum(shapes[i][1] * Nk for i in members)
        log(f"  wave {wv}: {len(members)} instances, batched {tb * 1e3:.1f} ms, one at a time {ts * 1e3:.1f} ms, "
            f"GF {tune}")
    return dict(instances=ninst, waves=len(waves), Nk=Nk,
                shapes=f"N in U{{{N_range[0]}..{N_range[1]}}}, NE in U{{{NE_range[0]}..{NE_range[1]}}} (seeded)",
                batched_ms_per_iteration=t_batch * 1e3, one_by_one_ms_per_iteration=t_single * 1e3,
                speedup=t_single / t_batch, blocks_per_iteration=solves,
                batched_solves_per_s=solves / t_batch, autotune=tunes, total_s=time.perf_counter() - t_all,
                per_instance_error_vs_fp64="0 (batched = one-at-a-time FP64 iteration bit for bit)")


def batch_bench(reps=3, ninst=16, Nk=16, N_range=(12, 48), NE_range=(1024, 4096)):
    """Uneven sweep: ninst instances with N and NE uniform in the given ranges,
    Nk momenta, two families each; one batched current_green call against
    the same instances solved one call at a time."""
    import numpy as np
    import torch

    from qclhip import api, synthetic as syn

    rng = np.random.default_rng(syn.SEED + 4)
    grids, hs, fams, shapes = [], [], [], []
    for i in range(ninst):
        N, NE = int(rng.integers(N_range[0], N_range[1] + 1)), int(rng.integers(NE_range[0], NE_range[1] + 1))
        w = syn.gf_workload(N=N, NE=NE, Nk=Nk, seed=syn.SEED + 10 + i)
        grids.append(api.Grid(w["eps"], w["wE"], w["wk"], N))
        hs.append(api.h_to_device(w["h"]))
        fams.append([syn.diag_fields_device(syn.injection_family(w)), syn.diag_fields_device(w)])
        shapes.append((N, NE))
    batch = api.Batch(grids)
    targets = [1.0] * ninst
    states, _ = api.batch_current_green(batch, hs, fams, targets)
    tune = api.batch_autotune(batch, hs, fams, targets, states)  # persistent vs per-instance launches
    nblk = sum(NE * Nk for _, NE in shapes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        api.batch_current_green(batch, hs, fams, targets, states=states)
    grids[0].ctx.synchronize()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        for i in range(ninst):
            api.current_green_production(grids[i], hs[i], fams[i], targets[i], state=states[i], spectral=False,
                                         return_scalars=False)
    grids[0].ctx.synchronize()
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t0) / reps
    return dict(instances=ninst, Nk=Nk, solves=nblk, batch_ms=tb * 1e3, batch_solves_per_s=nblk / tb, launch=tune,
                one_by_one_ms=ts * 1e3, one_by_one_solves_per_s=nblk / ts,
                shapes=f"N in U{{{N_range[0]}..{N_range[1]}}}, NE in U{{{NE_range[0]}..{NE_range[1]}}}")


def cpu_baseline(seconds=15.0, repeats=5):
    """Oracle GF solves on a bounded sample (all NE energies, the first k
    momenta): one warm-up run, then the median of `repeats` timed runs."""
    import numpy as np

    from oracle import oracle as orc
    from qclhip import synthetic as syn

    w = syn.gf_workload()
    threads = orc.num_threads()

    def run(k):
        sub = dict(w)
        sub["Nk"] = k
        sub["gamma"] = w["gamma"][:, :k]
        SR, SL, SG = syn.diag_fields_host(sub)
        inj = syn.injection_family(w)
        inj["Nk"], inj["gamma"] = k, inj["gamma"][:, :k]
        PR, PL, PG = syn.diag_fields_host(inj)
        h = np.asfortranarray(w["h"][:k])
        t0 = time.perf_counter()
        orc.current_green(w["eps"], h, [PR, SR], [PL, SL], [PG, SG], w["wE"], w["wk"][:k], 2, 1e-3)
        return time.perf_counter() - t0

    t1 = run(1)  # warm-up (and the per-momentum cost)
    k = int(max(1, min(w["Nk"], round(seconds / repeats / max(t1, 1e-3)))))
    times = sorted(run(k) for _ in range(repeats))
    tk = times[len(times) // 2]
    blocks = w["NE"] * k
    return dict(value=blocks / tk, unit="GF solves/s", cores=threads, kind="port",
                sample=f"N=24 NE=2048, first {k} of 64 momenta ({blocks} blocks); median of {repeats} runs after "
                       f"a warm-up ({tk:.2f} s; runs {', '.join(f'{t:.2f}' for t in times)} s); C++ oracle of the "
                       f"production Dyson/Keldysh/normalization with A, OpenMP {threads} threads; the reference's "
                       f"per-block cond(D) SVD (production_operators.jl:1380) is not computed, so the reference "
                       f"CPU path would be slower")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scba-iters", type=int, default=5)
    ap.add_argument("--no-scba", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the full-iteration period workload (N=32)")
    ap.add_argument("--full-iters", type=int, default=3)
    ap.add_argument("--only-full", action="store_true", help="run only the full-iteration workload (profiling)")
    ap.add_argument("--no-gradient", action="store_true", help="skip the gradient leg")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense-kernel contraction leg")
    ap.add_argument("--only-dense", action="store_true")
    ap.add_argument("--no-long-axis", action="store_true", help="skip the NE = 30721 Hilbert leg")
    ap.add_argument("--only-long-axis", action="store_true")
    ap.add_argument("--only-gradient", action="store_true")
    ap.add_argument("--no-batch-scba", action="store_true", help="skip the batched SCBA sweep")
    ap.add_argument("--batch-instances", type=int, default=24)
    ap.add_argument("--only-batch-scba", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--variants", default=",".join(f"{p}_{e}" if p != "fp32_products" else p for p, e in VARIANTS),
                    help="comma-separated GF-solve variants to run (profiling runs select one)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--only-pair", type=int, default=0, metavar="REPS",
                    help="run only the FP64 vector Dyson + Keldysh launch pair REPS times (kernel counters)")
    ap.add_argument("--only-residuals", type=int, default=0, metavar="REPS",
                    help="run only the residual suite and its state pass REPS times (kernel counters)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (the reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the multi-GPU path is not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs the MI355X", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device-to-device copies")
    if args.only_pair:
        print(json.dumps({"pair": pair_only(args.only_pair)}), flush=True)
        return
    if args.only_residuals:
        print(json.dumps({"residuals": residuals_only(args.only_residuals)}, default=float), flush=True)
        return
    if args.only_dense:
        print(json.dumps({"dense_contraction": dense_contraction_bench()}, default=float), flush=True)
        return
    if args.only_long_axis:
        print(json.dumps({"long_axis_hilbert": long_axis_hilbert_bench()}, default=float), flush=True)
        return
    if args.only_batch_scba:
        log(f"batched SCBA sweep: {args.batch_instances} uneven instances, Nk=64")
        print(json.dumps({"batch_scba": batch_scba_bench(args.batch_instances)}, default=float), flush=True)
        return
    if args.only_gradient:
        log("gradient (forward converged, adjoint iteration-capped) with a finite-difference check: "
            "N=24 NE=2048 Nk=64")
        print(json.dumps({"gradient": gradient_bench(fd_check=True)}, default=float), flush=True)
        return
    if args.only_full:
        log("full SCBA iteration: GaAs/Al0.15Ga0.85As period, N=32 NE=4096 Nk=128, five mechanisms")
        full = scba_full_bench(args.full_iters, 1)
        print(json.dumps({"scba_full": full}, default=float), flush=True)
        return
    log("gf batch: N=24 NE=2048 Nk=64")
    wanted = set(args.variants.split(","))
    sel = [(p, e) for p, e in VARIANTS if (f"{p}_{e}" if p != "fp32_products" else p) in wanted]
    if not sel or "fp64_vector" not in wanted:
        print("--variants must include fp64_vector (the reference arithmetic the others are checked against)",
              file=sys.stderr)
        sys.exit(2)
    gf = gf_bench(args.steps, args.warmup, sel, autotune=not args.no_autotune)
    for name, v in gf["variants"].items():
        log(f"{name}: step {v['step_s']*1e3:.3f} ms ({v['solves_per_s']/1e6:.2f} M solves/s), Dyson+Keldysh "
            f"{v['kernel_ms']:.3f} ms, J {v['J']:.12g}, rel. current error {v['current_rel_err']}, "
            f"FP64 fallback {v['fp64_fallback_fraction']:.4f}")
    sc = None
    if not args.no_scba:
        log("scba iteration: N=24 NE=2048 Nk=64, 3 mechanisms + embedding")
        sc = scba_bench(args.scba_iters, 1)
        log(f"scba {sc['iter_s']*1e3:.1f} ms/iter {sc['phases_ms']}")
        log(f"poisson step {sc['outer']['step_ms']:.2f} ms {sc['outer']}")
        if "adjoint_iter_ms" in sc:
            log(f"adjoint {sc['adjoint_iter_ms']:.1f} ms/iter")
    full = None
    if not args.no_full:
        log("full SCBA iteration: GaAs/Al0.15Ga0.85As period, N=32 NE=4096 Nk=128, five mechanisms")
        full = scba_full_bench(args.full_iters, 1)
        log(f"full iteration {full['iter_s']*1e3:.1f} ms/iter, phases "
            f"{ {k: round(v['ms'], 2) for k, v in full['phases'].items()} }, fp32 vs fp64 {full['fp32_vs_fp64']}")
    dense = None
    if not args.no_dense:
        log("dense kernel contraction: N=24 NE=2048 Nk=64 (Kflat 21.7 GB)")
        dense = dense_contraction_bench()
        log(f"dense contraction {dense['ms']:.1f} ms, {dense['tflops']:.1f} TFLOP/s")
    longax = None
    if not args.no_long_axis:
        log("long energy axis: S^R with the witness at NE=30721 (L=65536), Nk=32, N=10")
        longax = long_axis_hilbert_bench()
        log(f"long axis {longax['long_axis']['ms']:.2f} ms, {longax['long_axis']['GBs']:.0f} GB/s")
    bsc = None
    if not args.no_batch_scba:
        log(f"batched SCBA sweep: {args.batch_instances} uneven instances, Nk=64")
        bsc = batch_scba_bench(args.batch_instances)
        log(f"batched SCBA: {bsc['batched_ms_per_iteration']:.1f} ms/iteration batched, "
            f"{bsc['one_by_one_ms_per_iteration']:.1f} one at a time")
    grad = None
    if not args.no_gradient:
        log("gradient (forward converged, adjoint iteration-capped): N=24 NE=2048 Nk=64")
        grad = gradient_bench()
    gr = None
    if not args.no_graph:
        log("graph replay of small SCBA iterations")
        gr = graph_bench()
        log(f"graph {gr}")
    bt = None
    if not args.no_batch:
        bt = {}
        for name, kw in (("uneven_large", dict(ninst=16, Nk=16)),
                         ("small_instances", dict(ninst=64, Nk=5, N_range=(3, 8), NE_range=(49, 200), reps=10))):
            log(f"batched sweep {name}: {kw}")
            r = batch_bench(**kw)
            bt[name] = r
            log(f"batch {r['batch_ms']:.2f} ms ({r['batch_solves_per_s']/1e6:.2f} M solves/s), one call per "
                f"instance {r['one_by_one_ms']:.2f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle)")
        cpu = cpu_baseline(args.cpu_seconds)
        log(f"cpu {cpu['value']:.1f} solves/s on {cpu['cores']} threads")
    # Headline: the fastest variant whose current is within 1e-8 of FP64.
    ok = {k: v for k, v in gf["variants"].items()
          if v["current_rel_err"] is not None and v["current_rel_err"] <= CURRENT_TOL}
    best_name = max(ok, key=lambda k: ok[k]["solves_per_s"])
    best = ok[best_name]
    k_s = best["kernel_ms"] * 1e-3
    achieved = gf["flops"] / k_s / 1e12
    mixed = best["precision"] == "fp32_products"
    peak = PEAK_FP32_MATRIX_TFLOPS if mixed else PEAK_FP64_TFLOPS
    kernels = {"fp64_vector": "k_dyson<3,8> + k_keldysh<3,4>",
               "fp64_matrix": "k_dyson<3,8> + k_keldysh_mfma<double,3,4>",
               "fp32_refine_vector": "k_dyson_mixed<3,4> + k_dyson<3,4> (flagged tiles) + k_keldysh<3,4>",
               "fp32_refine_matrix": "k_dyson_mixed<3,4> + k_dyson<3,4> (flagged tiles) + "
                                     "k_keldysh_mfma<double,3,4>",
               "fp32_products": "k_dyson_mixed<3,4> + k_dyson<3,4> (flagged tiles) + k_keldysh_mfma<float,3,4>"}
    line = {
        "metric": METRIC,
        "value": best["solves_per_s"],
        "unit": "GF solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32_refine": "f32 factorization + f64 refinement and products",
                  "fp32_products": "f32 factorization + f64 refinement, f32 products"}[best["precision"]],
        "data": "synthetic",
        "config": {"workload": "GF-solve batch: N=24 basis, NE=2048 energies on [-0.1,0.3] eV, Nk=64 momenta; "
                               "diagonal scattering family (Im Sigma^R -U[0.5,5] meV, Fermi 77 K at mu) plus an "
                               "injecting family (U[0.5,2] meV, Fermi at mu+50 meV); current_green step",
                   "N": gf["N"], "NE": gf["NE"], "Nk": gf["Nk"], "solves_per_step": gf["nblk"],
                   "variant": best_name, "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": read_traffic() if best_name == "fp64_vector" else None,
                     "kernel": kernels[best_name] + " (one GF-solve launch set)",
                     "kernel_ms": best["kernel_ms"],
                     "hbm_achieved_GBs": gf["bytes"] / k_s / 1e9,
                     "hbm_frac": gf["bytes"] / k_s / 1e9 / PEAK_HBM_GBS},
        "number_check": best["number_check"],
        "variants": {k: {kk: vv for kk, vv in v.items() if kk != "precision"} for k, v in gf["variants"].items()},
        "autotune": gf["autotune"],
    }
    if gf.get("padded"):
        pm = gf["padded"]["kernel_ms"]
        line["pair_padded_planes"] = {
            "layout": "planes NE*Nk + 16 elements apart (qclhip_dyson_keldysh_strided); value above: contiguous",
            "kernel": kernels["fp64_vector"], "kernel_ms": pm,
            "achieved_tflops": gf["flops"] / (pm * 1e-3) / 1e12,
            "frac": gf["flops"] / (pm * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
            "speedup_vs_contiguous": gf["variants"]["fp64_vector"]["kernel_ms"] / pm}
    if sc is not None:
        line["scba_iter_per_s"] = 1.0 / sc["iter_s"]
        line["scba_ms_per_iter"] = sc["iter_s"] * 1e3
        line["scba_phases_ms"] = sc["phases_ms"]
        line["scba_config"] = ("N=24 NE=2048 Nk=64; momentum-independent (acoustic-like), separable IFR (J=3), "
                               "low-rank LO (r=4, thermal), bulk embedding, FFT Hilbert with 2L witness, "
                               "fixed-point measurement every iteration (residual suite r_D r_A r_K r_PSD r_caus; "
                               "self-energy residuals fused into the linear mixing); FP64")
        line["scba_last"] = {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in sc["last"].items()
                             if k != "populations"}
        line["poisson_step"] = sc["outer"]
        if "adjoint_iter_ms" in sc:
            line["adjoint"] = {"iter_ms": sc["adjoint_iter_ms"], "gradient_with_1_iter_ms": sc["adjoint_gradient_1iter_ms"],
                               "config": "N=24 NE=2048 Nk=64, same problem as the SCBA iteration: one damped "
                                         "fixed-point step = candidate reverse + current_green reverse + mixing"}
    if grad is not None:
        line["gradient"] = grad
    if bsc is not None:
        line["batch_scba"] = bsc
    if dense is not None:
        line["dense_contraction"] = dense
    if longax is not None:
        line["long_axis_hilbert"] = longax
    if full is not None:
        line["scba_full_iter_per_s"] = full["iter_per_s"]
        line["scba_full"] = dict(full, config="six-layer GaAs/Al0.15Ga0.85As period (thicknesses U[2,20] nm, seeded), "
                                               "N=32 NE=4096 Nk=128, LO (36 meV, thermal, low rank r=6), acoustic "
                                               "and alloy (momentum independent), ionized impurity (3e16 cm^-3, "
                                               "low rank r=4), IFR (6 interfaces, separable, Lambda=8 nm); FFT "
                                               "Hilbert L=8192 with 2L witness; GF solve FP32+2 FP64 refinement "
                                               "steps, FP64 mixing/normalization/Poisson (N_z=512); "
                                               "memory-lean order")
    if gr is not None:
        line["graph"] = gr
    if bt is not None:
        line["batch"] = bt
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
