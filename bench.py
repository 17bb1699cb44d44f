"""SR steps/s of the VMC step (sample + E_loc + O + SR solve + update) on one
MI355X: 1D transverse-field Ising chain, N=100, real RBM alpha=4 (40,500
parameters), 16,384 samples per step.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
    python bench.py --batched [--steps K] [--warmup W]   (32 machines, see batched())

Timed path (value): mixed precision -- sampling and E_loc in float, O kept
factored (spins and tanh(theta)), matrix-free CG whose products run as
split-bf16 MFMA with float accumulation, FP64 iterative refinement to a 1e-8
relative residual, FP64 update.  The inner CG runs as captured HIP-graph
chunks that stop themselves on the device.  The CG starts from zero without a
preconditioner, as the reference's solveCG does (--precond/--warm-start
switch on the Jacobi preconditioner and warm starts).

Every path below runs the same warm-up and timed steps from the same
parameters and chain seeds, so their rates and CG iteration counts cover the
same step window: the plain FP64 path (explicit S = O_c^T O_c / Ns and
Cholesky, the reference's default solver), the FP64 factored CG (FP64
sampling, FP64 MFMA products), and the mixed solver fed by FP64 sampling
(its chains are those of the FP64 paths, so its energies differ from theirs
only through the mixed solve).  Every timed solve must reach the 1e-8 FP64
residual: a step that does not makes the run exit with status 1.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, ALPHA = 100, 4
M = N * ALPHA
DIM = N * M + N + M
CHAINS = 16384           # one sample per chain per step
LAMBDA, LR, TOL = 1e-3, 0.01, 1e-8
SEED = 20241203
NS = CHAINS
KP, NP = 128, 416        # padded (n+1) and (m+1) of the MFMA grids

# Algorithmic HBM bytes per launch: the per-unit figures of the workload's
# measurement table times the units one launch processes.  A product of the
# factored operator (O_c u or O_c^T y) streams the spins as int8 (Ns x n) and
# T = tanh(theta) as float (Ns x m) once, reads its dim- or Ns-long input and
# writes its Ns- or dim-long result; the CG vector update reads and writes 11
# dim-length float vectors and writes the three bf16 layers of the split
# vector.  Chunk partial sums, bf16 spin copies and padding are this
# implementation's own traffic: they appear in the measured `traffic`, not
# here.
FACTORS = NS * N + NS * M * 4
KERNEL_WORK = {
    "fwd_mfma": FACTORS + DIM * 4 + NS * 4,
    "bwd_mfma": FACTORS + NS * 4 + DIM * 4,
    "cg_step": 11 * DIM * 4 + 3 * KP * NP * 2,
    # sampler: spins in/out and the recorded samples (spins, float theta)
    "sample": NS * N * 3 + NS * M * 4,
    # E_loc ratios: theta (float) and spins in, ratios (double) out
    "eloc_ratio": NS * M * 4 + NS * N + NS * N * 8,
}
VALU_KERNELS = ("sample", "eloc_ratio")
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) of this script, each counter scaled by what it reads
# for a copy of known size (bench.py --copy-probe); written with the command
# lines that produced it
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_tfi_n100_a4.json")
# VALU utilisation of the sampler and E_loc kernels from a rocprofv3 SQ
# counter pass (same provenance rules)
VALU_FILE = os.path.join(ROOT, "profiles", "pmc_valu_sample_eloc_tfi_n100_a4.json")


AUTOTUNE = False
GRAPH_CHUNK = 16  # CG iterations per captured graph of the mixed solve


def _ctx(yq, precision):
    ctx = yq.Context(0)
    ctx.autotune(AUTOTUNE)
    ctx.graph_mode(GRAPH_CHUNK)
    ctx.rbm_create(N, M)
    ctx.init_random(SEED, 0.01)
    ctx.ham_set(yq.TFI, N, J=-1.0, h=-1.0)   # TFIsing(N, -1, -h) at h = 1
    ctx.sampler_config(yq.LOCAL, CHAINS, precision=precision, seed=SEED)
    return ctx


def run_path(yq, sample_precision, solver, warmup, steps, solve_precision=None, precond=0,
             warm_start=0, prof_label=None):
    """warmup + steps SR epochs from the same parameters and chain seeds with
    sampling in sample_precision and the solve in solve_precision (default:
    the same); returns the timed seconds, the per-step outputs of the timed
    steps, the energies of all steps and the HIP-event kernel timings."""
    ctx = _ctx(yq, sample_precision)
    kw = dict(n_sweeps=1, solver=solver,
              precision=sample_precision if solve_precision is None else solve_precision,
              lam=LAMBDA, cg_tol=TOL, cg_max_iter=50000, lr=LR, precond=precond,
              warm_start=warm_start, allow_unconverged=True)
    energies, outs = [], []
    for i in range(warmup):
        o = ctx.sr_step(n_discard=20 if i == 0 else 0, **kw)
        energies.append(o["energy"])
    if prof_label:
        ctx._L.yq_prof_enable(ctx._h, prof_label.encode())
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        o = ctx.sr_step(n_discard=0 if (warmup or i) else 20, **kw)
        o["wall_ms"] = 1e3 * (time.perf_counter() - ts)
        outs.append(o)
        energies.append(o["energy"])
    t1 = time.perf_counter()   # every sr_step returns after its stream synchronised
    prof = None
    if prof_label:
        prof = {}
        tot, cnt = C.c_double(), C.c_longlong()
        for lab in prof_label.split(","):
            ctx._L.yq_prof_read(ctx._h, lab.encode(), C.byref(tot), C.byref(cnt))
            prof[lab] = (tot.value, cnt.value)
        ctx._L.yq_prof_enable(ctx._h, None)
    ctx.close()
    return (t1 - t0), outs, energies, prof


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(params, steps=3, threads=16):
    """The oracle's threaded C++ restatement of the reference CG step, timed
    at full size: per step SamplerMT.sample(1, 0) over 16,384 chains, E_loc,
    constructFromSamples with O materialised (16,384 x 40,500 doubles), CG on
    O_c to the same 1e-8 relative residual, SGD update -- RunRBM's loop body.
    Starts from the GPU run's initial parameters and randomised chains (no
    thermalisation: every sweep costs the same); returns the median step."""
    import oracle as orc
    L = orc.lib()
    rbm = orc.RBM(N, M)
    rbm.set_params(params)
    H = orc.ham(orc.TFI, N, J=-1.0, h=-1.0)
    streams = np.arange(CHAINS, dtype=np.uint32)
    sig = orc.randomize(SEED, 0, streams, N)
    ctr = np.zeros(CHAINS, np.uint64)
    acc = np.zeros(CHAINS, np.uint64)
    O = np.empty((CHAINS, DIM))
    om, F, x, el = np.empty(DIM), np.empty(DIM), np.empty(DIM), np.empty(CHAINS)
    it, err = C.c_int(), C.c_double()
    times, phases, iters = [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        sig, ctr, acc, ss, st = orc.sample_replay(rbm, sig, seed=SEED, n_sweeps=1, n_discard=0,
                                                  ctr=ctr, acc=acc, threads=threads)
        t1 = time.perf_counter()
        L.bl_eloc(C.byref(H), CHAINS, N, M, orc._p(rbm.W), orc._p(rbm.a), orc._p(ss),
                  orc._p(st), orc._p(el), threads)
        t2 = time.perf_counter()
        L.bl_sr_build(CHAINS, N, M, 1, orc._p(ss), orc._p(st), orc._p(el), orc._p(O),
                      orc._p(om), orc._p(F), threads)
        t3 = time.perf_counter()
        L.bl_cg_solve(CHAINS, DIM, orc._p(O), LAMBDA, orc._p(F), TOL, 50000, orc._p(x),
                      C.byref(it), C.byref(err), threads)
        rbm.set_params(rbm.params() - LR * x)
        t4 = time.perf_counter()
        times.append(t4 - t0)
        phases.append([t1 - t0, t2 - t1, t3 - t2, t4 - t3])
        iters.append(it.value)
        print(f"# cpu baseline step: {t4 - t0:.1f} s, {it.value} CG iterations", file=sys.stderr,
              flush=True)
    k = int(np.argsort(times)[len(times) // 2])
    return {
        "value": 1.0 / float(np.median(times)), "unit": "SR steps/s", "cores": threads,
        "kind": "port",
        "sample": (f"{steps} full steps timed, median taken: {CHAINS} chains x 1 sweep at "
                   f"N={N}, M={M}, E_loc, O materialised ({CHAINS} x {DIM} doubles), centring "
                   f"and F, CG on O_c to a {TOL:g} relative residual, SGD update; "
                   f"{threads} threads"),
        "step_s": times, "cg_iters": iters,
        "median_step_phase_s": dict(zip(["sample", "eloc", "build_O_F", "cg_and_update"],
                                        phases[k])),
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                 "threads_used": threads},
    }


def copy_probe():
    """Calibration load for the PMC traffic counters: four device copies of
    1 GiB each (1 GiB read and 1 GiB written per copy)."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_copy": 4 * n, "copies": 4}), flush=True)


def stream_probe(reps=400):
    """What plain streaming kernels achieve at the byte counts of one CG
    product: torch's sum over the product's algorithmic bytes (a read), and
    copies of that size and of the backward product's chunk partials (a read
    and a write each), timed back to back with events on torch's stream."""
    import torch
    out = {}
    sizes = {"sum_28MB": KERNEL_WORK["bwd_mfma"], "copy_28MB": KERNEL_WORK["bwd_mfma"],
             "copy_13.6MB": 64 * KP * NP * 4}
    for name, nbytes in sizes.items():
        a = torch.ones(nbytes // 4, device="cuda")
        b = torch.empty_like(a)
        op = (lambda: a.sum()) if name.startswith("sum") else (lambda: b.copy_(a))
        moved = nbytes * (1 if name.startswith("sum") else 2)
        for _ in range(20):
            op()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            op()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        out[name] = {"bytes_moved": moved, "avg_us": us, "gbs": moved / us / 1e3,
                     "hbm_frac": moved / us / 1e3 / HBM_PEAK_GBS}
    print(json.dumps({"stream_probe": out, "reps": reps}), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC measurement, or
    None when there is none for it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["hbm_bytes_per_launch"]


def precision_errors(yq):
    """The mixed path's error against FP64 on identical configurations.
    TFI N=100, alpha=4, 16,384 samples: E (mean and per-sample E_loc), F, the
    operator S v (random v) and the SR update delta-theta ((S + 1e-3 I) x = F,
    the refined mixed solve vs an FP64 solve to 1e-12).  Complex XXZ N=64,
    alpha=2, 4096 chains x 4 sweeps (Sz = 0, exchange moves): E_loc mean and F
    with float theta and float E_loc, S v with the split-bf16 explicit S, and
    delta-theta from its FP64-refined Cholesky solve, against the FP64 explicit
    build.  Prints one JSON line."""
    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))
    out = {}
    ctx = yq.Context(0)
    ctx.rbm_create(N, M)
    ctx.init_random(SEED, 0.01)
    ctx.ham_set(yq.TFI, N, J=-1.0, h=-1.0)
    ctx.sampler_config(yq.LOCAL, CHAINS, precision=yq.FP64, seed=SEED)
    ctx.sample(1, 20)
    sig, _ = ctx.samples()
    v = np.random.default_rng(SEED).normal(size=DIM)
    res = {}
    for prec in (yq.FP64, yq.MIXED):
        ctx.samples_set(sig, precision=prec)
        e, mean, _ = ctx.eloc()
        ctx.sr_build(yq.SOLVE_CG_FACTORED, precision=prec)
        F = ctx.sr_get(yq.SR_F)
        Sv = ctx.sr_apply(0.0, v)
        x, _, it = ctx.sr_solve(LAMBDA, tol=1e-12 if prec == yq.FP64 else TOL, max_iter=50000)
        res[prec] = (e, mean, F, Sv, x, it)
    a, b = res[yq.MIXED], res[yq.FP64]
    out["tfi_n100_a4"] = {
        "samples": int(sig.shape[0]),
        "energy_rel_err": abs(a[1] - b[1]) / abs(b[1]), "eloc_rel_err": rel(a[0], b[0]),
        "F_rel_err": rel(a[2], b[2]), "Sv_rel_err": rel(a[3], b[3]),
        "dtheta_rel_err": rel(a[4], b[4]), "cg_iters": [b[5], a[5]]}
    ctx.close()
    ctx = yq.Context(0)
    n2, m2 = 64, 128
    ctx.rbm_create(n2, m2, dtype=np.complex128)
    ctx.init_random(SEED, 0.01)
    ctx.ham_set(yq.XXZ, n2, J=1.0, delta=1.0)
    ctx.sampler_config(yq.EXCHANGE, 4096, precision=yq.FP64, seed=SEED)
    ctx.randomize(n2 // 2)
    ctx.sample(4, 20)
    sig, _ = ctx.samples()
    vc = None  # a random complex vector for S v, drawn once the dimension is known
    res = {}
    for prec in (yq.FP64, yq.MIXED):
        ctx.samples_set(sig, precision=prec)
        e, mean, _ = ctx.eloc()
        # the explicit build: FP64 S, or split-bf16 S with the FP64-refined solve
        ctx.sr_build(yq.SOLVE_EXPLICIT, precision=prec)
        if vc is None:
            r = np.random.default_rng(SEED)
            vc = r.normal(size=ctx.sr_dim) + 1j * r.normal(size=ctx.sr_dim)
        Sv = ctx.sr_apply(0.0, vc)
        x, _, _ = ctx.sr_solve(LAMBDA, tol=1e-12 if prec == yq.FP64 else TOL, max_iter=1)
        res[prec] = (e, mean, ctx.sr_get(yq.SR_F), Sv, x)
    a, b = res[yq.MIXED], res[yq.FP64]
    out["xxz_n64_a2_complex"] = {
        "samples": int(sig.shape[0]),
        "energy_rel_err": abs(a[1] - b[1]) / abs(b[1]), "eloc_rel_err": rel(a[0], b[0]),
        "F_rel_err": rel(a[2], b[2]), "Sv_rel_err": rel(a[3], b[3]),
        "dtheta_rel_err": rel(a[4], b[4])}
    ctx.close()
    print(json.dumps(out), flush=True)


def ff_bench(yq):
    """The FeedForward per-sample Jacobian at its configuration: Conv1D(1 -> 8,
    width 5) -> LnCosh -> FullyConnected(320 -> 64) -> Tanh -> Mean on N = 40
    spins (20,520 parameters), float, a batch of 8192 Sz = 0 configurations;
    kernel time from HIP events on the launch, achieved bytes (the Jacobian
    written once, spins and parameters read) against the HBM peak.  Prints one
    JSON line."""
    n, C_, K, H, ns = 40, 8, 5, 64, 8192
    ctx = yq.Context(0)
    ctx.ff_create(n, C_, K, H, dtype=np.float32)
    ctx.ff_init_random(SEED, 3)
    rng = np.random.default_rng(SEED)
    sig = np.ones((ns, n), np.int8)
    for s in range(ns):
        sig[s, rng.choice(n, n // 2, replace=False)] = -1
    npar = ctx.ff_npar
    ctx.ff_jacobian(sig)  # warm-up
    ctx._L.yq_prof_enable(ctx._h, b"ff_jacobian")
    for _ in range(5):
        ctx.ff_jacobian(sig)
    tot, cnt = C.c_double(), C.c_longlong()
    ctx._L.yq_prof_read(ctx._h, b"ff_jacobian", C.byref(tot), C.byref(cnt))
    ctx._L.yq_prof_enable(ctx._h, None)
    avg = tot.value / 1e3 / cnt.value
    nbytes = ns * npar * 4 + ns * n + npar * 4 + ns * 4
    out = {"workload": "FeedForward Conv1D(1->8,k5)-LnCosh-FC(320->64)-Tanh-Mean, N=40, "
                       f"{npar} params, float, batch {ns}",
           "jacobian_us": avg * 1e6, "jacobian_bytes": nbytes,
           "jacobian_gbs": nbytes / avg / 1e9, "hbm_frac": nbytes / avg / 1e9 / HBM_PEAK_GBS}
    ctx.close()
    print(json.dumps(out), flush=True)


def cg_probe(yq):
    """64 mixed CG iterations on 16,384 random configurations at N=100, M=400,
    launched eagerly (no graphs), as a short load for counter passes."""
    ctx = yq.Context(0)
    ctx.graph_mode(0)
    ctx.rbm_create(N, M)
    ctx.init_random(3, 0.05)
    ctx.ham_set(yq.TFI, N, J=-1.0, h=-1.0)
    sig = np.random.default_rng(0).choice([-1, 1], size=(NS, N)).astype(np.int8)
    ctx.samples_set(sig, precision=yq.MIXED)
    ctx.sr_build(yq.SOLVE_CG_FACTORED, precision=yq.MIXED)
    _, _, it = ctx.sr_solve(LAMBDA, tol=1e-30, max_iter=64, allow_unconverged=True)
    print(json.dumps({"cg_probe_iterations": it}), flush=True)
    ctx.close()


def syrk_probe(yq):
    """One FP64 explicit SR build at this workload (16,384 samples after 20
    thermalisation sweeps, S = O_c^T O_c / Ns on the FP64 MFMA SYRK) and its
    Cholesky solve, as a short load for counter passes; prints the SYRK time
    from HIP events on its launch."""
    ctx = _ctx(yq, yq.FP64)
    ctx.sample(1, 20)
    ctx._L.yq_prof_enable(ctx._h, b"syrk64,dpotrf")
    ctx.sr_build(yq.SOLVE_EXPLICIT)
    _, res, _ = ctx.sr_solve(LAMBDA)
    out = {"resid": res}
    tot, cnt = C.c_double(), C.c_longlong()
    for lab in ("syrk64", "dpotrf"):
        ctx._L.yq_prof_read(ctx._h, lab.encode(), C.byref(tot), C.byref(cnt))
        out[lab + "_ms"] = tot.value / max(cnt.value, 1)
    flop = float(DIM) * DIM * NS   # the lower half: dim^2 Ns / 2 multiply-adds
    out["syrk64_tflops"] = flop / out["syrk64_ms"] / 1e9
    out["syrk64_frac_of_fp64_mfma_peak"] = out["syrk64_tflops"] / FP64_MFMA_PEAK_TFLOPS
    ctx._L.yq_prof_enable(ctx._h, None)
    ctx.close()
    print(json.dumps(out), flush=True)


def batched(yq, warmup, steps):
    """Many machines in one context: 32 TFI machines, N ~ U{16, 24, ..., 128},
    alpha ~ U{1, 2, 4}, h evenly spaced over [0.5, 1.5], 2,048 samples each,
    float sampling and the mixed factored SR (one persistent sampler launch and
    one E_loc launch for all machines, the CG iterations of all machines as one
    set of graph-captured launches; also timed with one CG per machine), against
    32 sequential plain-FP64 GPU runs of the same machines (FP64 sampling and
    FP64 factored CG) from the same parameters and chain seeds.  Prints one
    JSON line; the energies of every machine and step are compared."""
    rng = np.random.default_rng(1234)
    Ns = rng.choice(np.arange(16, 129, 8), 32)
    alphas = rng.choice([1, 2, 4], 32)
    hs = np.linspace(0.5, 1.5, 32)
    ms = [dict(n=int(n), m=int(a * n), chains=2048, seed=77 + q,
               ham=(yq.TFI, dict(J=-1.0, h=-float(h))))
          for q, (n, a, h) in enumerate(zip(Ns, alphas, hs))]
    kw = dict(n_sweeps=1, solver=yq.SOLVE_CG_FACTORED, lam=LAMBDA, cg_tol=TOL,
              cg_max_iter=50000, lr=LR, allow_unconverged=True)

    def run_batched(sample_precision, grouped=True):
        ctx = yq.Context(0)
        ctx.multi_create(ms, sweeper=yq.LOCAL, precision=sample_precision)
        ctx.multi_init_random(11, 0.01)
        ctx.multi_set_grouped(grouped)
        p0 = [ctx.multi_get_params(q) for q in range(len(ms))]
        e_b, iters = [], []
        for i in range(warmup):
            e_b.append([o["energy"] for o in ctx.multi_sr_step(n_discard=20 if i == 0 else 0,
                                                               precision=yq.MIXED, **kw)])
        t0 = time.perf_counter()
        for i in range(steps):
            outs = ctx.multi_sr_step(n_discard=0 if (warmup or i) else 20, precision=yq.MIXED,
                                     **kw)
            e_b.append([o["energy"] for o in outs])
            iters.append([o["cg_iters"] for o in outs])
        dt = time.perf_counter() - t0
        ctx.close()
        return dt, np.array(e_b).T, iters, p0

    dt_b, eb, iters, p0 = run_batched(yq.MIXED)
    # the same step with one CG per machine (concurrent streams): the same
    # energies bit for bit
    dt_u, eu, _, _ = run_batched(yq.MIXED, grouped=False)
    same_as_ungrouped = bool(np.array_equal(eb, eu))
    # the same batched launch with FP64 sampling: its chains follow the FP64
    # runs' chains, so the energies differ only through the mixed SR solve
    dt_b64, eb64, _, _ = run_batched(yq.FP64)
    e_s, dt_s = [], 0.0
    for q, mc in enumerate(ms):
        c = yq.Context(0)
        c.rbm_create(mc["n"], mc["m"])
        c.set_params(p0[q])
        c.ham_set(yq.TFI, mc["n"], **mc["ham"][1])
        c.sampler_config(yq.LOCAL, mc["chains"], precision=yq.FP64, seed=mc["seed"])
        e = []
        for i in range(warmup):
            e.append(c.sr_step(n_discard=20 if i == 0 else 0, precision=yq.FP64, **kw)["energy"])
        t0 = time.perf_counter()
        for i in range(steps):
            e.append(c.sr_step(n_discard=0 if (warmup or i) else 20, precision=yq.FP64,
                               **kw)["energy"])
        dt_s += time.perf_counter() - t0
        e_s.append(e)
        c.close()
        print(f"# sequential FP64 machine {q}: N={mc['n']} M={mc['m']} done", file=sys.stderr,
              flush=True)
    es = np.array(e_s)            # machines x steps
    rel = np.abs(eb - es) / np.abs(es)
    rel64 = np.abs(eb64 - es) / np.abs(es)
    dims = [mc["n"] * mc["m"] + mc["n"] + mc["m"] for mc in ms]
    line = {
        "metric": "batched SR steps/s, 32 TFI machines of uneven N (one step advances all 32)",
        "value": steps / dt_b, "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * dt_b / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "32 x 1D TFI PBC, N ~ U{16..128 step 8}, alpha ~ U{1,2,4}, "
                               "h = linspace(0.5, 1.5, 32), 2048 chains x 1 sweep each, "
                               "persistent chains, params N(0, 0.01^2)",
                   "machines": [[mc["n"], mc["m"]] for mc in ms], "parameters": int(sum(dims)),
                   "solver": "mixed factored CG (split-bf16 MFMA) refined in FP64 to 1e-8",
                   "lambda": LAMBDA, "lr": LR},
        "ungrouped_ms_per_step": 1e3 * dt_u / steps,
        "grouped_equals_ungrouped": same_as_ungrouped,
        "sequential_fp64_ms_per_step": 1e3 * dt_s / steps,
        # the comparison at the matched energy tolerance (1e-4 relative) is the
        # batched path with FP64 sampling; float sampling drifts past it
        "speedup_vs_sequential_fp64_at_tolerance": dt_s / dt_b64,
        "fp64_sampling_within_1e-4": bool(rel64.max() < 1e-4),
        "speedup_vs_sequential_fp64_float_sampling": dt_s / dt_b,
        "float_sampling_within_1e-4": bool(rel.max() < 1e-4),
        "energy_rel_diff_max": float(rel.max()),
        "energy_rel_diff_median": float(np.median(rel)),
        "energy_rel_diff_max_per_step": [float(v) for v in rel.max(axis=0)],
        "energy_note": "per machine and step (warm-up steps included), batched float "
                       "sampling vs the FP64 runs: float theta changes a few accept "
                       "decisions, so the chains and then the parameters drift apart; with "
                       "FP64 sampling in the same batched launch only the mixed solve differs",
        "fp64_sampling_ms_per_step": 1e3 * dt_b64 / steps,
        "fp64_sampling_energy_rel_diff_max": float(rel64.max()),
        "fp64_sampling_energy_rel_diff_max_per_step": [float(v) for v in rel64.max(axis=0)],
        "cg_iters_range": [int(np.min(iters)), int(np.max(iters))],
    }
    print(json.dumps(line))


def measured_valu(kernel):
    """VALU figures of a VALU-bound kernel from the committed SQ counter pass,
    or None."""
    try:
        with open(VALU_FILE) as f:
            return json.load(f).get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return None


def _rel_diff(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.abs(a - b) / np.abs(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--fp64-explicit-steps", type=int, default=-1,
                    help="timed steps of the FP64 explicit path (default: --steps)")
    ap.add_argument("--precond", type=int, default=0)
    ap.add_argument("--warm-start", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=0,
                    help="time the mixed products' tile shapes per operand shape")
    ap.add_argument("--graph-chunk", type=int, default=16,
                    help="CG iterations per captured HIP graph of the mixed solve (even)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--stream-probe", action="store_true",
                    help="time plain streaming kernels at the CG products' byte counts")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the copy that calibrates the PMC traffic counters")
    ap.add_argument("--batched", action="store_true",
                    help="the 32-machine batched step against 32 sequential FP64 runs")
    ap.add_argument("--ff", action="store_true",
                    help="the FeedForward Jacobian kernel at N=40, batch 8192")
    ap.add_argument("--errors", action="store_true",
                    help="the mixed path's error against FP64 on identical configurations")
    ap.add_argument("--syrk-probe", action="store_true",
                    help="only one FP64 explicit SR build and solve at this shape (a short "
                         "load for counter passes of the SYRK)")
    ap.add_argument("--cg-probe", action="store_true",
                    help="only 64 eager mixed CG iterations at this shape (a short load for "
                         "counter passes of the CG kernels)")
    args = ap.parse_args()
    if args.stream_probe:
        stream_probe()
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import yannq_amd as yq
    if args.cg_probe:
        cg_probe(yq)
        return
    if args.syrk_probe:
        syrk_probe(yq)
        return
    if args.errors:
        precision_errors(yq)
        return
    if args.ff:
        ff_bench(yq)
        return
    if args.batched:
        batched(yq, args.warmup, args.steps)
        return
    global AUTOTUNE, GRAPH_CHUNK
    AUTOTUNE = bool(args.autotune)
    GRAPH_CHUNK = args.graph_chunk
    opts = dict(precond=args.precond, warm_start=args.warm_start)
    W, K = args.warmup, args.steps

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    # pick the dominant kernel from a profiled warm-up of the timed path
    _, _, _, prof0 = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, max(W, 1), 2,
                              prof_label=",".join(KERNEL_WORK), **opts)
    dominant = max(prof0, key=lambda k: prof0[k][0])
    log(f"kernel selection pass done: dominant kernel {dominant}")

    # timed pass: no per-kernel events, so the CG runs as captured graph chunks
    dt, outs, energies, _ = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, W, K, **opts)
    ms = 1e3 * dt / K
    value = K / dt
    iters = [o["cg_iters"] for o in outs]
    errs = [o["cg_err"] for o in outs]
    log(f"mixed pass done: {value:.2f} SR steps/s, CG iterations {min(iters)}..{max(iters)}")
    # kernel timing pass: the same steps again with HIP events around every
    # launch of the dominant kernel (eager launches, on the kernel's stream)
    dtk, _, _, prof = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, W, K, prof_label=dominant,
                               **opts)
    tot_ms, launches = prof[dominant]
    amount = KERNEL_WORK[dominant]
    avg_s = tot_ms / 1e3 / max(launches, 1)
    achieved = amount / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(dominant), "avg_launch_us": avg_s * 1e6,
                "launches": launches,
                "algorithmic_bytes_per_launch": amount,
                "share_of_step": tot_ms / (1e3 * dtk),
                "timing": "HIP events bound to each launch of the kernel (hipExtLaunchKernel) "
                          "in a second pass over the same steps; the average includes the "
                          "launches that return at once after the CG converged"}

    # every kernel of the step from the profiled pass (2 steps): average launch,
    # launches and time per step, algorithmic bytes and HBM fraction; the
    # sampler and the E_loc ratios are VALU-bound (log-cosh differences)
    per_kernel = {}
    for lab, (t_ms, cnt) in prof0.items():
        if cnt == 0:
            continue
        avg = t_ms / 1e3 / cnt
        gbs = KERNEL_WORK[lab] / avg / 1e9
        d = {"avg_launch_us": avg * 1e6, "launches_per_step": cnt / 2, "ms_per_step": t_ms / 2,
             "algorithmic_bytes_per_launch": KERNEL_WORK[lab], "achieved_gbs": gbs,
             "hbm_frac": gbs / HBM_PEAK_GBS,
             "bound": "valu" if lab in VALU_KERNELS else "hbm"}
        if lab in VALU_KERNELS:  # one sweep / one pass: Ns * n * m terms
            d["logcosh_diff_terms_per_s"] = NS * N * M / avg
            d["valu_counters"] = measured_valu(lab)
        per_kernel[lab] = d

    extra = {}
    if not args.no_fp64:
        # the FP64 paths and the mixed solver on FP64 samples, over the same
        # warm-up and timed steps (the same step window as the mixed pass)
        dtc, outc, enc, profc = run_path(yq, yq.FP64, yq.SOLVE_CG_FACTORED, W, K,
                                         prof_label="factored_fwd64,factored_bwd64", **opts)
        log(f"FP64 factored CG pass done: {K / dtc:.2f} SR steps/s")
        dtc2, _, _, _ = run_path(yq, yq.FP64, yq.SOLVE_CG_FACTORED, W, K, **opts)
        dts, outs_s, ens, _ = run_path(yq, yq.FP64, yq.SOLVE_CG_FACTORED, W, K,
                                       solve_precision=yq.MIXED, **opts)
        log("mixed solver on FP64 samples done")
        ke = K if args.fp64_explicit_steps < 0 else args.fp64_explicit_steps
        dte, oute, ene, _ = run_path(yq, yq.FP64, yq.SOLVE_EXPLICIT, W, ke)
        log(f"FP64 explicit pass done: {ke / dte:.3f} SR steps/s")
        fp64_iters = [o["cg_iters"] for o in outc]
        errs += [o["cg_err"] for o in outc + outs_s + oute]
        # per-launch FP64 products of the comparator (one label covers the MFMA
        # product and its small reduction kernel) against their FP64-MFMA floor
        comp = {}
        for lab, (t_ms, cnt) in profc.items():
            if cnt:
                avg_us = 1e3 * t_ms / cnt
                flop = 2.0 * NS * N * M
                comp[lab] = {"avg_launch_us": avg_us, "launches": cnt,
                             "algorithmic_flop": flop,
                             "floor_us": flop / (FP64_MFMA_PEAK_TFLOPS * 1e6),
                             "frac_of_fp64_mfma_peak": flop / (avg_us * 1e-6) /
                             (FP64_MFMA_PEAK_TFLOPS * 1e12)}
        ws = len(energies)
        rel_f = _rel_diff(energies, enc)
        rel_s = _rel_diff(ens, enc)
        rel_e = _rel_diff(ene[:ws], enc[:len(ene)])
        last = outs[-1]
        last_c = outc[-1]
        # the timed steps' energy differences against their Monte Carlo
        # errors (16,384 independent chains per step)
        zs = [abs(a["energy"] - b["energy"]) / np.sqrt((a["energy_var"] + b["energy_var"]) / NS)
              for a, b in zip(outs, outc)]
        em = float(np.mean([o["energy"] for o in outs]))
        ef = float(np.mean([o["energy"] for o in outc]))
        extra = {
            "fp64_factored_cg_steps_per_s": K / dtc,
            "fp64_factored_cg_steps_per_s_repeat": K / dtc2,
            "fp64_factored_cg_timing": "first figure from a pass with HIP events on the FP64 "
                                       "products, the repeat without them",
            "fp64_explicit_steps_per_s": ke / dte,
            "fp64_explicit_timed_steps": ke,
            "mixed_vs_fp64_factored_cg": value / (K / dtc2),
            "mixed_vs_fp64_explicit": value / (ke / dte),
            "ratio_note": "same warm-up and timed steps (from the same parameters and chain "
                          "seeds) for every path; the FP64 comparator's MFMA products run at "
                          "the fraction of the FP64 MFMA peak given in fp64_comparator_products",
            "fp64_cg_iters": fp64_iters,
            "fp64_comparator_products": comp,
            "mixed_solver_fp64_sampling_steps_per_s": K / dts,
            # energies of every step (warm-up and timed) of each path
            "energy_rel_diff_mixed_vs_fp64_max": float(rel_f.max()),
            "energy_rel_diff_mixed_vs_fp64_per_step": [float(v) for v in rel_f],
            "energy_rel_diff_mixed_solver_vs_fp64_max": float(rel_s.max()),
            "energy_rel_diff_fp64_explicit_vs_fp64_cg_max": float(rel_e.max()),
            "energy_comparison": (
                "mixed_vs_fp64: the timed mixed path (float sampling) against the FP64 "
                "factored CG path, all %d steps; float theta changes some accept decisions, so "
                "the two Markov chains and then the parameters drift apart stochastically. "
                "mixed_solver_vs_fp64: the mixed solve fed by the FP64 path's own chains, all "
                "%d steps: the difference the mixed arithmetic itself makes" % (ws, ws)),
            "energy_timed_window": {
                "mixed_mean": em, "fp64_mean": ef, "rel_diff_of_means": abs(em - ef) / abs(ef),
                "diff_over_mc_err_max": float(max(zs)),
                "diff_over_mc_err_mean": float(np.mean(zs)),
                "note": "per timed step |E_mixed - E_fp64| / sqrt((var_mixed + var_fp64) / Ns): "
                        "the two paths' energies differ by what their Monte Carlo errors "
                        "allow"},
            "energy_last_step": {"mixed": energies[-1], "fp64": enc[-1],
                                 "mixed_solver_fp64_sampling": ens[-1],
                                 "mc_std_err_fp64": float(np.sqrt(last_c["energy_var"] / NS)),
                                 "mc_std_err_mixed": float(np.sqrt(last["energy_var"] / NS))},
        }
    err_max = float(max(errs))
    converged = err_max < TOL
    cpu = None
    if not args.no_cpu_baseline:
        p0 = None
        c0 = _ctx(yq, yq.FP64)
        p0 = c0.get_params()
        c0.close()
        cpu = cpu_baseline(p0, steps=args.cpu_steps)
    line = {
        "metric": "SR steps/sec (sample+E_loc+O+SR solve), TFI N=100 RBM alpha=4, mixed vs FP64",
        "value": value, "unit": "SR steps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1D TFI PBC N=100 h=1 (J=-1), real RBM alpha=4 (40500 params), "
                               "16384 chains x 1 sweep = 16384 samples/step, persistent chains",
                   "precision": "mixed: f32 sampling/E_loc, split-bf16x3 MFMA CG products "
                                "(f32 accumulate), f64 refinement to 1e-8, f64 update",
                   "solver": "factored CG, Jacobi preconditioner=%d, warm start=%d, "
                             "tile autotune=%d, graph chunk=%d iterations"
                             % (args.precond, args.warm_start, args.autotune, args.graph_chunk),
                   "lambda": LAMBDA, "lr": LR, "cg_tol": TOL},
        "samples_per_s": value * CHAINS,
        "median_step_ms": float(np.median([o["wall_ms"] for o in outs])),
        "cg_iters": iters,
        "cg_err_max": err_max,
        "cg_converged": converged,
        "energy": energies[-1],
        "phase_ms": {"sample": float(np.mean([o["sample_ms"] for o in outs])),
                     "eloc_build_solve_update": float(np.mean([o["solve_ms"] for o in outs]))},
        "roofline": roofline,
        "per_kernel": per_kernel,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)
    if not converged:
        print(f"bench.py: a timed solve ended above the {TOL:g} relative residual "
              f"(max {err_max:.3g})", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
