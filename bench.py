"""SR steps/s of the VMC step (sample + E_loc + O + SR solve + update) on one
MI355X: 1D transverse-field Ising chain, N=100, real RBM alpha=4 (40,500
parameters), 16,384 samples per step.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
    python bench.py --batched [--steps K] [--warmup W]   (32 machines, see batched())

Timed path (value): mixed precision -- sampling and E_loc in float, O kept
factored (spins and tanh(theta)), matrix-free CG whose products run as
split-bf16 MFMA with float accumulation, FP64 iterative refinement to a 1e-8
relative residual, FP64 update.  The inner CG runs as captured HIP-graph
chunks of 8 iterations that stop themselves on the device.  The CG starts from zero without a
preconditioner, as the reference's solveCG does (--precond/--warm-start
switch on the Jacobi preconditioner and warm starts; both solve the same
system to the same tolerance, and on this workload both cost iterations).
Also timed on the same start state: the plain FP64 path (explicit S =
O_c^T O_c / Ns and Cholesky, the reference's default solver), the FP64
factored CG, and the mixed path with both solver options.  Prints one JSON
line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
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

# algorithmic HBM bytes per launch of the kernels that can dominate a step
KERNEL_WORK = {
    # forward product: T' (Ns x (m+1) float) and sigma' (Ns x (n+1) bf16) in, y partials out
    "fwd_mfma": NS * (M + 1) * 4 + NS * (N + 1) * 2 + NS * 4,
    # backward product: T'^T and sigma'^T and y in, the (n+1) x (m+1) chunk partials out
    "bwd_mfma": NS * (M + 1) * 4 + NS * (N + 1) * 2 + NS * 4 + (NS // 256) * KP * NP * 4,
    # CG step: the chunk partials of the dim parameters in; x, r, p, s, u, omean
    # in; x, r, p, s, u and the three bf16 layers of V' out
    "cg_step": (NS // 256) * DIM * 4 + DIM * 4 * 11 + 3 * KP * NP * 2,
    # sampler: spins in/out and the recorded samples (spins, float theta)
    "sample": NS * N * 3 + NS * M * 4,
    # E_loc ratios: theta (float) and spins in, ratios (double) out
    "eloc_ratio": NS * M * 4 + NS * N + NS * N * 8,
}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) of this script, each counter scaled by what it reads
# for a copy of known size (bench.py --copy-probe); written with the command
# lines that produced it
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_tfi_n100_a4.json")


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


def run_path(yq, precision, solver, warmup, steps, precond=0, warm_start=0, prof_label=None):
    """warmup + steps SR epochs from the same seeds; returns the timed seconds,
    the per-step outputs, all energies and the HIP-event kernel timings."""
    ctx = _ctx(yq, precision)
    kw = dict(n_sweeps=1, solver=solver, precision=precision, lam=LAMBDA, cg_tol=TOL,
              cg_max_iter=50000, lr=LR, precond=precond, warm_start=warm_start,
              allow_unconverged=True)
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


def cpu_baseline(iters_per_step):
    """The oracle's threaded C++ restatement of the reference CG step
    (SamplerMT + constructDelta/constructObs + SRMat::solveCG), timed on a
    bounded sample of chains and scaled linearly in samples."""
    import oracle as orc
    threads = min(16, orc.lib().bl_threads())
    rng = np.random.default_rng(SEED)
    rbm = orc.RBM.random(N, M, 0.01, rng)
    H = orc.ham(orc.TFI, N, J=-1.0, h=-1.0)
    c = 2048
    sig = orc.randomize(SEED, 0, np.arange(c), N)
    t0 = time.perf_counter()
    _, _, _, ss, st = orc.sample_replay(rbm, sig, seed=SEED, n_sweeps=1, n_discard=0,
                                        threads=threads)
    t_samp = time.perf_counter() - t0
    el = np.empty(c)
    t0 = time.perf_counter()
    orc.lib().bl_eloc(C.byref(H), c, N, M, orc._p(rbm.W), orc._p(rbm.a), orc._p(ss),
                      orc._p(st), orc._p(el), threads)
    t_eloc = time.perf_counter() - t0
    O = np.empty((c, DIM))
    om, F = np.empty(DIM), np.empty(DIM)
    t0 = time.perf_counter()
    orc.lib().bl_sr_build(c, N, M, 1, orc._p(ss), orc._p(st), orc._p(el), orc._p(O),
                          orc._p(om), orc._p(F), threads)
    t_build = time.perf_counter() - t0
    x = np.empty(DIM)
    it, err = C.c_int(), C.c_double()
    t0 = time.perf_counter()
    orc.lib().bl_cg_solve(c, DIM, orc._p(O), LAMBDA, orc._p(F), 1e-30, 20, orc._p(x),
                          C.byref(it), C.byref(err), threads)
    t_cg_iter = (time.perf_counter() - t0) / max(it.value, 1)
    scale = CHAINS / c
    t_step = scale * (t_samp + t_eloc + t_build + t_cg_iter * iters_per_step)
    return {
        "value": 1.0 / t_step, "unit": "SR steps/s", "cores": threads, "kind": "port",
        "sample": (f"{c} chains x 1 sweep, E_loc, O build and {it.value} CG applies on "
                   f"{c} samples at N={N}, M={M}, scaled x{scale:.0f} to {CHAINS} samples "
                   f"with {iters_per_step} CG iterations per step (the GPU FP64 CG count)"),
        "phase_s_at_full_size": {"sample": scale * t_samp, "eloc": scale * t_eloc,
                                 "build_O_F": scale * t_build,
                                 "cg": scale * t_cg_iter * iters_per_step},
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
    with float theta and float E_loc (the SR build is FP64 for complex
    machines).  Prints one JSON line."""
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
    res = {}
    for prec in (yq.FP64, yq.MIXED):
        ctx.samples_set(sig, precision=prec)
        e, mean, _ = ctx.eloc()
        ctx.sr_build(yq.SOLVE_CG, precision=yq.FP64)
        res[prec] = (e, mean, ctx.sr_get(yq.SR_F))
    a, b = res[yq.MIXED], res[yq.FP64]
    out["xxz_n64_a2_complex"] = {
        "samples": int(sig.shape[0]),
        "energy_rel_err": abs(a[1] - b[1]) / abs(b[1]), "eloc_rel_err": rel(a[0], b[0]),
        "F_rel_err": rel(a[2], b[2])}
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
        "speedup_vs_sequential_fp64": dt_s / dt_b,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--precond", type=int, default=0)
    ap.add_argument("--warm-start", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=0,
                    help="time the mixed products' tile shapes per operand shape")
    ap.add_argument("--graph-chunk", type=int, default=16,
                    help="CG iterations per captured HIP graph of the mixed solve (even)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the copy that calibrates the PMC traffic counters")
    ap.add_argument("--batched", action="store_true",
                    help="the 32-machine batched step against 32 sequential FP64 runs")
    ap.add_argument("--ff", action="store_true",
                    help="the FeedForward Jacobian kernel at N=40, batch 8192")
    ap.add_argument("--errors", action="store_true",
                    help="the mixed path's error against FP64 on identical configurations")
    ap.add_argument("--cg-probe", action="store_true",
                    help="only 64 eager mixed CG iterations at this shape (a short load for "
                         "counter passes of the CG kernels)")
    args = ap.parse_args()
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

    # pick the dominant kernel from a profiled warm-up of the timed path
    _, _, _, prof0 = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, max(args.warmup, 1), 2,
                              prof_label=",".join(KERNEL_WORK), **opts)
    dominant = max(prof0, key=lambda k: prof0[k][0])

    # timed pass: no per-kernel events, so the CG runs as captured graph chunks
    dt, outs, energies, _ = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, args.warmup,
                                     args.steps, **opts)
    ms = 1e3 * dt / args.steps
    value = args.steps / dt
    iters = [o["cg_iters"] for o in outs]
    # kernel timing pass: the same steps again with HIP events around every
    # launch of the dominant kernel (eager launches, on the kernel's stream)
    dtk, _, _, prof = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, args.warmup,
                               args.steps, prof_label=dominant, **opts)
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
             "bound": "valu" if lab in ("sample", "eloc_ratio") else "hbm"}
        if lab in ("sample", "eloc_ratio"):  # one sweep / one pass: Ns * n * m terms
            d["logcosh_diff_terms_per_s"] = NS * N * M / avg
        per_kernel[lab] = d

    extra = {}
    fp64_iters = iters
    if not args.no_fp64:
        k = args.fp64_steps
        dte, _, ene, _ = run_path(yq, yq.FP64, yq.SOLVE_EXPLICIT, args.warmup, k)
        dtc, outc, _, _ = run_path(yq, yq.FP64, yq.SOLVE_CG_FACTORED, args.warmup, k, **opts)
        dtp, outp, _, _ = run_path(yq, yq.MIXED, yq.SOLVE_CG_FACTORED, args.warmup, k,
                                   precond=1, warm_start=1)
        fp64_iters = [o["cg_iters"] for o in outc]
        extra = {
            "fp64_explicit_steps_per_s": k / dte,
            "fp64_factored_cg_steps_per_s": k / dtc,
            "mixed_jacobi_warm_cg_steps_per_s": k / dtp,
            "mixed_jacobi_warm_cg_iters": [o["cg_iters"] for o in outp],
            "mixed_vs_fp64_explicit": value / (k / dte),
            "mixed_vs_fp64_factored_cg": value / (k / dtc),
            "fp64_cg_iters": fp64_iters,
            # same seeds and start state: energies of the first steps of each path
            "energy_rel_diff_mixed_vs_fp64": float(np.max(np.abs(
                np.array(energies[:len(ene)]) - np.array(ene)) / np.abs(np.array(ene)))),
        }
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(int(np.median(fp64_iters)))
    line = {
        "metric": "SR steps/sec (sample+E_loc+O+SR solve), TFI N=100 RBM alpha=4, mixed vs FP64",
        "value": value, "unit": "SR steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
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
        "energy": energies[-1],
        "phase_ms": {"sample": float(np.mean([o["sample_ms"] for o in outs])),
                     "eloc_build_solve_update": float(np.mean([o["solve_ms"] for o in outs]))},
        "roofline": roofline,
        "per_kernel": per_kernel,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
