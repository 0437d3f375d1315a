"""Block H.psi throughput of the Smolyak SG4 operator on one MI355X.

Workload (the configuration the headline metric is quoted on): 10D
Henon-Heiles (lambda = 0.111803), Smolyak SG type 4 with L = LB = LG = 8,
HO bases nq(l) = nb(l) = 1 + 2l, block of 16 seeded orthonormal vectors.
A step is one block application H.psi (sub_TabOpPsi_FOR_SGtype4) with the
input block already resident in HBM.

The reported value is the split-precision mode (SG4_PREC_F32X2: mode
contractions on the matrix cores with three bf16 limbs per FP32 operand and
FP32 accumulation; kinetic part, term sum and result in FP64); the plain FP64
apply of the same operator is timed beside it with the same number of steps,
and their ratio is reported.  The eigenvalue error condition of the metric is
measured by block Davidson runs for the 20 lowest levels: FP64, split
precision followed by an FP64 phase, and split precision alone, each
compared with the FP64 levels; each run is timed after an untimed run of the
same kind (first-use set-up of the block widths it applies).

One JSON line: the metric, the roofline of the dominant kernel (the term
kernel's main launch: algorithmic flops per launch over the launch time
measured with HIP events on the plan's stream, against the peak of the unit
the kernel computes on; HBM traffic per launch from the committed rocprofv3
counter summary when one exists for this workload) and the CPU baseline (the
oracle, OpenMP over terms, on a bounded sample of the same workload, with the
host CPU model).  For the 10D L=8 block-16 workload the launch parameters are
the best ones the autotuner found on MI355X (pinned, so that runs and
profiles use the same launch; --autotune re-tunes); other workloads are
autotuned before the timed region (--no-autotune keeps the defaults).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# peak of the unit each mode's term kernel computes on (TFLOP/s, MI355X dense)
PEAK = {"fp64": 78.6, "f32": 157.3, "f32x2": 2500.0}
PEAK_HBM_GBPS = 8000.0
UNIT = {"fp64": "FP64 VALU (= FP64 MFMA peak)", "f32": "FP32 VALU",
        "f32x2": "bf16 MFMA (dense)"}
# sg4_plan_autotune's best launches for 10D L=8 block 16 on MI355X:
# threads, LDS budget, persistent grid, item order (3: classes of terms with
# equal levels in modes 3..D-1 on one XCD; 2: modes 2..D-1)
PINNED = {(10, 8, 16): {"fp64": (512, 163840, 2048, 3), "f32": (768, 163840, 2048, 2),
                        "f32x2": (1024, 163840, 512, 2)}}
# the propagation workload's 2-column applies (64-instance batch below):
# sg4_plan_autotune's best launches, timed over whole applies on MI355X
PINNED_PROP = {"f32x2": (256, 32768, 2048, 3), "fp64": (256, 65536, 0, 3)}
PREC = {"fp64": 0, "f32": 1, "f32x2": 2}
DTYPE = {"fp64": "f64", "f32": "f32", "f32x2": "bf16x3-split/f32-acc"}
KERNEL = {"fp64": "k_hpsi_t1s_f64", "f32": "k_hpsi_t1s_f32", "f32x2": "k_hpsi_mfma"}


def cpu_model():
    """'model name' of /proc/cpuinfo with the family/model/stepping numbers
    (a virtualised host may report only the brand)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}; {os.cpu_count()} logical CPUs visible)")


def cpu_baseline(D, L, nvec, stride, min_seconds=10.0):
    """Oracle (OpenMP over terms) on every `stride`-th term of the same
    workload, repeated until at least `min_seconds` of CPU work (at most 5
    times); the rate is scaled by the sampled share of the algorithmic flops."""
    from oracle import oracle as O
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    nq = np.diff(o.qoff).astype(np.float64)
    per_term = np.array([2.0 * nq[t] * np.sum(1 + 2 * o.lev[t]) for t in range(o.nterms)])
    frac = per_term[::stride].sum() / per_term.sum()
    psi = np.random.default_rng(2026).standard_normal((o.nb_ba, nvec))
    reps, dt = 0, 0.0
    while reps < 5 and (reps == 0 or dt < min_seconds):
        t0 = time.perf_counter()
        o.hpsi(psi, V, -0.5 * np.ones(D), t0=0, t1=-1, tstride=stride)
        dt += time.perf_counter() - t0
        reps += 1
    return {"value": reps * nvec * frac / dt, "unit": "H.psi/s", "cores": O.num_threads(),
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{reps} x every {stride}th Smolyak term ({frac:.3f} of the flops) of the "
                      f"same {D}D L={L} block-{nvec} FP64 apply, {dt:.1f} s in total, "
                      f"scaled by the flop share"}


def measured_traffic(kernel, D, L, nvec):
    """HBM bytes per launch of `kernel` from profiles/traffic_<kernel>_<D>D_L<L>_block<nvec>.md
    (line 'traffic_bytes_per_launch: <n>'), or None."""
    f = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{D}D_L{L}_block{nvec}.md")
    if not os.path.exists(f):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(f).read())
    return float(m.group(1)) if m else None


def reference_flops(b, nvec):
    """Flops of one block apply by the reference's algorithm (B->G, kinetic
    on the grid, G->B; sub_OpPsi_SG4.f90:1352-1666): F1 = 3 sum_t S_t per
    vector, S_t = sum_k 2 N_t n_k over the terms t of the basis b."""
    D, L = b.D, b.L
    lev = b.levels()
    ntab = {(k, l): b.basis1d(k, l)["nq"] for k in range(D) for l in range(L + 1)}
    nq = np.array([[ntab[k, int(l)] for k, l in enumerate(row)] for row in lev], dtype=np.float64)
    return 3.0 * float(np.sum(2.0 * np.prod(nq, axis=1) * nq.sum(axis=1))) * nvec


def propagation_bench(E, torch, args):
    """64 uneven instances (D ~ U{4..12}, L ~ U{4..8}, numpy seed 7) in one
    batched plan, dipole sum_i q_i, 200 piecewise-constant N(0, 0.01)
    amplitudes (seed 7), steps of dt = 0.1 spread evenly over the segments;
    forward propagation of one complex state per instance with the device
    propagator, then J and dJ/deps over the first segments."""
    from elvibrot_amd.propagation import (ChebyshevPropagator, spectral_bounds, propagate_device,
                                          field_gradient_device)
    prec = PREC[args.prec]
    t0 = time.perf_counter()
    rng = np.random.default_rng(7)
    Ds = rng.integers(4, 13, size=64)
    Ls = rng.integers(4, 9, size=64)
    bases = [E.SG4Basis(int(d), int(l)) for d, l in zip(Ds, Ls)]
    op = E.SG4Batch(bases, max_nvec=4)
    eps = np.random.default_rng(7).normal(0.0, 0.01, 200)
    mu = op.grid_dipole()
    setup = time.perf_counter() - t0
    if args.autotune or args.prec not in PINNED_PROP:
        for nv in (2, 4):
            op.autotune(nv, prec)
            if prec != PREC["fp64"]:
                op.autotune(nv, PREC["fp64"])
    else:       # pinned 2-column launches; the 4-column applies inherit them
        for pn in {args.prec, "fp64"}:
            op.set_config(2, PREC[pn], *PINNED_PROP[pn])
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + np.abs(eps).max() * mu.abs().max().item()
    prop = ChebyshevPropagator(op, 0.1, lo - w, hi + w, mu=mu, prec=prec)
    nb = op.basis.nb_ba
    blocks = list(zip(op.row0[:-1], op.row0[1:]))
    psi0 = torch.zeros(1, nb, dtype=torch.complex128, device="cuda")
    target = torch.zeros_like(psi0)
    for r0, r1 in blocks:
        psi0[0, r0] = 1.0
        target[0, r0] = np.sqrt(0.5)
        target[0, r0 + 1] = np.sqrt(0.5)
    spp = max(1, args.propagation // len(eps))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    psiT = psi0
    for c in range(0, len(eps), 10):     # 10 segments per call, a progress line each
        psiT = propagate_device(prop, psiT, eps[c:c + 10], spp)
        print(f"propagation: {(c + 10) * spp} steps, {time.perf_counter() - t1:.1f} s",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t1
    nsteps = spp * len(eps)
    gseg = max(1, args.grad_steps // spp)
    t2 = time.perf_counter()
    J, g = field_gradient_device(prop, psi0, target, eps[:gseg], spp, checkpoint_every=5,
                                 blocks=blocks)
    tg = time.perf_counter() - t2
    napp = nsteps * (prop.nmax - 1)
    return {"workload": "64 uneven Henon-Heiles instances (seed 7), 200-segment field, "
                        f"{nsteps} Chebyshev steps of dt = 0.1",
            "precision": args.prec, "rows": nb, "grid_points": op.basis.sum_nq,
            "setup_s": setup, "spectral_range": [lo, hi], "orders_per_step": prop.nmax,
            "forward_s": tf, "steps_per_s": nsteps / tf,
            "batch_hpsi_per_s": napp / tf, "norm_after": torch.linalg.norm(psiT).item() / 8.0,
            "gradient": {"segments": gseg, "steps": gseg * spp, "seconds": tg, "J": J,
                         "dJ_deps_first": [float(x) for x in g[:4]]},
            "launch": {"nvec2": op.config(2, prec), "nvec4": op.config(4, prec)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", default="f32x2", choices=list(PREC),
                    help="precision of the reported value; FP64 is timed beside it "
                         "(the FP32 VALU mode beside FP64)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune even where a pinned launch exists for the workload")
    ap.add_argument("--launch", default=None,
                    help="threads,lds_budget,grid,xcd for the reported precision (skips its autotune)")
    ap.add_argument("--nlev", type=int, default=20,
                    help="levels of the Davidson eigenvalue check (0: skip it)")
    ap.add_argument("--host-entry", action="store_true",
                    help="also time sg4_apply_host (host buffers, PCIe included)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (x5), for scaling HBM counters")
    ap.add_argument("--no-other", action="store_true", help="do not time the other precision")
    ap.add_argument("--no-graph", action="store_true", help="do not time the graph-replayed apply")
    ap.add_argument("--propagation", type=int, default=0, metavar="NSTEPS",
                    help="only time the batched propagation workload: 64 seeded uneven "
                         "Henon-Heiles instances, 200-segment field, NSTEPS Chebyshev steps of "
                         "dt = 0.1 (forward), plus the field gradient over the first "
                         "--grad-steps steps")
    ap.add_argument("--grad-steps", type=int, default=25)
    ap.add_argument("--D", type=int, default=10)
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--nvec", type=int, default=16)
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(terms would be sharded across GPUs by flops, with one all-reduce of "
              "the packed block per apply)", file=sys.stderr)
        sys.exit(2)

    import torch
    import elvibrot_amd as E

    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, device="cuda:0", dtype=torch.float64)
        c = torch.empty_like(a)
        for _ in range(5):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "bytes_written": n * 8, "launches": 5}))
        return

    if args.propagation:
        print(json.dumps(propagation_bench(E, torch, args)), flush=True)
        return

    D, L, nvec = args.D, args.L, args.nvec
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    g = torch.Generator(device="cuda:0").manual_seed(2026)
    X = torch.randn(b.nb_ba, nvec, device="cuda:0", dtype=torch.float64, generator=g)
    X, _ = torch.linalg.qr(X)                       # orthonormal block
    Xc = X.T.contiguous()                            # (nvec, nb_ba) = column-major block
    Y = torch.empty_like(Xc)

    def choose_launch(pname):
        pinned = PINNED.get((D, L, nvec), {}).get(pname)
        if pname == args.prec and args.launch:
            op.set_config(nvec, PREC[pname], *(int(v) for v in args.launch.split(",")))
        elif pinned and not args.autotune and not args.no_autotune:
            op.set_config(nvec, PREC[pname], *pinned)
        elif not args.no_autotune:
            op.autotune(nvec, PREC[pname])

    def timed(p, steps, warmup):
        for _ in range(warmup):
            op.apply(Xc, Y, p)
        torch.cuda.synchronize()
        op.timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            op.apply(Xc, Y, p)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kt = op.kernel_times()
        op.timing(False)
        return dt, float(np.mean(kt)) * 1e-3

    # the reported mode
    choose_launch(args.prec)
    split = op.split_info()[0]
    prec = PREC[args.prec]
    dt, kavg = timed(prec, args.steps, args.warmup)
    st = op.stats(nvec)
    mf, ngm = op.mfma_stats() if args.prec == "f32x2" else (0.0, -1)
    Yrep = Y.clone()
    # the other precision, same number of steps
    other = None
    oname = "fp64" if args.prec != "fp64" else "f32"
    if not args.no_other:
        choose_launch(oname)
        dto, kavgo = timed(PREC[oname], args.steps, args.warmup)
        sto = op.stats(nvec)
        other = {"precision": oname, "H·psi/s": nvec * args.steps / dto,
                 "ms_per_step": dto / args.steps * 1e3, "kernel": KERNEL[oname],
                 "kernel_ms": kavgo * 1e3,
                 "roofline": {"achieved": sto["flops"] / kavgo / 1e12, "peak": PEAK[oname],
                              "frac": sto["flops"] / kavgo / 1e12 / PEAK[oname],
                              "unit": "TFLOP/s"}}
        if oname == "fp64":   # H.psi error of the reported mode against FP64, same input
            e2 = (torch.linalg.norm(Yrep - Y, dim=1) / torch.linalg.norm(Y, dim=1)).max().item()
            other["hpsi_rel_err_2norm_of_reported_vs_fp64"] = e2
    # the apply replayed from one captured HIP graph, in the rows layout the
    # device-side drivers use (kinetic rows and term kernel in one graph)
    graph = None
    if not args.no_graph:
        Xr = X.contiguous()                        # (nb_ba, nvec): psi[row*nvec + v]
        Yr = torch.empty_like(Xr)
        graph = {"layout": "rows", "steps": args.steps}
        for pname in dict.fromkeys([args.prec] + ([oname] if other is not None else [])):
            choose_launch(pname)
            op.apply_graph(Xr, Yr, PREC[pname])    # capture (not timed)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                op.apply_graph(Xr, Yr, PREC[pname])
            torch.cuda.synchronize()
            tg = time.perf_counter() - t0
            graph[pname] = {"H·psi/s": nvec * args.steps / tg, "ms_per_step": tg / args.steps * 1e3}
    F1 = reference_flops(b, nvec)
    kname = KERNEL[args.prec] if split or args.prec == "f32x2" else (
        "k_hpsi_t1_" + ("f64" if args.prec == "fp64" else "f32"))
    traffic = measured_traffic(kname, D, L, nvec)
    achieved = st["flops"] / kavg / 1e12
    roof = {"bound": "mfma" if args.prec == "f32x2" else "valu",
            "achieved": achieved, "peak": PEAK[args.prec], "unit": "TFLOP/s",
            "frac": achieved / PEAK[args.prec], "traffic": traffic,
            "kernel": kname, "compute_unit": UNIT[args.prec], "split_path": split,
            "kernel_ms": kavg * 1e3,
            "algorithmic_flops_per_launch": st["flops"],
            "algorithmic_bytes_per_launch": st["bytes"],
            "hbm_GBps_algorithmic": st["bytes"] / kavg / 1e9,
            "F1_reference_flops_per_apply": F1,
            "F1_over_kernel_TFLOPs": F1 / kavg / 1e12}
    if args.prec == "f32x2":
        # the split kernel's algorithmic intensity (flops over compulsory
        # bytes) is below the bf16 ridge (peak flops over peak HBM bandwidth),
        # so its roofline is the HBM one: algorithmic bytes against the HBM
        # peak, the flop rates beside it
        roof.update({"bound": "hbm", "achieved": st["bytes"] / kavg / 1e9, "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": st["bytes"] / kavg / 1e9 / PEAK_HBM_GBPS,
                     "intensity_flop_per_byte": st["flops"] / st["bytes"],
                     "ridge_flop_per_byte": PEAK["f32x2"] * 1e12 / (PEAK_HBM_GBPS * 1e9),
                     "limit": "latency of the LDS passes (one dependency chain per tile, a "
                              "barrier per pass, 4 waves per SIMD); the scatter's memory-side "
                              "FP64 atomics add ~20 %",
                     "atomic_adds_per_launch": b.sum_nb * nvec,   # every map entry (LB = L)
                     "traffic_GBps": (traffic / kavg / 1e9) if traffic else None,
                     "algorithmic_TFLOPs": achieved,
                     "algorithmic_frac_of_bf16_peak": achieved / PEAK["f32x2"]})
        roof["mfma_executed_flops_per_launch"] = mf
        roof["mfma_executed_TFLOPs"] = mf / kavg / 1e12
        roof["mfma_executed_frac_of_peak"] = mf / kavg / 1e12 / PEAK["f32x2"]
        roof["group_matrices"] = ngm
    line = {
        "metric": "H·psi/s 10D Smolyak L=8 block-16: split-FP32 MFMA vs FP64, @1e-8 rel eigen err",
        "value": nvec * args.steps / dt,
        "unit": "H·psi/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.prec],
        "data": "synthetic",
        "config": {"workload": f"{D}D Henon-Heiles SG4 Smolyak L={L} block-{nvec} H.psi",
                   "D": D, "L": L, "nvec": nvec, "nterms": b.nterms, "nb_ba": b.nb_ba,
                   "precision": args.prec, "items": st["items"],
                   "launch": op.config(nvec, prec)},
        "roofline": roof,
        "reference_algorithm": {"flops_per_apply": F1,
                                "effective_TFLOPs": F1 / (dt / args.steps) / 1e12},
    }
    if graph is not None:
        line["graph_replay"] = graph
    if other is not None:
        line["other_precision"] = other
        line[f"speedup_vs_{oname}"] = (nvec * args.steps / dt) / other["H·psi/s"]
    if args.nlev and args.prec != "fp64":
        from elvibrot_amd.davidson import sub_propagation_Davidson
        runs = {}
        legs = (("fp64", PREC["fp64"], True), (f"{args.prec}_then_fp64", prec, True),
                (f"{args.prec}_only", prec, False))
        for name, p, refine in legs:     # untimed: first-use set-up of every block width
            sub_propagation_Davidson(op, args.nlev, max_it=400, conv_resi=1e-6, conv_ene=1e-10,
                                     prec=p, refine=refine)
        for name, p, refine in legs:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = sub_propagation_Davidson(op, args.nlev, max_it=400, conv_resi=1e-6,
                                         conv_ene=1e-10, prec=p, refine=refine)
            torch.cuda.synchronize()
            runs[name] = {"seconds": time.perf_counter() - t0, "iterations": r.it,
                          "n_hpsi": r.n_hpsi, "n_hpsi_reduced": getattr(r, "n_hpsi_low", 0),
                          "E": np.array(r.E)}
        e64 = runs["fp64"]["E"]
        eig = {"levels": args.nlev, "E_fp64": [float(e) for e in e64]}
        for name in runs:
            if name == "fp64":
                eig["fp64"] = {k: v for k, v in runs[name].items() if k != "E"}
                continue
            d = {k: v for k, v in runs[name].items() if k != "E"}
            d["max_rel_eigen_err_vs_fp64"] = float(np.max(np.abs(runs[name]["E"] - e64) / np.abs(e64)))
            eig[name] = d
        eig["time_to_solution_speedup"] = (runs["fp64"]["seconds"] /
                                           runs[f"{args.prec}_then_fp64"]["seconds"])
        line["eigen_check"] = eig
    if args.host_entry:
        hp = np.ascontiguousarray(X.cpu().numpy())
        op.apply_host(hp, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            op.apply_host(hp, prec)
        line["host_entry_H·psi/s"] = nvec * 5 / (time.perf_counter() - t0)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(D, L, nvec, args.cpu_stride)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
