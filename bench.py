"""Block H.psi throughput of the Smolyak SG4 operator on one MI355X.

Workload (the configuration the headline metric is quoted on): 10D
Henon-Heiles (lambda = 0.111803), Smolyak SG type 4 with L = LB = LG = 8,
HO bases nq(l) = nb(l) = 1 + 2l, block of 16 seeded orthonormal vectors.
A step is one block application H.psi (sub_TabOpPsi_FOR_SGtype4) with the
input block already resident in HBM.

Prints ONE JSON line with the metric, the roofline of the H.psi kernel
(algorithmic flops per launch over the launch time measured with HIP events on
the plan's stream) and the CPU baseline (the oracle, OpenMP over terms, on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "f32": 157.3, "f32x2": 157.3}  # TFLOP/s (MI355X dense)
PREC = {"fp64": 0, "f32": 1, "f32x2": 2}
DTYPE = {"fp64": "f64", "f32": "f32", "f32x2": "f32x2"}


def cpu_baseline(D, L, nvec, stride):
    """Oracle (OpenMP) on every `stride`-th term of the same workload; the
    rate is scaled by the sampled share of the algorithmic flops."""
    from oracle import oracle as O
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    nq = np.diff(o.qoff).astype(np.float64)
    per_term = np.array([2.0 * nq[t] * np.sum(1 + 2 * o.lev[t]) for t in range(o.nterms)])
    frac = per_term[::stride].sum() / per_term.sum()
    psi = np.random.default_rng(2026).standard_normal((o.nb_ba, nvec))
    t0 = time.perf_counter()
    o.hpsi(psi, V, -0.5 * np.ones(D), t0=0, t1=-1, tstride=stride)
    dt = time.perf_counter() - t0
    return {"value": nvec * frac / dt, "unit": "H.psi/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"every {stride}th Smolyak term ({frac:.3f} of the flops) of the same "
                      f"{D}D L={L} block-{nvec} FP64 apply, {dt:.1f} s, scaled by the flop share"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="fp64", choices=list(PREC),
                    help="precision of the reported value; the other mode is timed beside it")
    ap.add_argument("--autotune", action="store_true", help="autotune launch parameters first")
    ap.add_argument("--no-other", action="store_true", help="do not time the other precision")
    ap.add_argument("--D", type=int, default=10)
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--nvec", type=int, default=16)
    ap.add_argument("--cpu-stride", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(terms would be sharded across GPUs with one all-reduce of the packed "
              "block per apply; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import elvibrot_amd as E

    D, L, nvec = args.D, args.L, args.nvec
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    g = torch.Generator(device="cuda:0").manual_seed(2026)
    X = torch.randn(b.nb_ba, nvec, device="cuda:0", dtype=torch.float64, generator=g)
    X, _ = torch.linalg.qr(X)                       # orthonormal block
    Xc = X.T.contiguous()                            # (nvec, nb_ba) = column-major block
    Y = torch.empty_like(Xc)
    prec = PREC[args.prec]
    if args.autotune:
        op.autotune(nvec, prec)

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

    dt, kavg = timed(prec, args.steps, args.warmup)
    st = op.stats(nvec)
    achieved = st["flops"] / kavg / 1e12
    other = None
    if not args.no_other:
        op_name = "fp64" if args.prec != "fp64" else "f32"
        if args.autotune:
            op.autotune(nvec, PREC[op_name])
        dto, kavgo = timed(PREC[op_name], max(3, args.steps // 2), 2)
        other = {"precision": op_name, "H·psi/s": nvec * max(3, args.steps // 2) / dto,
                 "kernel_ms": kavgo * 1e3}
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
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK[args.prec],
                     "unit": "TFLOP/s", "frac": achieved / PEAK[args.prec], "traffic": None,
                     "kernel": "k_hpsi_t1_f64" if args.prec == "fp64" else "k_hpsi_t1_f32",
                     "kernel_ms": kavg * 1e3,
                     "flops_per_launch": st["flops"], "algorithmic_bytes": st["bytes"],
                     "hbm_GBps_algorithmic": st["bytes"] / kavg / 1e9},
    }
    if other is not None:
        line["other_precision"] = other
        line["speedup_vs_other"] = (nvec * args.steps / dt) / other["H·psi/s"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(D, L, nvec, args.cpu_stride)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
