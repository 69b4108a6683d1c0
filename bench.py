#!/usr/bin/env python3
"""Benchmark of the batched split-step Pade range march on one MI355X.

Workload (default): the 10,000-ray Munk batch (nz = 5303 rows, np = 6 Pade
terms, dr = 10 m, dz = 1 m, 100 Hz; cupyram_amd.workloads.munk_batch).  A
bench "step" is one range step of the march over the whole batch.  The self-
starter and ``--warmup`` steps run untimed; exactly ``--steps`` march steps are
timed between device synchronisations, with inputs resident in HBM.

value = sum over rays of nz x timed steps / wall time / 1e6 (Mcell-steps/s)
for the complex64 march; the complex128 march of the same batch is reported
beside it.  The march kernel keeps each column in registers, so it is bound
by VALU issue, not HBM: roofline.achieved is its FP32 arithmetic, 140 np
flop per cell-step, over the kernel's mean launch time measured with
hipEvents on the plan's stream, against the FP32 vector peak; the HBM
figures (the minimum traffic of 36 B per cell-step, the profiled traffic and
the reference design's 408 B) are reported beside it.  cpu_baseline times the
FP64 CPU oracle (the PyRAM restatement, built -O3 -march=native on this host)
on a bounded sample of the same batch on every host core the job may use.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-steps/s (nz×range steps) c64 vs c128 at ≤0.5 dB TL err; HBM GB/s"


# HBM bytes per cell-step of the complex64 team march on the 10,000-ray Munk
# batch, from rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a 1 GiB copy
# (factors 2.0 and 1.0), and the VALU-busy fraction from the SQ counters of
# the same kernel (4 x SQ_ACTIVE_INST_VALU / SIMDs / busy cycles); keyed by
# (precision, team rows, team lanes).  These are profile figures, reported
# beside the live timing, not measured by this script.
PROFILED = {(64, 21, 256): {"fetch": 0.526, "write": 0.224, "valu_busy": 0.866,
                            "source": "profiles/team_march_munk10k.md"}}

FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak
FP32_SINGLE_ISSUE_TFLOPS = 78.6   # 256 CUs x 64 lanes x 2 flop (FMA) x 2.4 GHz, no packed FP32
FP64_PEAK_TFLOPS = 78.6           # MI355X published FP64 vector peak


def flop_per_cell_step(np_pade: int) -> int:
    """Arithmetic of one cell-step: about 140 flop per Pade term (assembly,
    right-hand side, elimination and back substitution of one row)."""
    return 140 * np_pade


def min_bytes_per_cell_step(precision: int) -> int:
    """Lower bound of the HBM traffic: the field in and out and the medium
    (f1, f2, f3, ksq) read once per step."""
    s = 8 if precision == 64 else 16
    r = 4 if precision == 64 else 8
    return 2 * s + 3 * r + s


def alg_bytes_per_cell_step(np_pade: int, precision: int) -> int:
    """Traffic of the reference's design (Thomas workspace and medium in HBM,
    per Pade term); this build keeps them on chip, so the figure is reported
    for comparison only."""
    s = 8 if precision == 64 else 16     # complex element
    r = 4 if precision == 64 else 8      # real medium element
    return np_pade * (6 * s + 3 * r + s)


class MarchTimer:
    """One plan of a single-grid workload: self-starter and `warmup` steps
    untimed, then `steps` march steps timed between device synchronisations
    (wall clock, and the kernels' hipEvent time on the plan's stream)."""

    def __init__(self, model):
        model.setup()
        buckets = model._buckets()
        assert len(buckets) == 1, "the timed workload has one grid"
        idx, shape, variant = buckets[0]
        self.model = model
        self.plan = model._make_plan(idx, shape=shape, variant=variant)

    def warm(self, warmup):
        self.warmup = warmup
        self.plan.selfs()
        if warmup:
            self.plan.march(0, warmup)

    def timed(self, steps, launch_steps):
        plan, model, warmup = self.plan, self.model, self.warmup
        try:
            plan.sync()
            ms0, n0 = plan.march_time()
            t0 = time.perf_counter()
            s = warmup
            while s < warmup + steps:
                e = min(s + launch_steps, warmup + steps)
                plan.march(s, e)
                s = e
            plan.sync()
            wall = time.perf_counter() - t0
            ms1, n1 = plan.march_time()
            variant = plan.describe(0)
            tll = plan.read_outputs()[0]
        finally:
            plan.close()
        # every ray of the timed plan has stored the same number of outputs (an
        # output never written holds p = 0, so TL >= 400 dB; a team that marched
        # another ray's column would leave its own outputs unwritten)
        written = np.sum(np.isfinite(tll) & (tll < 399.0), axis=1)
        if written.min() != written.max() or written.max() == 0:
            raise RuntimeError(f"timed march left rays unmarched: outputs per ray {written.min()}..{written.max()}")
        cells = float(np.sum(model._rays["nz"].astype(np.float64)))
        return {"wall_s": wall, "kernel_ms": ms1 - ms0, "launches": n1 - n0, "cells": cells,
                "mcs": cells * steps / wall / 1e6, "variant": variant, "outputs_per_ray": int(written.max())}


def host_cpu():
    """The host cores this process may use and how they are described: the
    affinity mask, the cgroup CPU quota and OMP_NUM_THREADS (the share the
    box allots to one GPU's job) bound the thread count."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = aff
    if quota is not None:
        threads = min(threads, max(1, int(quota)))
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": threads, "nproc": os.cpu_count(), "affinity_cpus": aff, "cgroup_cpu_quota": quota,
            "omp_num_threads": omp, "model": model,
            "pinning": "none (OpenMP threads unpinned, dynamic schedule over rays)"}


def _oracle_baseline(m, npick, what):
    """Time the FP64 oracle built with -O3 -march=native on this host over
    the rays of a set-up CuPyRAM instance m, one ray per OpenMP thread at a
    time; returns the cpu_baseline record and the TL lines."""
    import tempfile
    from oracle import ram_oracle as O
    from tests import envs as E
    host = host_cpu()
    rays = [E.oracle_for(m, e, 0) for e in range(npick)]
    native = O.native_lib(tempfile.mkdtemp(prefix="cpr_oracle_"))
    t0 = time.perf_counter()
    tl_ref = O.run_batch_tl(rays, nthreads=host["threads"], library=native)
    t = time.perf_counter() - t0
    cell_steps = float(np.sum(m._rays["nz"].astype(np.float64) * m._rays["nsteps"]))
    cb = {"value": cell_steps / t / 1e6, "unit": "Mcell-steps/s", "cores": host["threads"], "kind": "port",
          "per_core": cell_steps / t / 1e6 / host["threads"],
          "sample": f"{what}; FP64 oracle (ram_oracle.c, -O3 -march=native -fno-tree-vectorize, built on this host), OpenMP over rays, "
                    f"{npick / host['threads']:.1f} rays per thread ({cell_steps:.3g} cell-steps in {t:.1f} s)",
          "host": host}
    return cb, tl_ref


def _tl_errors(skw, tl_ref, npick, precision_runs):
    from cupyram_amd import CuPyRAM
    errs = {}
    for prec, variant in precision_runs.items():
        g = CuPyRAM(**skw, precision=prec, variant=variant).run()["TL Line"]
        d = np.concatenate([g[i, :len(tl_ref[i])] - tl_ref[i] for i in range(npick)])
        d = d[np.isfinite(d)]
        errs[prec] = {"rms_db": float(np.sqrt(np.mean(d ** 2))), "max_db": float(np.max(np.abs(d))),
                      "frac_gt_0.5db": float(np.mean(np.abs(d) > 0.5))}
    return errs


def cpu_baseline(kw, sample_rays, sample_rmax, precision_runs):
    """Oracle (FP64 PyRAM restatement) on a seeded sample of the batch with a
    shortened range, on every host core this process may use; also the TL
    error of the GPU marches on the same sample."""
    from cupyram_amd import CuPyRAM
    rng = np.random.default_rng(7)
    n = kw["batch_size"]
    threads = host_cpu()["threads"]
    npick = min(max(sample_rays, 4 * threads), n)
    pick = np.sort(rng.choice(n, size=npick, replace=False))
    skw = dict(kw)
    for k in ("cb", "rhob", "attn", "rbzb"):
        if isinstance(kw[k], list):
            skw[k] = [kw[k][i] for i in pick]
    skw["batch_size"] = len(pick)
    skw["rmax"] = sample_rmax
    m = CuPyRAM(**skw)
    m.setup()
    cb, tl_ref = _oracle_baseline(m, npick, f"{npick} seeded rays of the batch, rmax {sample_rmax:.0f} m "
                                            f"({int(m._rays['nsteps'][0])} steps), full nz")
    return cb, _tl_errors(skw, tl_ref, npick, precision_runs)


def uneven_cpu_baseline(kw, nz, per_decile=26, sample_rmax=5000.0, seed=11):
    """Oracle on a sample of the uneven batch stratified by column depth
    (per_decile seeded rays from each tenth of the rays sorted by nz), with
    the range cut to sample_rmax so that it finishes in seconds, and the TL
    error of both GPU marches on the same rays."""
    from cupyram_amd import CuPyRAM
    rng = np.random.default_rng(seed)
    order = np.argsort(nz, kind="stable")
    pick = np.sort(np.concatenate([rng.choice(part, size=min(per_decile, len(part)), replace=False)
                                   for part in np.array_split(order, 10)]))
    skw = dict(kw)
    for k in ("z_ss", "cw", "cb", "rhob", "attn", "rbzb"):
        skw[k] = [kw[k][i] for i in pick]
    for k in ("freq", "dz", "dr"):
        skw[k] = np.asarray(kw[k])[pick]
    skw["batch_size"] = len(pick)
    skw["rmax"] = sample_rmax
    m = CuPyRAM(**skw)
    m.setup()
    cb, tl_ref = _oracle_baseline(m, len(pick), f"{len(pick)} seeded rays stratified by nz "
                                                f"({int(m._rays['nz'].min())}-{int(m._rays['nz'].max())} rows), "
                                                f"rmax {sample_rmax:.0f} m")
    return cb, _tl_errors(skw, tl_ref, len(pick), {"c64": "auto", "c128": "auto"})


def bench_uneven(args):
    """The uneven batch (per-ray frequency and grid): the whole job after the
    self-starter, every nz-bucket on its own stream, march launches of
    --launch-steps range steps captured in one hipGraph per bucket."""
    from cupyram_amd import CuPyRAM, workloads
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    kw = workloads.uneven_batch(nrays=args.rays)
    log(f"uneven: generated {args.rays} rays")
    out = {}
    for prec in (("c64", "c128") if args.with_c128 else ("c64",)):
        m = CuPyRAM(**kw, precision=prec)
        m.setup()
        buckets = m._buckets()
        log(f"uneven {prec}: setup done, {len(buckets)} buckets")
        plans = []
        try:
            for idx, shape, variant in buckets:
                plans.append((idx, m._make_plan(idx, shape=shape, variant=variant)))
            for _, p in plans:
                p.selfs()
            for _, p in plans:
                p.sync()
            log(f"uneven {prec}: self-starters done, marching")
            t0 = time.perf_counter()
            for _, p in plans:
                p.march_graph(args.launch_steps)
            for i, (_, p) in enumerate(plans):
                p.sync()
                log(f"uneven {prec}: bucket {i} done at {time.perf_counter() - t0:.1f} s")
            wall = time.perf_counter() - t0
            per_bucket = []
            for idx, p in plans:
                # the whole job: every ray has stored all its outputs
                tll = p.read_outputs()[0]
                nvr = m._rays["nvr"][idx]
                written = np.sum(np.isfinite(tll) & (tll < 399.0), axis=1)
                if np.any(written < nvr):
                    raise RuntimeError(f"uneven {prec}: {int(np.sum(written < nvr))} rays left unmarched")
                ms, n = p.march_time()
                v, r, l = p.describe(0)
                cs = float(np.sum(m._rays["nz"][idx].astype(np.float64) * m._rays["nsteps"][idx]))
                per_bucket.append({"rays": len(idx), "variant": v, "rows": r, "lanes": l,
                                   "nz_max": int(m._rays["nz"][idx].max()), "kernel_s": ms / 1e3,
                                   "cell_steps": cs, "mcs": cs / (ms / 1e3) / 1e6 if ms else None})
        finally:
            for _, p in plans:
                p.close()
        cells = float(np.sum(m._rays["nz"].astype(np.float64) * m._rays["nsteps"]))
        out[prec] = {"value": cells / wall / 1e6, "wall_s": wall, "cell_steps": cells, "buckets": per_bucket}
        nz_all = m._rays["nz"].copy()
    line = {"metric": METRIC, "value": out["c64"]["value"], "unit": "Mcell-steps/s", "n_gpus": 1,
            "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (complex64 march; FP64 Pade and self-starter)",
            "data": "synthetic (seeded uneven batch)",
            "config": {"workload": f"uneven_batch: {args.rays} rays, depth logU 100-5000 m, freq logU 25-500 Hz, "
                                   "dz=lambda/10, dr=lambda/2, rmax 20 km, np 6 (whole job)"},
            "c64": out["c64"], "c128": out.get("c128")}
    if "c128" in out:
        line["speedup_c64_vs_c128"] = out["c64"]["value"] / out["c128"]["value"]
    if not args.no_cpu:
        log("uneven: CPU oracle on a stratified sample")
        cb, errs = uneven_cpu_baseline(kw, nz_all)
        line["cpu_baseline"] = cb
        line["tl_err_vs_fp64_oracle"] = errs
        line["speedup_vs_cpu"] = line["value"] / cb["value"]
    print(json.dumps(line), flush=True)
    return 0


ROUTER_WORKLOADS = {
    # domain, default query count, query seed, workload text
    "wedge": ("wedge", 100000, 5, "wedge_queries: {n} queries, 200->20 m over 10 km, sediment 5-30 m over "
                                  "basement, freq logU 50-1000 Hz, np 6, output 200 x 64 cell-mean TL"),
    "deep": ("deep_munk", 20000, 7, "deep_queries: {n} queries, Munk 5000 m perturbed at 32 knots, 16-knot "
                                   "bathymetry 3500-5000 m over 50 km, freq logU 80-125 Hz, np 6, output 200 x 64 "
                                   "cell-mean TL"),
}


def bench_router(args):
    """Surrogate-first answering of a query set with complex64 PE fallback,
    beside the plain complex64 PE march of every query; TL error of both
    against the complex128 march on the domain grid (cell-mean TL clipped to
    20-140 dB)."""
    import torch
    from cupyram_amd.surrogate import dataset as D
    from cupyram_amd.surrogate.router import Router
    from cupyram_amd.surrogate.train import field_rms, load_ensemble, pe_fields, summary
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    dname, n_default, qseed, text = ROUTER_WORKLOADS[args.workload]
    dom = D.DOMAINS[dname]
    n = args.rays if args.rays != 10000 else n_default
    gen = D.deep_queries if dname == "deep_munk" else D.wedge_queries
    q = gen(n, qseed)
    members = load_ensemble(dom.name)
    log(f"{dname}: {n} queries, {len(members)} ensemble members")
    router = Router(dom, members, threshold_db=3.0, spot_fraction=0.01, seed=qseed, precision="c64", log=None)
    router.ensemble(q.take(np.arange(min(n, 256))))          # warm-up (buffers, kernels)
    torch.cuda.synchronize()
    res = router.run(q)
    log(f"{dname}: routed {res.frac_routed:.3f} in {res.seconds['total']:.1f} s")
    t0 = time.perf_counter()
    tl64, t64, cells = pe_fields(dom, q, "c64", chunk=20000, log=log)
    torch.cuda.synchronize()
    wall64 = time.perf_counter() - t0
    tl128, t128, _ = pe_fields(dom, q, "c128", chunk=20000, log=log)
    e2e = summary(field_rms(dom, res.tl, tl128))
    pe_err = summary(field_rms(dom, tl64, tl128))
    acc = ~res.routed
    sur = summary(field_rms(dom, res.tl[torch.as_tensor(np.nonzero(acc)[0], device="cuda")],
                            tl128[torch.as_tensor(np.nonzero(acc)[0], device="cuda")])) if acc.any() else None
    line = {"metric": f"{dname} queries/s, surrogate-first with complex64 PE fallback (TL error vs complex128)",
            "value": n / res.seconds["total"], "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
            "data": f"synthetic (seeded {dname} queries, seed {qseed})",
            "config": {"workload": text.format(n=n),
                       "members": len(members), "threshold_db": 3.0, "spot_fraction": 0.01},
            "frac_routed_to_pe": res.frac_routed, "batch_fallback": res.batch_fallback,
            "seconds": res.seconds, "ensemble_spread_median_db": float(np.nanmedian(res.spread_db)),
            "spot_checked": int(len(res.spot)),
            "spot_err_rms_db": float(np.sqrt(np.mean(res.spot_err_db ** 2))) if len(res.spot) else None,
            "tl_err_end_to_end_vs_c128": e2e, "tl_err_accepted_surrogate_vs_c128": sur,
            "pe_only_c64": {"queries_per_s": n / wall64, "march_s": t64, "wall_s": wall64, "cell_steps": cells,
                            "tl_err_vs_c128": pe_err},
            "c128_march_s": t128}
    print(json.dumps(line), flush=True)
    return 0


def fno_flops(cfg, S):
    """Multiply-add flops (x2) of one forward pass over S samples: products on
    MFMA, depth transforms on the vector ALUs."""
    R = S * cfg.depth_points
    C, K, F = cfg.width, cfg.modes, cfg.features
    mfma = 2 * R * F * C + cfg.layers * (2 * K * S * (2 * C) ** 2 + 2 * R * C * C) + 2 * R * C * C + 2 * R * C
    valu = cfg.layers * 2 * (2 * S * cfg.depth_points * C * 2 * K)
    return mfma, valu


def bench_fno(args):
    """Forward (inference) and training-step throughput of the surrogate
    network on synthetic bf16 inputs (64 queries x 200 range bins); `value`
    is the router's inference form (first layer from the encoding's tables of
    64 seeded deep-water queries), `inference_fused` the same kernels on an
    encoded input."""
    import torch
    from cupyram_amd.surrogate import FNO, FNOConfig
    cfg = FNOConfig()
    nq = 64
    S = nq * 200
    g = torch.Generator().manual_seed(0)
    x = (torch.randn(S, cfg.depth_points, cfg.features, generator=g) * 0.5).to(torch.bfloat16).cuda()
    tgt = torch.randn(S * cfg.depth_points, generator=g).cuda()
    mask = torch.ones_like(tgt)
    fno = FNO(cfg, seed=0)
    mf, vf = fno_flops(cfg, S)
    # the router's form: the first layer from the encoding's tables of 64
    # seeded deep-water queries (no encoded input, no first-layer product)
    from cupyram_amd.surrogate import dataset as D
    from cupyram_amd.surrogate.encode import lift_inputs
    inp = lift_inputs(D.DEEP, D.deep_queries(nq, 0))

    def fwd():
        fno.forward(x, train=False)

    def fused():
        fno.forward_fused(x)

    def fused_lift():
        fno.forward_fused_lift(inp)

    def train():
        pred = fno.forward(x, train=True)
        _, dpred = fno.loss_grad(pred, tgt, mask)
        fno.backward(dpred)
        fno.step(1e-4)

    out = {}
    for name, fn, n in (("inference_fused_lift", fused_lift, args.steps), ("inference_fused", fused, args.steps),
                        ("inference", fwd, args.steps), ("train_step", train, max(1, args.steps // 3))):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / n
        mult = 3 if name == "train_step" else 1
        # the table form does neither the product with the encoded input nor
        # the second layer's pointwise product (both from the first layer's tables)
        R = S * cfg.depth_points
        mfn = mf - (2 * R * cfg.width * (cfg.features + cfg.width) if name == "inference_fused_lift" else 0)
        out[name] = {"ms": dt * 1e3, "fields_per_s": nq / dt, "mfma_tflops": mult * mfn / dt / 1e12,
                     "valu_tflops": mult * vf / dt / 1e12}
    line = {"metric": "surrogate fields/s (200 x 64 TL field per query)",
            "value": out["inference_fused_lift"]["fields_per_s"],
            "unit": "fields/s", "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 (FP32 accumulation)",
            "data": "synthetic inputs", "steps": args.steps,
            "config": {"workload": "FNO 4 layers, width 128, 16 modes over 64 depths, 200 range bins, batch 64",
                       "params": fno.n_params(), "forward_gflop_per_field": (mf + vf) / nq / 1e9},
            **out}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--rays", type=int, default=10000)
    ap.add_argument("--variant", default="auto")
    ap.add_argument("--launch-steps", type=int, default=100, help="range steps per kernel launch")
    ap.add_argument("--no-c128", action="store_true")
    ap.add_argument("--with-c128", action="store_true", help="uneven workload: also march in complex128")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=256, help="CPU baseline sample (at least 4 rays per thread)")
    ap.add_argument("--cpu-rmax", type=float, default=20000.0)
    ap.add_argument("--workload", default="munk", choices=["munk", "uneven", "wedge", "deep", "fno"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (torch) so that HBM counters can be scaled")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(rays are independent: shard the ray list across processes)"}))
        return 2
    if args.copy_calibration:
        import torch
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload in ROUTER_WORKLOADS:
        return bench_router(args)
    if args.workload == "fno":
        return bench_fno(args)
    if args.workload == "uneven":
        # one stream per nz-bucket (20 here): HIP maps streams onto
        # GPU_MAX_HW_QUEUES hardware queues (4 unless set), and kernels of
        # streams sharing a queue run one after another; set before the
        # first HIP call (with 24 queues the batch measured 4 % faster)
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "24")
        if args.rays == 10000:
            args.rays = 50000
        if args.launch_steps == 100:
            args.launch_steps = 1000
        return bench_uneven(args)
    from cupyram_amd import CuPyRAM, workloads
    kw = workloads.munk_batch(nrays=args.rays)
    np_pade = kw["np"]
    # Both precisions are set up and warmed before either is timed, so that
    # the complex64 timing does not start on an idle (down-clocked) GPU:
    # after 50-200 ms idle a 20-step launch measured 21.1 ms instead of 18.9.
    t64 = MarchTimer(CuPyRAM(**kw, precision="c64", variant=args.variant))
    t128 = None if args.no_c128 else MarchTimer(CuPyRAM(**kw, precision="c128", variant=args.variant))
    # The plans run on streams of their own and the self-starter is
    # asynchronous: the complex128 warm-up completes before the complex64 one
    # starts, so that the complex64 timing follows its own warm-up steps and
    # does not share the GPU with other work.
    if t128:
        t128.warm(args.warmup)
        t128.plan.sync()
    t64.warm(args.warmup)
    t64.plan.sync()
    r64 = t64.timed(args.steps, args.launch_steps)
    out_c128 = None
    if t128:
        r128 = t128.timed(args.steps, args.launch_steps)
        l128 = r128["kernel_ms"] / max(r128["launches"], 1) / 1e3
        a128 = flop_per_cell_step(np_pade) * r128["cells"] * args.steps / max(r128["launches"], 1) / l128 / 1e12
        out_c128 = {"value": r128["mcs"], "ms_per_step": 1e3 * r128["wall_s"] / args.steps,
                    "variant": r128["variant"],
                    "kernel_ms_per_launch": l128 * 1e3,
                    "roofline": {"bound": "valu", "achieved": a128, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": a128 / FP64_PEAK_TFLOPS}}
    per_launch_steps = args.steps / max(r64["launches"], 1)
    cell_steps_per_launch = r64["cells"] * per_launch_steps
    avg_launch_s = r64["kernel_ms"] / max(r64["launches"], 1) / 1e3
    flop = flop_per_cell_step(np_pade)
    achieved = flop * cell_steps_per_launch / avg_launch_s / 1e12
    b_min = min_bytes_per_cell_step(64)
    prof = PROFILED.get((64, r64["variant"][1], r64["variant"][2])) if r64["variant"][0] == 2 else None
    traffic = (prof["fetch"] + prof["write"]) * cell_steps_per_launch if prof else None
    line = {
        "metric": METRIC, "value": r64["mcs"], "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * r64["wall_s"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (complex64 march; FP64 Pade and self-starter)",
        "data": "synthetic (seeded Munk batch, random bathymetry and bottom)",
        "config": {"workload": "munk_batch: 10,000-ray Munk 5000 m, 100 Hz, nz 5303, np 6, dr 10 m, dz 1 m",
                   "rays": args.rays, "nz": int(r64["cells"] / args.rays), "np": np_pade,
                   "variant": r64["variant"], "steps_per_launch": per_launch_steps},
        # The team march keeps each ray's column in registers and its
        # coefficients in LDS for a whole launch, so HBM carries little more
        # than the field once per launch; its bound is VALU issue.  achieved =
        # 140 np FP32 flop per cell-step (the march's arithmetic) x the
        # launch's cell-steps / its hipEvent-timed duration, against the
        # published FP32 vector peak.
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic,
                     "flop_per_cell_step": flop,
                     "kernel_ms_per_launch": avg_launch_s * 1e3,
                     "frac_of_single_issue_fp32_peak": achieved / FP32_SINGLE_ISSUE_TFLOPS,
                     "traffic_source": (f"{prof['source']}: rocprofv3 FETCH_SIZE + WRITE_SIZE per cell-step of this "
                                        "kernel and shape, scaled by a copy kernel, times this launch's cell-steps")
                     if prof else None,
                     "hbm": {"b_min_per_cell_step": b_min,
                             "b_min_gbs": b_min * cell_steps_per_launch / avg_launch_s / 1e9,
                             "b_min_frac": b_min * cell_steps_per_launch / avg_launch_s / 1e9 / HBM_PEAK_GBS,
                             "profiled_bytes_per_cell_step": (prof["fetch"] + prof["write"]) if prof else None,
                             "profiled_gbs": traffic / avg_launch_s / 1e9 if traffic else None,
                             "alg_bytes_reference_design_per_cell_step": alg_bytes_per_cell_step(np_pade, 64),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s"},
                     "valu_busy_profiled": prof["valu_busy"] if prof else None},
        "c128": out_c128,
    }
    if out_c128:
        line["speedup_c64_vs_c128"] = r64["mcs"] / out_c128["value"]
    if not args.no_cpu:
        cb, errs = cpu_baseline(kw, args.cpu_rays, args.cpu_rmax,
                                {"c64": args.variant, "c128": args.variant})
        line["cpu_baseline"] = cb
        line["tl_err_vs_fp64_oracle"] = errs
        line["speedup_vs_cpu"] = r64["mcs"] / cb["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
