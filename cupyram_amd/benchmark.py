"""Batch-scaling benchmark of the march on the local GPU.

    python -m cupyram_amd.benchmark [-f 100 [-f 125 ...]] [--rmax 2000]
        [--max-workers 8] [--n-steps 9] [--source-depth 10] [--receiver-depth 10]
        [--precision c128|c64]

The counterpart of the reference's ``cupyram-benchmark`` command
(cupyram/benchmark.py:282-518), with the same options, synthetic environment
and report: it sizes a series of batches from the free device memory, marches
each one through ``CuPyRAM.run()`` and prints, per batch, the march time on the
device (hipEvents around the march launches, the reference's
'propagation' time), the whole run() time, both throughputs in fields/s and
the overhead.  Differences: the memory per ray is measured (the device bytes
a plan of the same grid holds) instead of estimated, the timings are
wall-clock (the reference's 'GPU time' is CPU process time), and
``--precision`` selects the march precision.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

MAX_BATCH = 50000


def synthetic_environment(rmax: float) -> dict:
    """Isovelocity 1500 m/s water 3000 m deep over a 1700 m/s, 1.5 g/cm3,
    0.5 dB/lambda bottom, tabulated every km out to rmax (the benchmark
    environment of cupyram/benchmark.py:154-194)."""
    r = np.arange(0.0, rmax + 1.0, 1000.0)
    z_ss = np.array([0.0, 100.0, 500.0, 1000.0, 2000.0, 3000.0])
    n = len(r)
    return dict(z_ss=z_ss, rp_ss=r, cw=np.full((len(z_ss), n), 1500.0), z_sb=np.array([0.0]), rp_sb=r,
                cb=np.full((1, n), 1700.0), rhob=np.full((1, n), 1.5), attn=np.full((1, n), 0.5),
                rbzb=np.column_stack([r, np.full(n, 3000.0)]), c0=1500.0)


def batch_sizes(bytes_per_ray: float, free_bytes: int, n_freqs: int, n_steps: int) -> list:
    """Batches from 40 % to 120 % of what 90 % of the free memory holds (at
    most MAX_BATCH), rounded to multiples of 128; a linear series from 32
    when fewer than 512 rays fit (the series of cupyram/benchmark.py:86-151)."""
    fit = int(free_bytes * 0.9 // max(bytes_per_ray * n_freqs, 1.0))
    top = max(min(fit, MAX_BATCH), 128)
    if top < 512:
        return list(range(32, top + 1, max(32, top // 8)))
    sizes = np.linspace(int(top * 0.4), int(top * 1.2), n_steps, dtype=int)
    out = []
    for b in sizes:
        b = int(round(b / 128) * 128)
        if b not in out:
            out.append(b)
    return out


def _model(batch, freqs, rmax, zs, zr, max_workers, precision):
    from .cupyram import CuPyRAM
    env = synthetic_environment(rmax)
    c0 = env.pop("c0")
    lists = {k: [v] * batch for k, v in env.items()}
    return CuPyRAM(freqs if len(freqs) > 1 else freqs[0], zs, zr, lists["z_ss"], lists["rp_ss"], lists["cw"],
                   lists["z_sb"], lists["rp_sb"], lists["cb"], lists["rhob"], lists["attn"], lists["rbzb"],
                   rmax=rmax, zmplt=100.0, compute_grids=False, batch_size=batch, max_workers=max_workers,
                   c0=c0, benchmark_propagation=True, precision=precision)


def bytes_per_ray(freqs, rmax, zs, zr, max_workers, precision, probe=256) -> float:
    """Device bytes per calculation (ray x frequency) of a plan on this grid."""
    m = _model(probe, freqs, rmax, zs, zr, max_workers, precision)
    m.setup()
    total = 0
    for idx, shape, variant in m._buckets():
        plan = m._make_plan(idx, shape=shape, variant=variant)
        try:
            total += plan.device_bytes()
        finally:
            plan.close()
    return total / m._total_batch


def run_batch(batch, freqs, rmax, zs, zr, max_workers, precision) -> dict:
    t0 = time.perf_counter()
    m = _model(batch, freqs, rmax, zs, zr, max_workers, precision)
    m.run()
    total = time.perf_counter() - t0
    fields = batch * len(freqs)
    return {"batch_size": batch, "total_fields": fields, "time_total": total,
            "time_propagation": m.propagation_time, "dr": float(m._dr), "dz": float(m._dz)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--frequency", "-f", type=float, action="append", help="Hz; repeat for multi-frequency (100)")
    ap.add_argument("--rmax", type=float, default=2000.0)
    ap.add_argument("--max-workers", type=int, default=8)
    ap.add_argument("--n-steps", type=int, default=9)
    ap.add_argument("--source-depth", type=float, default=10.0)
    ap.add_argument("--receiver-depth", type=float, default=10.0)
    ap.add_argument("--precision", default="c128", choices=["c128", "c64"])
    args = ap.parse_args(argv)
    freqs = args.frequency or [100.0]
    from . import _abi
    if _abi.device_count() < 1:
        print("no HIP device available", file=sys.stderr)
        return 1
    dev_name, free, total = _abi.device_info(0)
    print("=" * 70)
    print("CuPyRAM batch scaling benchmark (MI355X march)")
    print("=" * 70)
    print(f"Device: {dev_name}; memory {total / 2**30:.1f} GiB, free {free / 2**30:.1f} GiB")
    bpr = bytes_per_ray(freqs, args.rmax, args.source_depth, args.receiver_depth, args.max_workers, args.precision)
    sizes = batch_sizes(bpr, free, len(freqs), args.n_steps)
    print(f"Frequencies {freqs} Hz, rmax {args.rmax / 1000:.1f} km, zs {args.source_depth} m, "
          f"zr {args.receiver_depth} m, precision {args.precision}")
    print(f"Device memory per ray: {bpr / 2**20:.2f} MiB (measured); batch sizes {sizes}")
    results = []
    for i, b in enumerate(sizes):
        print(f"[{i + 1}/{len(sizes)}] batch {b} ...", end="", flush=True)
        try:
            r = run_batch(b, freqs, args.rmax, args.source_depth, args.receiver_depth, args.max_workers,
                          args.precision)
        except Exception as exc:  # noqa: BLE001 - report and go on, as the reference does
            print(f" error: {exc}")
            if "memory" in str(exc).lower():
                break
            continue
        r["throughput_propagation"] = r["total_fields"] / max(r["time_propagation"], 1e-12)
        r["throughput_total"] = r["total_fields"] / r["time_total"]
        r["overhead_pct"] = 100.0 * (r["time_total"] - r["time_propagation"]) / r["time_total"]
        results.append(r)
        print(f" march {r['time_propagation']:.3f} s, total {r['time_total']:.2f} s "
              f"(dr={r['dr']:.2f} m, dz={r['dz']:.4f} m)", flush=True)
    if not results:
        print("no successful runs", file=sys.stderr)
        return 1
    best = max(r["throughput_propagation"] for r in results)
    print(f"\n{'Batch':>7} {'March (s)':>10} {'March tput':>12} {'Eff.':>7} {'Total (s)':>10} "
          f"{'Total tput':>12} {'Overhead':>9}")
    for r in results:
        print(f"{r['batch_size']:>7} {r['time_propagation']:>10.3f} {r['throughput_propagation']:>12.1f} "
              f"{100 * r['throughput_propagation'] / best:>6.1f}% {r['time_total']:>10.2f} "
              f"{r['throughput_total']:>12.1f} {r['overhead_pct']:>8.0f}%")
    opt = max(results, key=lambda r: r["throughput_propagation"])
    print(f"\nPeak: batch {opt['batch_size']}, {opt['throughput_propagation']:.1f} fields/s (march), "
          f"{opt['throughput_total']:.1f} fields/s (run), grid dr={opt['dr']:.2f} m, dz={opt['dz']:.4f} m "
          f"(~{int(args.rmax / opt['dr'])} steps)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
