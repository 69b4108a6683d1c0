"""Autotuner of the team march shape (rows per lane M, lanes per ray T).

For each column depth nz of a sweep and each precision, every supported M
(VGPR-resident or streamed coefficients) is timed on a synthetic batch of
about four full-GPU waves of teams (steady state, not a tail), with T = the
smallest multiple of 64 lanes that
covers nz (T <= 512).  The winner per nz is written to a JSON table
(``tuning/gfx950.json``); ``CuPyRAM`` buckets rays with it, falling back to
the library's static cost model (``cpr_team_shape``) when no table exists.

Every candidate marches the same rays from the same start field; results are
not compared here (the GPU tests cover every shape against the oracle).

Usage: ``python -m cupyram_amd.autotune [--steps 40] [--out path] [--precision 64 128]
[--nz depth ...] [--merge]`` (``--merge`` replaces only the swept depths'
entries of the table at ``--out``).
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "gfx950.json")
ROWS = {64: (4, 8, 12, 14, 16, 20, 21, 24, 28, 32, 40, 48, 64), 128: (2, 4, 6, 8, 10, 11, 12, 14, 16, 20, 24, 32, 40)}
NZ_SWEEP = (200, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 5376, 6144, 7168, 8192, 10240, 12288, 14336,
            16384, 20480)

_table_cache = None


def lanes_for(nz: int, rows: int) -> int:
    t = -(-nz // rows)
    return 64 if t <= 64 else -(-t // 64) * 64


def load_table(path: str = TABLE):
    global _table_cache
    if _table_cache is None:
        try:
            with open(path) as fh:
                _table_cache = json.load(fh)
        except (OSError, ValueError):
            _table_cache = {}
    return _table_cache


def tuned_shape(precision: int, nz: int):
    """(rows, lanes) from the tuning table, or None when the table has no
    entry covering nz."""
    rows = None
    for nz_hi, m in load_table().get(str(precision), []):
        if nz <= nz_hi:
            rows = m
            break
    if rows is None:
        return None
    lanes = lanes_for(nz, rows)
    if lanes > 512:
        return None
    if lanes == 448 and (precision == 64 or rows <= 14):
        # seven waves load the four SIMDs of a CU unevenly: the sweep's
        # complex64 rates at 20, 24, 28 and 40 rows per lane are 19-30 %
        # higher with 512 lanes than with 448, more than the 14 % of rows 512
        # lanes add; complex128 3-6 % higher at 8, 10 and 14 rows, while its
        # deeper lanes keep 448 (20 and 24 rows: within 2 %; 32 rows: 6.0
        # against 5.0 G cell-steps/s)
        lanes = 512
    return rows, lanes


def _synthetic(nz: int, nrays: int, rmax: float = 600.0):
    """1000 m Munk-like column with nz rows (dz chosen for it) over a bottom
    sloping by about two grid rows per range step, so that the seabed row
    moves every step as in the sloping-bathymetry workloads."""
    from .workloads import munk_speed
    z = np.linspace(0.0, 1000.0, 21)
    c = munk_speed(z * 5.0).reshape(-1, 1)
    freq = 100.0
    c0 = float(c.mean())
    zmax = 1000.0 + 20 * c0 / freq
    dz = zmax / (nz + 1.5)
    dr = 10.0
    drop = 2.0 * dz / dr * 1e6
    rbzb = np.array([[0.0, 1000.0], [1e6, max(1000.0 - drop, 100.0)]])
    return dict(freq=freq, zs=50.0, zr=60.0, z_ss=z, rp_ss=np.array([0.0]), cw=c,
                z_sb=np.array([0.0]), rp_sb=np.array([0.0]), cb=np.array([[1700.0]]), rhob=np.array([[1.6]]),
                attn=np.array([[0.5]]), rbzb=rbzb, rmax=rmax, dr=dr, dz=dz, np=6, batch_size=nrays,
                compute_grids=False)


def time_shape(model, shape, steps, warmup=10):
    plan = model._make_plan(np.arange(model._total_batch), shape=shape, variant="team")
    try:
        plan.selfs()
        plan.march(0, warmup)
        plan.sync()
        t0 = time.perf_counter()
        plan.march(warmup, warmup + steps)
        plan.sync()
        dt = time.perf_counter() - t0
    finally:
        plan.close()
    cells = float(np.sum(model._rays["nz"])) * steps
    return cells / dt / 1e6


def sweep(precisions=(64, 128), nz_list=NZ_SWEEP, steps=40, lanes_target=4 * 131072, log=print):
    from .cupyram import CuPyRAM
    table = {}
    for prec in precisions:
        entries = []
        for nz in nz_list:
            best = None
            for m in ROWS[prec]:
                lanes = lanes_for(nz, m)
                if lanes > 512:
                    continue
                nrays = max(64, lanes_target // lanes)
                model = CuPyRAM(**_synthetic(nz, nrays), precision=prec)
                model.setup()
                try:
                    mcs = time_shape(model, (m, lanes), steps)
                except _abi.CprError as exc:
                    log(f"  c{prec} nz={nz} M={m} T={lanes}: {exc}")
                    continue
                log(f"  c{prec} nz={nz} M={m} T={lanes} rays={nrays}: {mcs:.0f} Mcell-steps/s")
                if best is None or mcs > best[1]:
                    best = (m, mcs)
            if best is not None:
                entries.append([nz, best[0], round(best[1], 1)])
                log(f"c{prec} nz<={nz}: M={best[0]} ({best[1]:.0f} Mcell-steps/s)")
        table[str(prec)] = [[e[0], e[1]] for e in entries]
        table[f"{prec}_mcs"] = entries
    return table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--precision", type=int, nargs="*", default=[64, 128])
    ap.add_argument("--nz", type=int, nargs="*", default=list(NZ_SWEEP), help="column depths to sweep")
    ap.add_argument("--merge", action="store_true",
                    help="replace only the swept depths' entries of the table already at --out")
    args = ap.parse_args(argv)
    table = sweep(tuple(args.precision), nz_list=tuple(args.nz), steps=args.steps,
                  log=lambda s: print(s, flush=True))
    if args.merge and os.path.exists(args.out):
        with open(args.out) as fh:
            table = merge_tables(json.load(fh), table)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    print("wrote", args.out)


def merge_tables(old: dict, new: dict) -> dict:
    """`old` with the entries of `new` replacing those of the same precision
    and depth bound (entries sorted by depth)."""
    out = dict(old)
    for key, ents in new.items():
        if not key.endswith("_mcs"):
            continue
        prec = key[:-4]
        rows = {e[0]: list(e) for e in old.get(key, [])}
        rows.update({e[0]: list(e) for e in ents})
        merged = [rows[k] for k in sorted(rows)]
        out[key] = merged
        out[prec] = [[e[0], e[1]] for e in merged]
    return out


if __name__ == "__main__":
    main()
