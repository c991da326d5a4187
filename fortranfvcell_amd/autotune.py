"""Project autotuner: picks the launch shape (threads per block, row chunks =
blocks, grid-barrier kind) and the storage format (ELLPACK or CSRD) for a
system by timing short, iteration-capped solves of that system on the GPU,
and remembers the winner per (device, method, format family, size bucket).

The kernels process one row per lane (64 rows per wavefront); "rows per
block" is rows-per-chunk = neq / blocks, which is what the launch shape
changes.  A cached choice is reused for every system in the same bucket.

Cache file: $FVC_TUNE_CACHE, default ~/.cache/fvcell_tune.json.
"""
from __future__ import annotations

import json
import math
import os

from . import GMRES, DeviceSystem, device_info

CACHE_PATH = os.environ.get("FVC_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache",
                                                            "fvcell_tune.json"))

BLOCK_SIZES = (256, 512, 1024)
BARRIERS = (0, 1)


def _bucket(neq: int) -> int:
    return int(round(math.log2(max(neq, 1))))


def _key(method: int, fmt_family: str, neq: int) -> str:
    return f"{device_info()['name']}|m{method}|{fmt_family}|2^{_bucket(neq)}"


def load_cache() -> dict:
    try:
        with open(CACHE_PATH) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict) -> None:
    os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
    tmp = CACHE_PATH + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, CACHE_PATH)


def candidates(cus: int, neq: int):
    """Launch shapes worth timing: blocks from half the CUs to four per CU,
    never fewer than 64 rows per block."""
    out = []
    for bs in BLOCK_SIZES:
        for per_cu in (0.5, 1, 2, 4):
            G = int(cus * per_cu)
            if G < 8 or neq // G < 64:
                continue
            for bar in BARRIERS:
                out.append(dict(block_size=bs, grid_blocks=G, barrier=bar))
    return out


def time_config(ds: DeviceSystem, method: int, cfg: dict, iters: int = 60, reps: int = 2) -> float:
    """Seconds per iteration of an iteration-capped solve (maxit = iters)."""
    best = float("inf")
    for _ in range(reps):
        st = ds.solve(method, tol=1e-30, maxit=iters, x0_rule=2, gmres_k=30, **cfg)
        it = st["inner_iterations"] if method == GMRES else st["iterations"]
        if st["status"] not in (0, 1) or it <= 0:
            return float("inf")
        best = min(best, st["seconds"] / it)
    return best


def tune(fmts: dict, method: int, iters: int = 60, use_cache: bool = True, verbose: bool = False) -> dict:
    """Best configuration for solving this system with `method`.

    fmts: layouts of the same system, e.g. {"ell": to_ell(s), "csrd": to_csrd(s)}.
    Returns {"layout": name, "options": {...}, "s_per_iter": t}.
    """
    first = next(iter(fmts.values()))
    neq = int(first["neq"])
    key = _key(method, "+".join(sorted(fmts)), neq)
    cache = load_cache() if use_cache else {}
    if key in cache and cache[key].get("layout") in fmts:
        return cache[key]
    cus = device_info()["compute_units"]
    best = None
    for name, f in fmts.items():
        ds = DeviceSystem(f)
        for cfg in candidates(cus, neq):
            t = time_config(ds, method, cfg, iters)
            if verbose:
                print(f"tune {name} {cfg} {t * 1e6:.1f} us/it", flush=True)
            if best is None or t < best["s_per_iter"]:
                best = {"layout": name, "options": dict(cfg), "s_per_iter": t}
    if use_cache and best is not None:
        cache[key] = best
        save_cache(cache)
    return best


def tune_batch(batch: dict, precision: int = 0, verbose: bool = False) -> dict:
    """Best (threads per block, block-allocation exponent) for a batch, by
    timing whole batched solves (each system runs to its tolerance) of the
    batch resident in HBM."""
    from . import DeviceBatch
    db = DeviceBatch(batch)
    best = None
    for bs in BLOCK_SIZES:
        for exp10 in (10, 15, 20):
            sts = db.solve(tol=1e-10, precision=precision, stop_rule=1, block_size=bs, batch_work_exp10=exp10)
            t = sts[0]["seconds"]
            if verbose:
                print(f"tune batch bs={bs} exp={exp10 / 10} {t * 1e3:.2f} ms", flush=True)
            if best is None or t < best["seconds"]:
                best = {"options": dict(block_size=bs, batch_work_exp10=exp10), "seconds": t}
    return best
