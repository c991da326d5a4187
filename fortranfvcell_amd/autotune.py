"""Project autotuner: picks the launch shape (threads per block, row chunks =
blocks, grid-barrier kind) and the storage format (ELLPACK or CSRD) for a
system by timing short, iteration-capped solves of that system on the GPU,
and remembers the winner per (device, method, format family, size bucket);
for a batch it picks the layout (ELLPACK or CSRD), threads per block, the
block-allocation exponent and the rows per wavefront of each system's block
group by timing whole batched solves.

The kernels process one row per lane; a wavefront's rows per sweep are 64,
and its rows in all are neq / (blocks x waves per block), which is what the
launch shape (single systems) or the allocation (batches) changes.  A cached
choice is reused for every system in the same bucket.

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


ROWS_PER_WAVE = (16, 64, 256)
# batches: systems per launch as a fraction of the batch (1/3 .. 1/8), against
# the default of half the co-resident blocks
LAUNCH_DIVISORS = (3, 4, 6, 8)


def tune_batch(batch, precision: int = 0, verbose: bool = False) -> dict:
    """Best layout and launch shape for a batch, by timing whole batched
    solves (each system runs to its tolerance) of the batch resident in HBM.

    batch: one concatenated batch (concat_ell or concat_csrd), or a dict of
    layouts of the same batch, e.g. {"ell": concat_ell(fmts), "csrd":
    concat_csrd(fmts)}.  The search is staged, since each probe is a whole
    batched solve: (1) threads per block x block-allocation exponent on the
    first layout; (2) at the winning exponent, each block size with a
    third, a quarter, a sixth and an eighth of the batch per launch
    (batch_launch_systems; the launches run one after another, each on all
    blocks); (3) at the
    winner, the rows per wavefront of each system's block group
    (batch_rows_per_wave: a system gets at most neq / (rows x waves per
    block) blocks, so small systems keep their wavefronts full); (4) the
    other layouts at the winning options.  Returns {"layout": name,
    "options": {...}, "seconds": t, "trials": [...]}."""
    from . import DeviceBatch
    layouts = batch if "nsys" not in batch else {"ell" if int(batch.get("matriz", 4)) == 4 else "csrd": batch}
    names = list(layouts)
    trials = []

    def timed(name, db, opts):
        sts = db.solve(tol=1e-10, precision=precision, stop_rule=1, **opts)
        t = sts[0]["seconds"] if all(q["status"] == 0 for q in sts) else float("inf")
        trials.append({"layout": name, "options": dict(opts), "seconds": t})
        if verbose:
            print(f"tune batch {name} {opts} {t * 1e3:.2f} ms", flush=True)
        return t

    best = None
    db = DeviceBatch(layouts[names[0]])
    for bs in BLOCK_SIZES:
        for exp10 in (10, 15, 20):
            opts = dict(block_size=bs, batch_work_exp10=exp10)
            t = timed(names[0], db, opts)
            if best is None or t < best["seconds"]:
                best = {"layout": names[0], "options": opts, "seconds": t}
    # systems per launch at every block size (how many launches a batch
    # takes depends on how many blocks of that size a CU holds)
    nsys = int(layouts[names[0]]["nsys"])
    exp10 = best["options"]["batch_work_exp10"]
    for bs in BLOCK_SIZES:
        for div in LAUNCH_DIVISORS:
            opts = dict(block_size=bs, batch_work_exp10=exp10, batch_launch_systems=max(1, -(-nsys // div)))
            t = timed(names[0], db, opts)
            if t < best["seconds"]:
                best = {"layout": names[0], "options": opts, "seconds": t}
    for rpw in ROWS_PER_WAVE:
        opts = dict(best["options"], batch_rows_per_wave=rpw)
        t = timed(names[0], db, opts)
        if t < best["seconds"]:
            best = {"layout": names[0], "options": opts, "seconds": t}
    del db
    for name in names[1:]:
        t = timed(name, DeviceBatch(layouts[name]), best["options"])
        if t < best["seconds"]:
            best = {"layout": name, "options": dict(best["options"]), "seconds": t}
    best["trials"] = trials
    return best
