"""Autotuner of the persistent decoder's launch shape per (base graph, lifting size).

For every (BG, Z) it times the fused PUSCH decode of a device-generated batch under each
candidate -- the row mode (one check row per thread, or two for even Z), G codeblocks
packed per workgroup (G * Z or G * Z/2 lanes of the workgroup's waves) and LDS staging of
the codeword LLRs on or off -- and keeps the fastest.  The table is saved as JSON and
applied to a context with lh_tune_set and lh_tune_set_rows: a class whose one-row mode
won keeps it (row_pairs 0); where row pairs won, the mode stays decided per launch
(row_pairs -1: row pairs only in launches that fill the GPU with such classes alone).

    python -m ocudu_hip.autotune --out tables/tune.json
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import Context
from . import LIFTING_SIZES
from .nms import DeviceBatch, rm_length

MAX_THREADS = 384
MAX_GROUPS = 64
# Representative codeblocks: rate 1/2 at a comfortable SNR and rate 0.9 near its
# threshold, 16-QAM (a few iterations each, both layer counts).
MIX = ((0.5, 4, 10.0), (0.9, 4, 15.0))


def lanes(Z, pairs):
    return Z // 2 if pairs else Z


def default_groups(Z, pairs=False):
    """Codeblocks per workgroup without tuning in a batch of one lifting size."""
    n = lanes(Z, pairs)
    return min(MAX_GROUPS, ((n + 63) // 64 * 64) // n)


def candidates(Z, pairs=False):
    n = lanes(Z, pairs)
    top = min(MAX_GROUPS, MAX_THREADS // n)
    gs = sorted({g for g in (1, 2, 4, 8, 16, 32, 64, top, default_groups(Z, pairs)) if 1 <= g <= top})
    return [(g, st) for g in gs for st in (1, 0)]


def tune_one(ctx, bg, Z, n_cb, seed=1, reps=3):
    """Times every (row mode, G, staging) candidate; returns the best and the default-shape
    time (mode from occupancy, G as many as fit, staging off)."""
    cbs = [(rm_length(bg, Z, r, q), q, snr) for r, q, snr in MIX for _ in range(n_cb // len(MIX))]
    b = DeviceBatch(ctx, bg, Z, cbs, seed)
    ctx.tune_set(bg, Z, 0, False)
    ctx.tune_rows(bg, Z, -1)
    b.timed_decode(ctx, reps=1)  # warm-up
    default = b.timed_decode(ctx, reps=reps)
    times = {}
    for pairs in ((False, True) if Z % 2 == 0 else (False,)):
        ctx.tune_rows(bg, Z, int(pairs))
        for g, st in candidates(Z, pairs):
            ctx.tune_set(bg, Z, g, st)
            b.timed_decode(ctx, reps=1)  # warm-up
            times[(pairs, g, st)] = b.timed_decode(ctx, reps=reps)
    best = min(times, key=times.get)
    pairs, g, st = best
    ctx.tune_set(bg, Z, g, st)
    ctx.tune_rows(bg, Z, -1 if pairs else 0)
    return best, times, default


def tune(ctx, bgs=(1, 2), sizes=LIFTING_SIZES, n_cb=1024, log=print):
    table = []
    for bg in bgs:
        for Z in sizes:
            (pairs, g, st), times, default = tune_one(ctx, bg, Z, n_cb)
            t = times[(pairs, g, st)]
            table.append({"bg": bg, "Z": Z, "row_pairs": -1 if pairs else 0, "groups": g, "stage_llrs": st,
                          "ms": round(t, 4), "default_ms": round(default, 4)})
            log(f"BG{bg} Z={Z}: {'row pairs' if pairs else 'one row'} groups {g} stage {st} {t:.3f} ms "
                f"(default {default:.3f} ms)")
    return table


def save(path, table, meta):
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text(json.dumps({"entries": table, "meta": meta}, indent=1))


def apply(ctx, path):
    """Loads a tuning table into a context; returns the number of entries."""
    d = json.loads(Path(path).read_text())
    for e in d["entries"]:
        ctx.tune_set(e["bg"], e["Z"], e["groups"], bool(e["stage_llrs"]))
        ctx.tune_rows(e["bg"], e["Z"], e.get("row_pairs", -1))
    return len(d["entries"])


def _main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--cbs", type=int, default=1024)
    args = ap.parse_args(argv)
    ctx = Context(0, max_codeblocks=max(4096, args.cbs))
    t0 = time.time()
    table = tune(ctx, n_cb=args.cbs, log=lambda m: print(m, flush=True))
    save(args.out, table, {"codeblocks_per_batch": args.cbs, "mix": MIX, "seconds": round(time.time() - t0, 1)})
    print(f"tuning table written to {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())
