"""Autotuner of the persistent decoder's launch shape per (base graph, lifting size).

For every (BG, Z) it times the fused PUSCH decode of a device-generated batch under each
candidate -- G codeblocks packed per workgroup (G * Z lanes of the workgroup's waves) and
LDS staging of the codeword LLRs on or off -- and keeps the fastest.  The table is saved
as JSON and applied to a context with lh_tune_set.

    python -m ocudu_hip.autotune --out tables/tune.json
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import Context
from .datagen import LIFTING_SIZES
from .nms import DeviceBatch, rm_length

MAX_THREADS = 384
MAX_GROUPS = 64
# Representative codeblocks: rate 1/2 at a comfortable SNR and rate 0.9 near its
# threshold, 16-QAM (a few iterations each, both layer counts).
MIX = ((0.5, 4, 10.0), (0.9, 4, 15.0))


def default_groups(Z):
    """Codeblocks per workgroup without tuning in a batch of one lifting size."""
    return min(MAX_GROUPS, ((Z + 63) // 64 * 64) // Z)


def candidates(Z):
    top = min(MAX_GROUPS, MAX_THREADS // Z)
    gs = sorted({g for g in (1, 2, 4, 8, 16, 32, 64, top, default_groups(Z)) if 1 <= g <= top})
    return [(g, st) for g in gs for st in (1, 0)]


def tune_one(ctx, bg, Z, n_cb, seed=1, reps=3):
    cbs = [(rm_length(bg, Z, r, q), q, snr) for r, q, snr in MIX for _ in range(n_cb // len(MIX))]
    b = DeviceBatch(ctx, bg, Z, cbs, seed)
    times = {}
    for g, st in candidates(Z):
        ctx.tune_set(bg, Z, g, st)
        b.timed_decode(ctx, reps=1)  # warm-up
        times[(g, st)] = b.timed_decode(ctx, reps=reps)
    best = min(times, key=times.get)
    ctx.tune_set(bg, Z, *best)
    return best, times


def tune(ctx, bgs=(1, 2), sizes=LIFTING_SIZES, n_cb=1024, log=print):
    table = []
    for bg in bgs:
        for Z in sizes:
            (g, st), times = tune_one(ctx, bg, Z, n_cb)
            default = times[(default_groups(Z), 1)]
            table.append({"bg": bg, "Z": Z, "groups": g, "stage_llrs": st, "ms": round(times[(g, st)], 4),
                          "default_ms": round(default, 4)})
            log(f"BG{bg} Z={Z}: groups {g} stage {st} {times[(g, st)]:.3f} ms (default {default:.3f} ms)")
    return table


def save(path, table, meta):
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text(json.dumps({"entries": table, "meta": meta}, indent=1))


def apply(ctx, path):
    """Loads a tuning table into a context; returns the number of entries."""
    d = json.loads(Path(path).read_text())
    for e in d["entries"]:
        ctx.tune_set(e["bg"], e["Z"], e["groups"], bool(e["stage_llrs"]))
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
