"""Synthetic PUSCH workloads on the device generator.

Uneven slot batch: codeblocks of many UEs in one persistent launch -- per UE a base graph
(50/50), a lifting size (uniform over the 51 NR sizes), a code rate (uniform in
[0.2, 0.93]), a modulation (QPSK..256QAM) and E = K_b Z / R rounded up to Qm, at the
Es/N0 where the plain decoder's BLER is about 10% for that UE's configuration.
"""
import numpy as np

from . import LIFTING_SIZES
from .nms import MixedBatch, bler, rm_length

QMS = (2, 4, 6, 8)


def ue_configs(n_ue=64, seed=0, rate_lo=0.2, rate_hi=0.93):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_ue):
        bg = int(rng.integers(1, 3))
        Z = int(rng.choice(LIFTING_SIZES))
        rate = float(rng.uniform(rate_lo, rate_hi))
        qm = int(rng.choice(QMS))
        out.append((bg, Z, rm_length(bg, Z, rate, qm), qm, rate))
    return out


def snr_for_bler(ctx, bg, Z, E, qm, target=0.10, n=128, seed=3, lo=-6.0, hi=36.0, steps=10, iters=None):
    """Es/N0 (dB) where the plain decoder's BLER crosses `target` (bisection)."""
    kw = {} if iters is None else {"iters": iters}
    for _ in range(steps):
        m = 0.5 * (lo + hi)
        e, _ = bler(ctx, bg, Z, E, qm, m, n, seed, **kw)
        if e > target:
            lo = m
        else:
            hi = m
    return 0.5 * (lo + hi)


def uneven_batch(ctx, n_ue=64, cbs_per_ue=64, seed=0, target=0.10, log=None):
    """(MixedBatch, per-UE configurations with their Es/N0)."""
    ues = []
    for bg, Z, E, qm, rate in ue_configs(n_ue, seed):
        snr = snr_for_bler(ctx, bg, Z, E, qm, target)
        ues.append({"bg": bg, "Z": Z, "E": E, "Qm": qm, "rate": round(rate, 4), "esn0_db": round(snr, 3)})
        if log:
            log(f"UE BG{bg} Z={Z} R={rate:.2f} Qm={qm} E={E}: {snr:.2f} dB")
    cbs = [(u["bg"], u["Z"], u["E"], u["Qm"], u["esn0_db"]) for u in ues for _ in range(cbs_per_ue)]
    return MixedBatch(ctx, cbs, seed + 1), ues
