"""Learned (neural) min-sum decoder: training, quantisation and evaluation.

The model is the layered normalised-offset min-sum decoder with one scale alpha[t][l] and
one offset beta[t][l] per iteration t and layer l (the reference decoders use the fixed
0.8 scaling of lib/phy/upper/channel_coding/ldpc/ldpc_decoder_impl.h:185).  Training
runs entirely on the GPU: codeblocks come from the device generator (encoder, rate
matcher, QAM mapper, AWGN, max-log demapper, int8 quantiser), are rate dematched by the
device dematcher, and the FP32 forward/backward kernels (lh_nms_train_step) and the Adam
kernel (lh_nms_adam) update the tables.  The trained tables are quantised to the int8
decoder's format (alpha, beta in 1/256) and deployed with lh_set_scale_table; the plain
decoder remains the fallback on CRC failure (lh_pusch_decode_batch_learned).

    python -m ocudu_hip.nms train --bg 2 --Z 208 --cbs 1048576 --out tables/nms_bg2.json
"""
import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import (LH_CRC16, LH_CRC24B, Context, DeviceBuffer, Event, GenDesc, PuschCbDesc, RmDesc, TrainDesc, k_b, msg_bytes,
               n_short)

M_LAYERS = {1: 46, 2: 42}
ITERS = 8
TABLE_DIR = Path(__file__).resolve().parent.parent / "tables"


class _View:
    """A sub-range of a device buffer."""

    def __init__(self, buf, off_bytes, nbytes):
        self.ptr = buf.ptr + off_bytes
        self.nbytes = nbytes


def rm_length(bg, Z, rate, qm):
    """E for code rate K_b*Z / E, rounded up to a multiple of Qm."""
    return int(math.ceil(k_b(bg) * Z / rate / qm)) * qm


def layers_for(bg, Z, E):
    """Layers the decoder runs for a rv0 codeblock without fillers (ldpc_decoder_impl.cpp:85-99)."""
    N = n_short(bg) * Z
    nsys = (k_b(bg) - 2) * Z
    mp = max(min(E, N), nsys)
    cbl = max(mp + 2 * Z, (k_b(bg) + 4) * Z)
    L = -(-cbl // Z) - k_b(bg)
    return int(min(max(L, 4), M_LAYERS[bg]))


def n0_of(esn0_db):
    return float(10.0 ** (-esn0_db / 10.0))


def cb_crc(bg, Z):
    """CRC attached to a synthetic codeblock: CRC24B, or CRC16 for codeblocks of at most 48 bits."""
    return (LH_CRC16, 16) if k_b(bg) * Z <= 48 else (LH_CRC24B, 24)


class MixedBatch:
    """Generated codeblocks of any (BG, Z) resident on the device: codeword LLRs,
    rate-dematched soft bits (HARQ buffers), messages and the decoder descriptors."""

    def __init__(self, ctx, cbs, seed, seeds=None):
        """cbs: list of (bg, Z, E, Qm, Es/N0 dB); the generator seeds of codeblock i are
        seeds[i] (message, noise), drawn from `seed` when not given."""
        self.n = len(cbs)
        gen, rms, self.pusch = [], [], []
        self.msg_off, self.mbs, self.info = [], [], []
        cw_off = rm_off = msg_off = 0
        if seeds is None:
            seeds = np.random.default_rng(seed).integers(0, 2 ** 63, (len(cbs), 2))
        for i, (bg, Z, E, qm, snr) in enumerate(cbs):
            poly, order = cb_crc(bg, Z)
            mb = msg_bytes(bg, Z)
            gen.append(GenDesc(cw_off, msg_off, int(seeds[i, 0]), int(seeds[i, 1]), n0_of(snr), E, bg, Z, 0, qm, 0, 0,
                               poly))
            rm = RmDesc(cw_off, rm_off, E, bg, Z, 0, qm, 0, 0)
            rms.append(rm)
            self.pusch.append(PuschCbDesc(rm, msg_off, order, poly))
            self.msg_off.append(msg_off)
            self.mbs.append(mb)
            self.info.append(k_b(bg) * Z - order)
            cw_off = (cw_off + E + 15) // 16 * 16
            rm_off += n_short(bg) * Z
            msg_off += mb
        self.d_cw = DeviceBuffer(cw_off)
        self.d_rm = DeviceBuffer(rm_off)
        self.d_msg = DeviceBuffer(msg_off)
        self.d_bits = DeviceBuffer(msg_off)
        # The decoder leaves the bits past K_b*Z in a last partial byte untouched.
        self.d_bits.zero()
        self.d_res = DeviceBuffer(4 * len(cbs))
        self.msg_bytes = msg_off
        for c0 in range(0, len(cbs), 4096):
            ctx.gen_codeblocks(gen[c0:c0 + 4096], self.d_cw, self.d_msg)
            ctx.rate_dematch(rms[c0:c0 + 4096], self.d_cw, self.d_rm, new_data=True)
        self.info = np.array(self.info)
        self.arr = (PuschCbDesc * len(self.pusch))(*self.pusch)

    def decode(self, ctx, learned=False, iters=ITERS, fb_iters=0, stream=None):
        if learned:
            ctx.pusch_decode_batch_learned(None, self.d_cw, self.d_rm, self.d_bits, self.d_res,
                                           nof_ldpc_iterations=iters, nof_fallback_iterations=fb_iters, stream=stream,
                                           arr=self.arr)
        else:
            ctx.pusch_decode_batch(None, self.d_cw, self.d_rm, self.d_bits, self.d_res, nof_ldpc_iterations=iters,
                                   stream=stream, arr=self.arr)

    def timed_decode(self, ctx, learned=False, iters=ITERS, fb_iters=0, reps=3):
        """Best event-timed duration (ms) of `reps` decodes of the batch (descriptor upload and
        kernels; inputs resident)."""
        e0, e1 = Event(), Event()
        best = float("inf")
        for _ in range(reps):
            e0.record()
            self.decode(ctx, learned, iters, fb_iters)
            e1.record()
            e1.sync()
            best = min(best, e0.elapsed_ms(e1))
        return best

    def outcome(self):
        """(block error flags, iterations) of the last decode: an error is a failed CRC or
        a passed CRC with wrong bits."""
        res = self.d_res.download(np.int32, self.n)
        bits = self.d_bits.download(np.uint8, self.msg_bytes)
        msg = self.d_msg.download(np.uint8, self.msg_bytes)
        wrong = np.array([not np.array_equal(bits[o:o + m], msg[o:o + m]) for o, m in zip(self.msg_off, self.mbs)])
        return (res <= 0) | wrong, res

    def good_info_bits(self, err):
        return int(self.info[~err].sum())


class DeviceBatch(MixedBatch):
    """Codeblocks of one base graph and lifting size, with the descriptors of the
    training step as well."""

    def __init__(self, ctx, bg, Z, cbs, seed):
        """cbs: list of (E, Qm, Es/N0 dB)."""
        super().__init__(ctx, [(bg, Z, E, qm, snr) for E, qm, snr in cbs], seed)
        self.bg, self.Z = bg, Z
        N = n_short(bg) * Z
        mb = msg_bytes(bg, Z)
        self.mb = mb
        self.train = [TrainDesc(i * N, i * mb, N, Z, 0, layers_for(bg, Z, E)) for i, (E, qm, snr) in enumerate(cbs)]
        self.info_bits = int(self.info[0]) if self.n else 0


class Trainer:
    """Adam on the FP32 (alpha, beta) tables of one base graph."""

    def __init__(self, ctx, bg, iters=ITERS, tau=0.125, lr=0.01, alpha0=0.8, beta0=0.0, weights=None):
        self.ctx, self.bg, self.T, self.tau, self.lr = ctx, bg, iters, tau, lr
        self.M = M_LAYERS[bg]
        self.n = 2 * self.T * self.M
        p = np.concatenate([np.full(self.T * self.M, alpha0), np.full(self.T * self.M, beta0)]).astype(np.float32)
        self.d_par = DeviceBuffer.from_array(p)
        self.d_grad, self.d_m, self.d_v = DeviceBuffer(4 * self.n), DeviceBuffer(4 * self.n), DeviceBuffer(4 * self.n)
        self.d_m.zero()
        self.d_v.zero()
        self.w = list(weights) if weights is not None else [1.0] * self.T
        self.step_no = 0
        self.d_ws = None
        self.d_loss = None

    def _bufs(self, batch):
        need = Context.nms_workspace_bytes(self.bg, batch.train, self.T)
        if self.d_ws is None or self.d_ws.nbytes < need:
            self.d_ws = DeviceBuffer(need)
        if self.d_loss is None or self.d_loss.nbytes < 4 * batch.n:
            self.d_loss = DeviceBuffer(4 * batch.n)

    def loss(self, batch):
        self._bufs(batch)
        tm = 4 * self.T * self.M
        self.ctx.nms_train_step(self.bg, batch.train, self.T, _View(self.d_par, 0, tm), _View(self.d_par, tm, tm),
                                self.w, self.tau, batch.d_rm, batch.d_msg, self.d_ws, self.d_loss, backward=False)
        return float(self.d_loss.download(np.float32, batch.n).sum())

    def step(self, batch):
        """One forward/backward pass over the batch and one Adam update; returns the mean loss."""
        self._bufs(batch)
        tm = 4 * self.T * self.M
        self.d_grad.zero()
        self.ctx.nms_train_step(self.bg, batch.train, self.T, _View(self.d_par, 0, tm), _View(self.d_par, tm, tm),
                                self.w, self.tau, batch.d_rm, batch.d_msg, self.d_ws, self.d_loss,
                                _View(self.d_grad, 0, tm), _View(self.d_grad, tm, tm), backward=True)
        self.step_no += 1
        self.ctx.nms_adam(self.d_par, self.d_grad, self.d_m, self.d_v, self.n, self.T * self.M, self.lr,
                          self.step_no)
        return float(self.d_loss.download(np.float32, batch.n).sum()) / batch.n

    def tables(self):
        p = self.d_par.download(np.float32, self.n)
        return p[:self.T * self.M].reshape(self.T, self.M), p[self.T * self.M:].reshape(self.T, self.M)

    def quantized(self):
        a, b = self.tables()
        aq = np.clip(np.rint(a * 256.0), 0, 511).astype(np.uint16)
        bq = np.clip(np.rint(b * 256.0), -32768, 32767).astype(np.int16)
        return aq, bq


def save_table(path, bg, aq, bq, meta):
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        json.dump({"base_graph": bg, "iterations": int(aq.shape[0]), "layers": int(aq.shape[1]),
                   "alpha_q": aq.astype(int).tolist(), "beta_q": bq.astype(int).tolist(), "meta": meta}, f)


def load_table(path):
    with open(path) as f:
        d = json.load(f)
    return d["base_graph"], np.array(d["alpha_q"], np.uint16), np.array(d["beta_q"], np.int16), d.get("meta", {})


def bler(ctx, bg, Z, E, qm, snr, n, seed, learned=False, iters=ITERS, fb_iters=0):
    """(BLER, mean iterations) of n codeblocks at one SNR."""
    b = DeviceBatch(ctx, bg, Z, [(E, qm, snr)] * n, seed)
    b.decode(ctx, learned=learned, iters=iters, fb_iters=fb_iters)
    err, res = b.outcome()
    its = np.where(res > 0, res, iters)
    return float(err.mean()), float(its.mean())


def measure(ctx, batch, learned, iters=ITERS, fb_iters=0, reps=3):
    """BLER, mean iterations, decode time (ms) and info Gbit/s (correct codeblocks only) of
    one decoder on a resident batch."""
    ms = batch.timed_decode(ctx, learned, iters, fb_iters, reps)
    err, res = batch.outcome()
    its = np.where(res > 0, res, iters)
    return {"bler": float(err.mean()), "iters": float(its.mean()), "ms": ms,
            "gbps": batch.good_info_bits(err) / (ms * 1e-3) / 1e9}


def operating_point(ctx, bg, Z, E, qm, lo, hi, n=4096, seed=5, step=0.05, target=0.10, iters=ITERS):
    """Lowest SNR on a `step` dB grid from `lo` at which the plain decoder's BLER <= target."""
    snr = lo
    while snr <= hi + 1e-9:
        e, _ = bler(ctx, bg, Z, E, qm, snr, n, seed, iters=iters)
        if e <= target:
            return round(snr, 4)
        snr += step
    return None


def calibrate(ctx, bg, Z, E, qm, n=512, seed=1, iters=ITERS, lo=-4.0, hi=30.0, targets=(0.30, 0.01)):
    """SNRs (dB) at which the plain decoder's BLER crosses each target (bisection)."""
    out = []
    for tgt in targets:
        a, b = lo, hi
        for _ in range(14):
            m = 0.5 * (a + b)
            e, _ = bler(ctx, bg, Z, E, qm, m, n, seed)
            if e > tgt:
                a = m
            else:
                b = m
        out.append(0.5 * (a + b))
    return out


def train(ctx, bg, Z, configs, total_cbs, batch, seed=0, iters=ITERS, lr=0.01, log=print, weights=None):
    """configs: list of (rate, Qm, snr_lo, snr_hi); each codeblock draws a configuration
    and an SNR uniformly.  weights: per-iteration loss weights (uniform by default; larger
    early weights reward early convergence, i.e. fewer iterations under CRC early stop).
    Returns the trainer."""
    tr = Trainer(ctx, bg, iters=iters, lr=lr, weights=weights)
    rng = np.random.default_rng(seed)
    steps = max(1, total_cbs // batch)
    t0 = time.time()
    for s in range(steps):
        pick = rng.integers(0, len(configs), batch)
        cbs = []
        for k in pick:
            rate, qm, lo, hi = configs[k]
            cbs.append((rm_length(bg, Z, rate, qm), qm, float(rng.uniform(lo, hi))))
        b = DeviceBatch(ctx, bg, Z, cbs, seed=int(rng.integers(0, 2 ** 62)))
        loss = tr.step(b)
        if s % max(1, steps // 20) == 0 or s == steps - 1:
            log(f"step {s + 1}/{steps} loss {loss:.4f} ({time.time() - t0:.1f} s)")
    return tr


def curves(ctx, bg, Z, configs, table, n=8192, points=12, iters=ITERS, seed=2024, log=print):
    """BLER curves of the plain and the learned int8 decoders: for each (rate, Qm, snr_30,
    snr_1) configuration, `points` SNRs from snr_30 - 0.3 dB to snr_1 + 0.5 dB.  Returns rows
    (rate, Qm, snr, plain dict, learned dict, learned + fallback dict)."""
    ctx.set_scale_table(bg, *table)
    rows = []
    for rate, qm, lo, hi in configs:
        E = rm_length(bg, Z, rate, qm)
        for snr in np.linspace(lo - 0.3, hi + 0.5, points):
            b = DeviceBatch(ctx, bg, Z, [(E, qm, float(snr))] * n, seed)
            p = measure(ctx, b, False, iters)
            l = measure(ctx, b, True, iters)
            lf = measure(ctx, b, True, iters, fb_iters=iters)
            rows.append((rate, qm, float(snr), p, l, lf))
            log(f"rate {rate} Qm {qm} {snr:6.2f} dB: plain {p['bler']:.4f} ({p['iters']:.2f} it, {p['gbps']:.2f} Gb/s)"
                f" learned {l['bler']:.4f} ({l['iters']:.2f} it, {l['gbps']:.2f} Gb/s)"
                f" +fallback {lf['bler']:.4f} ({lf['gbps']:.2f} Gb/s)")
    return rows


def curves_markdown(bg, Z, rows, n, command):
    out = [f"# BLER of the plain and learned int8 decoders, BG{bg} Z={Z}", "",
           f"{n} codeblocks per point from the device generator (CRC24B, rv0, AWGN, max-log "
           f"demapping), max {ITERS} iterations, CRC early stop; a block error is a failed CRC or wrong bits.",
           "Gbit/s = info bits of correct codeblocks / event-timed decode (descriptor upload + kernels).", "",
           f"Command: `{command}`", "",
           "| rate | Qm | Es/N0 dB | plain BLER | plain it | plain Gb/s | learned BLER | learned it | learned Gb/s "
           "| ratio | learned+fallback BLER | +fallback Gb/s |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for rate, qm, snr, p, l, lf in rows:
        ratio = l["gbps"] / p["gbps"] if p["gbps"] > 0 else float("nan")
        out.append(f"| {rate} | {qm} | {snr:.2f} | {p['bler']:.4f} | {p['iters']:.2f} | {p['gbps']:.2f} | "
                   f"{l['bler']:.4f} | {l['iters']:.2f} | {l['gbps']:.2f} | {ratio:.2f} | {lf['bler']:.4f} | "
                   f"{lf['gbps']:.2f} |")
    return "\n".join(out) + "\n"


def _main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--bg", type=int, default=2)
    t.add_argument("--Z", type=int, default=208)
    t.add_argument("--rates", default="0.3,0.5,0.7")
    t.add_argument("--qms", default="2,4")
    t.add_argument("--cbs", type=int, default=1 << 20)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=0.01)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    t.add_argument("--weights", help="per-iteration loss weights, comma-separated (default uniform)")
    t.add_argument("--snr-hi-extra", type=float, default=0.0,
                   help="extend each configuration's training SNR range upward by this many dB")
    t.add_argument("--curves", help="also write BLER curves (markdown) of the trained table here")
    args = ap.parse_args(argv)
    ctx = Context(0, max_codeblocks=max(8192, args.batch))

    def log(m):
        print(m, flush=True)

    rates = [float(x) for x in args.rates.split(",")]
    qms = [int(x) for x in args.qms.split(",")]
    configs = []
    for r in rates:
        for q in qms:
            E = rm_length(args.bg, args.Z, r, q)
            s30, s1 = calibrate(ctx, args.bg, args.Z, E, q)
            log(f"rate {r} Qm {q} E {E}: plain BLER 30% at {s30:.2f} dB, 1% at {s1:.2f} dB")
            configs.append((r, q, s30, s1))
    weights = [float(x) for x in args.weights.split(",")] if args.weights else None
    train_cfgs = [(r, q, lo, hi + args.snr_hi_extra) for r, q, lo, hi in configs]
    tr = train(ctx, args.bg, args.Z, train_cfgs, args.cbs, args.batch, seed=args.seed, lr=args.lr, log=log,
               weights=weights)
    aq, bq = tr.quantized()
    save_table(args.out, args.bg, aq, bq, {"Z": args.Z, "configs": configs, "codeblocks": args.cbs,
                                          "batch": args.batch, "lr": args.lr, "weights": weights,
                                          "snr_hi_extra": args.snr_hi_extra})
    log(f"table written to {args.out}")
    if args.curves:
        n = 8192
        rows = curves(ctx, args.bg, args.Z, configs, (aq, bq), n=n, log=log)
        cmd = "python -m ocudu_hip.nms " + " ".join(argv if argv is not None else sys.argv[1:])
        Path(args.curves).write_text(curves_markdown(args.bg, args.Z, rows, n, cmd))
        log(f"curves written to {args.curves}")
    return 0


if __name__ == "__main__":
    sys.exit(_main())
