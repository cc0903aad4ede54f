#!/usr/bin/env python3
"""Writes the golden fixtures from the reference's own code built from source: the
generic decoder's, and the UL-SCH demultiplexer's (ulsch_demux).

The reference CPU code is built from its sources into oracle/_ref/libref_ldpc.so
(oracle/Makefile.ref); this script feeds it seeded synthetic codeblocks and stores the
inputs (rate-matched int8 LLRs + dematching parameters) and the reference outputs
(rate-dematched buffer checksum, hard bits, iteration count / CRC verdict).  The
fixtures are data only; the tests compare the oracle and the HIP path against them.

    python tests/golden/gen_golden.py              # every fixture
    python tests/golden/gen_golden.py high_rate    # only high_rate_l4.npz
    python tests/golden/gen_golden.py demux        # only ulsch_demux.npz (the reference's
                                                   # UL-SCH demultiplexer, libref_demod.so)
"""
import sys
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import oracle as O  # noqa: E402


def run_case(dec, bg, Z, E, Qm, snr, seed, rv=0, F=0, max_iter=6):
    K = O.K_b(bg) * Z
    poly, cb = (O.CRC24B, 24) if K > 48 else (O.CRC16, 16)
    _, le = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=seed, rv=rv, F=F, crc_poly=poly, crc_bits=cb)
    rm = O.ref_rate_dematch(le[0], bg, Z, rv=rv, Qm=Qm, F=F)
    r, bits = O.ref_decode(rm, bg, Z, F=F, crc_poly=poly, max_iter=max_iter, crc_bits=cb, dec=dec)
    return le[0], rm, poly, cb, r, bits


def _pack(rows, llr_blob, bits_blob, path, columns):
    meta = np.array(rows, np.int64)
    offs = np.cumsum([0] + [a.size for a in llr_blob])
    boffs = np.cumsum([0] + [a.size for a in bits_blob])
    np.savez_compressed(path, meta=meta, llr=np.concatenate(llr_blob), llr_off=offs, bits=np.concatenate(bits_blob),
                        bits_off=boffs, columns=np.array(columns.split()))


# Es/N0 (dB) near the 8-iteration decoding threshold at code rate ~0.93, per base graph and
# bits per symbol (calibrated on the oracle's generator).
HIGH_RATE_SNR = {1: {2: 8.6, 4: 14.3, 6: 20.0, 8: 25.6}, 2: {2: 6.4, 4: 12.2, 6: 17.6, 8: 23.4}}


def high_rate(dec):
    """The high-rate regime: rv0 codeblocks whose soft bits end before (K_b + 2) Z, so the
    decoder runs the (K_b + 4) Z floor of four layers (ldpc_decoder_impl.cpp:90-99), with
    8 iterations near threshold.  First the headline slot shape (BG1 Z=384, 256-QAM,
    E=9088) at Es/N0 24.9..25.4 dB, then both base graphs over every third lifting size and
    every modulation, E within (K_b + 2) Z, with and without fillers."""
    rows, llr_blob, bits_blob = [], [], []

    def add(bg, Z, E, Qm, F, snr, seed):
        le, rm, poly, cb, r, b = run_case(dec, bg, Z, E, Qm, snr, seed, F=F, max_iter=8)
        rows.append((bg, Z, E, Qm, 0, F, poly, cb, r, zlib.crc32(rm.tobytes()), int(round(snr * 100))))
        llr_blob.append(le)
        bits_blob.append(b)

    for k, snr in enumerate((24.9, 25.0, 25.1, 25.2, 25.3, 25.4)):
        for c in range(8):
            add(1, 384, 9088, 8, 0, snr, 7000 + 8 * k + c)
    rng = np.random.default_rng(20260103)
    for bg in (1, 2):
        for Z in O.LIFTING_SIZES[::3] + [384]:
            nsys = (O.K_b(bg) - 2) * Z
            for Qm in (2, 4, 6, 8):
                F = int(rng.choice([0, 8, 16])) if nsys > 64 else 0
                E = int((O.K_b(bg) + 2) * Z * rng.uniform(0.93, 1.0) - F) // Qm * Qm
                snr = HIGH_RATE_SNR[bg][Qm] + float(rng.choice([-0.4, 0.0, 0.4]))
                add(bg, Z, E, Qm, F, snr, int(rng.integers(0, 2 ** 31)))
    _pack(rows, llr_blob, bits_blob, HERE / "high_rate_l4.npz", "bg Z E Qm rv F crc_poly crc_bits iters rm_crc32 snr_cdb")


def ulsch_demux():
    """UL-SCH demultiplexing by the reference's ulsch_demultiplex_impl built from source
    (oracle/_ref/libref_demod.so): 48 seeded configurations (tests/demux_util.py), small
    PRB counts to keep the fixture small, int8 codeword soft bits from a seeded generator;
    stored: the configuration words (+ c_init), the input and the four outputs."""
    sys.path.insert(0, str(HERE.parent))
    from demux_util import config_words, random_config

    rng = np.random.default_rng(20260301)
    words, llr, outs = [], [], {k: [] for k in ("sch", "harq_ack", "csi_part1", "csi_part2")}
    while len(words) < 48:
        cfg = random_config(rng, max_prb=6)
        sz = O.demux_sizes(cfg)
        x = rng.integers(-120, 121, sz["nof_llrs"]).astype(np.int8)
        c_init = int(rng.integers(0, 2 ** 31))
        got = O.ref_ulsch_demultiplex(cfg, x, c_init, block=int(rng.integers(1, 3000)))
        if got is None:
            continue
        words.append(config_words(cfg, c_init))
        llr.append(x)
        for k, v in zip(outs, got):
            outs[k].append(v)
    arrays = {"config": np.array(words, np.int64), "llr": np.concatenate(llr),
              "llr_off": np.cumsum([0] + [a.size for a in llr])}
    for k, v in outs.items():
        arrays[k] = np.concatenate(v)
        arrays[k + "_off"] = np.cumsum([0] + [a.size for a in v])
    np.savez_compressed(HERE / "ulsch_demux.npz", **arrays)


def main():
    if sys.argv[1:] == ["demux"]:
        if O.ref_demod() is None:
            raise SystemExit("build oracle/_ref first (make -f oracle/Makefile.ref)")
        ulsch_demux()
        return
    R = O.ref()
    if R is None:
        raise SystemExit("build oracle/_ref first (make -f oracle/Makefile.ref)")
    dec = R.ref_decoder_create(0, 0, 0)  # generic, force_decoding=false, early_stop_syndrome=false
    if sys.argv[1:] == ["high_rate"]:
        high_rate(dec)
        R.ref_decoder_destroy(dec)
        return

    # cfg1: 64 CBs BG1 Z=384, E=16896, QPSK, Es/N0 2 dB, rv0, 6 iterations, CRC24B.
    msgs, llrs = O.make_codeblocks(64, 1, 384, 16896, 2, 2.0, seed=20260101)
    res, bits = [], []
    for c in range(64):
        rm = O.ref_rate_dematch(llrs[c], 1, 384, Qm=2)
        r, b = O.ref_decode(rm, 1, 384, crc_poly=O.CRC24B, max_iter=6, dec=dec)
        res.append(r)
        bits.append(b)
    np.savez_compressed(HERE / "cfg1_bg1_z384_qpsk_2db.npz", llrs=llrs, msgs=msgs, iters=np.array(res, np.int32),
                        bits=np.array(bits), seed=20260101)

    # sweep: every (BG, Z), two rates, two SNRs, random rv/Qm/fillers.
    rng = np.random.default_rng(20260102)
    rows = []
    llr_blob = []
    bits_blob = []
    for bg in (1, 2):
        for Z in O.LIFTING_SIZES:
            N = O.N_short(bg) * Z
            nsys = (O.K_b(bg) - 2) * Z
            for frac in (0.45, 1.1):
                for snr_off in (0.0, 2.5):
                    Qm = int(rng.choice([2, 4, 6, 8]))
                    rv = int(rng.integers(0, 4))
                    F = int(rng.choice([0, 8, 24])) if nsys > 64 else 0
                    E = max(Qm, int(N * frac) // Qm * Qm)
                    snr = {2: 1.0, 4: 6.0, 6: 10.5, 8: 15.0}[Qm] + snr_off + (3.0 if frac < 1 else 0.0)
                    seed = int(rng.integers(0, 2 ** 31))
                    le, rm, poly, cb, r, b = run_case(dec, bg, Z, E, Qm, snr, seed, rv=rv, F=F)
                    rows.append((bg, Z, E, Qm, rv, F, poly, cb, r, zlib.crc32(rm.tobytes())))
                    llr_blob.append(le)
                    bits_blob.append(b)
    _pack([r[:10] for r in rows], llr_blob, bits_blob, HERE / "sweep_bg_z.npz",
          "bg Z E Qm rv F crc_poly crc_bits iters rm_crc32")
    high_rate(dec)
    R.ref_decoder_destroy(dec)
    ulsch_demux()
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()
