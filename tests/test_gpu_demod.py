"""Soft demodulation + descrambling on the device (lh_pusch_demodulate) against the
oracle (orc_pusch_demodulate, pinned to the reference's own scalar demappers and
pseudo-random generator in test_oracle_demod.py), bit-exact: every modulation, blocks of
several 256-symbol tiles with ragged ends, blocks continuing a sequence (seq_offset),
special noise variances (0, negative, NaN), near-zero symbols and symbols on the demapper's
interval boundaries, and one full 4-layer 256-QAM slot."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def _symbols(rng, n):
    iq = (rng.standard_normal(2 * n) * rng.choice([0.05, 0.3, 1, 3], 2 * n)).astype(np.float32)
    nv = rng.choice([1e-4, 1e-2, 0.1, 0.5, 2.0, 0.0, -1.0, np.nan], n).astype(np.float32)
    iq[:64] = 0
    iq[64:128] *= 1e-6
    for k, N in ((200, 42), (400, 170), (600, 10)):
        if k + 100 <= n:
            m = rng.integers(-9, 10, 100).astype(np.float32)
            iq[2 * k:2 * k + 100] = m * np.float32(2) * (np.float32(1) / np.sqrt(np.float32(N)))
    return iq, nv


def _run(blocks):
    """blocks: (iq, nv, modulation, c_init, seq_offset); one call, results per block."""
    descs, iqs, nvs, so, lo = [], [], [], 0, 0
    for iq, nv, mod, ci, off in blocks:
        descs.append(H.DemodDesc(so, lo, nv.size, mod, ci, off))
        iqs.append(iq)
        nvs.append(nv)
        so += nv.size
        lo += nv.size * O.bits_per_symbol(mod)
    ctx = H.Context(0, max_codeblocks=64)
    d_iq = H.DeviceBuffer.from_array(np.concatenate(iqs))
    d_nv = H.DeviceBuffer.from_array(np.concatenate(nvs))
    d_llr = H.DeviceBuffer(max(lo, 1))
    ctx.pusch_demodulate(descs, d_iq, d_nv, d_llr)
    got = d_llr.download(np.int8, lo)
    ctx.close()
    out, o = [], 0
    for iq, nv, mod, ci, off in blocks:
        n = nv.size * O.bits_per_symbol(mod)
        out.append(got[o:o + n])
        o += n
    return out


def test_demodulate_matches_oracle():
    require_gpu()
    rng = np.random.default_rng(21)
    blocks = []
    for mod in (0, 1, 2, 4, 6, 8):
        for n, off in ((1, 0), (700, 0), (1024, 13), (2500, 3 * 1024 + 5), (3276, 70000)):
            iq, nv = _symbols(rng, n)
            blocks.append((iq, nv, mod, int(rng.integers(0, 2**31)), off))
    print(f"case demodulation, {len(blocks)} blocks", flush=True)
    got = _run(blocks)
    for k, (iq, nv, mod, ci, off) in enumerate(blocks):
        want = O.pusch_demodulate(iq, nv, mod, ci, off)
        assert np.array_equal(got[k], want), (mod, nv.size, off, int((got[k] != want).sum()))


def test_full_slot_256qam():
    """One 4-layer 273-PRB slot: 48 blocks of 3276 symbols, each layer's sequence continued
    over its 12 OFDM symbols."""
    require_gpu()
    rng = np.random.default_rng(22)
    per = 273 * 12
    blocks = []
    for b in range(48):
        iq = (rng.standard_normal(2 * per) * 0.7).astype(np.float32)
        nv = rng.uniform(0.01, 0.3, per).astype(np.float32)
        blocks.append((iq, nv, 8, 0x4d2 * 2**15 + 500, (b % 12) * per * 8))
    print("case demodulation, full slot", flush=True)
    got = _run(blocks)
    for k, (iq, nv, mod, ci, off) in enumerate(blocks):
        assert np.array_equal(got[k], O.pusch_demodulate(iq, nv, mod, ci, off)), k


def test_invalid_modulation_is_rejected():
    require_gpu()
    ctx = H.Context(0, max_codeblocks=4)
    d = H.DeviceBuffer(64)
    with pytest.raises(H.LhError):
        ctx.pusch_demodulate([H.DemodDesc(0, 0, 4, 3, 0, 0)], d, d, d)
    ctx.close()
