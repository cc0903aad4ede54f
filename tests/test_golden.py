"""Golden fixtures written from the reference's own generic decoder (built from its
sources, tests/golden/gen_golden.py): the oracle (CPU) and the HIP path (GPU) must
reproduce the dematched buffer, hard bits and iteration count/CRC verdict."""
import zlib
from pathlib import Path

import numpy as np
import pytest

import oracle as O

G = Path(__file__).resolve().parent / "golden"


def _sweep(name="sweep_bg_z.npz"):
    d = np.load(G / name)
    out = []
    for k, row in enumerate(d["meta"]):
        bg, Z, E, Qm, rv, F, poly, cb, iters, rmcrc = (int(x) for x in row[:10])
        llr = d["llr"][d["llr_off"][k]:d["llr_off"][k + 1]]
        bits = d["bits"][d["bits_off"][k]:d["bits_off"][k + 1]]
        out.append(dict(bg=bg, Z=Z, E=E, Qm=Qm, rv=rv, F=F, poly=poly, cb=cb, iters=iters, rmcrc=rmcrc, llr=llr,
                        bits=bits))
    return out


def test_oracle_cfg1():
    d = np.load(G / "cfg1_bg1_z384_qpsk_2db.npz")
    for c in range(d["llrs"].shape[0]):
        rm = O.rate_dematch(d["llrs"][c], 1, 384, Qm=2)
        r, b = O.decode(rm, 1, 384, crc_poly=O.CRC24B, max_iter=6)
        assert r == d["iters"][c] and np.array_equal(b, d["bits"][c]), c
        if r:
            assert np.array_equal(b, d["msgs"][c])


def test_oracle_sweep():
    for k, c in enumerate(_sweep()):
        rm = O.rate_dematch(c["llr"], c["bg"], c["Z"], rv=c["rv"], Qm=c["Qm"], F=c["F"])
        assert zlib.crc32(rm.tobytes()) == c["rmcrc"], k
        r, b = O.decode(rm, c["bg"], c["Z"], F=c["F"], crc_poly=c["poly"], max_iter=6, crc_bits=c["cb"])
        assert r == c["iters"] and np.array_equal(b, c["bits"]), (k, c["bg"], c["Z"])


def test_oracle_high_rate():
    """Fixtures of the four-layer regime (soft bits end before (K_b + 2) Z: the (K_b + 4) Z
    floor of ldpc_decoder_impl.cpp:90-99), 8 iterations near threshold, incl. the headline
    slot shape BG1 Z=384 256-QAM E=9088."""
    cases = _sweep("high_rate_l4.npz")
    assert sum(c["iters"] == 0 for c in cases) >= 40 and sum(c["iters"] > 0 for c in cases) >= 40
    for k, c in enumerate(cases):
        rm = O.rate_dematch(c["llr"], c["bg"], c["Z"], rv=c["rv"], Qm=c["Qm"], F=c["F"])
        assert zlib.crc32(rm.tobytes()) == c["rmcrc"], k
        assert np.flatnonzero(rm).max() < (O.K_b(c["bg"]) + 2) * c["Z"], k  # four layers
        r, b = O.decode(rm, c["bg"], c["Z"], F=c["F"], crc_poly=c["poly"], max_iter=8, crc_bits=c["cb"])
        assert r == c["iters"] and np.array_equal(b, c["bits"]), (k, c["bg"], c["Z"])


def _gpu_fused(name, max_iter, rows):
    import ocudu_hip as H
    from gpu_util import make_ctx

    ctx = make_ctx(1024, rows)
    cases = _sweep(name)
    cbs, parts = [], []
    cw_off = rm_off = bo = 0
    for c in cases:
        N = O.N_short(c["bg"]) * c["Z"]
        cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, c["E"], c["bg"], c["Z"], c["rv"], c["Qm"], 0, c["F"]), bo,
                                 c["cb"], c["poly"]))
        parts.append((cw_off, c["llr"]))
        c["_o"] = (rm_off, N, bo, H.msg_bytes(c["bg"], c["Z"]))
        cw_off += c["E"]
        rm_off += N
        bo += H.msg_bytes(c["bg"], c["Z"])
    cw = np.zeros(cw_off, np.int8)
    for o, a in parts:
        cw[o:o + a.size] = a
    d_cw = H.DeviceBuffer.from_array(cw)
    d_rm, d_bits, d_res = H.DeviceBuffer(rm_off), H.DeviceBuffer(bo), H.DeviceBuffer(4 * len(cases))
    d_bits.zero()
    ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=max_iter)
    res = d_res.download(np.int32, len(cases))
    rm = d_rm.download(np.int8, rm_off)
    bits = d_bits.download(np.uint8, bo)
    for k, c in enumerate(cases):
        ro, N, b0, nb = c["_o"]
        assert zlib.crc32(rm[ro:ro + N].tobytes()) == c["rmcrc"], k
        assert res[k] == c["iters"] and np.array_equal(bits[b0:b0 + nb], c["bits"]), (k, c["bg"], c["Z"])
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rows", ["auto", "pairs"])
def test_gpu_sweep_fused(rows):
    print(f"case golden sweep rows={rows}", flush=True)
    _gpu_fused("sweep_bg_z.npz", 6, rows)


@pytest.mark.gpu
@pytest.mark.parametrize("rows", ["auto", "pairs"])
def test_gpu_high_rate_fused(rows):
    print(f"case golden high rate rows={rows}", flush=True)
    _gpu_fused("high_rate_l4.npz", 8, rows)


@pytest.mark.gpu
def test_gpu_cfg1():
    import ocudu_hip as H
    from gpu_util import decode_batch_gpu, require_gpu

    require_gpu()
    ctx = H.Context(0, 128)
    d = np.load(G / "cfg1_bg1_z384_qpsk_2db.npz")
    rms = [O.rate_dematch(d["llrs"][c], 1, 384, Qm=2) for c in range(64)]
    res, bits = decode_batch_gpu(ctx, rms, [(1, 384, 0, 24, 6)] * 64, [O.CRC24B] * 64)
    assert np.array_equal(res, d["iters"])
    for c in range(64):
        assert np.array_equal(bits[c], d["bits"][c]), c
    ctx.close()
