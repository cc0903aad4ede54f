"""Device generator (lh_gen_codeblocks) against its CPU statement (oracle.gen_codeblock).

Messages are bit-exact.  LLRs: the integer streams are identical, the Gaussian draws use
double-precision log/cos, whose device library may differ from the host libm in the last
ulp; that moves a quantisation boundary only in rare cases.  Tolerance: every LLR within
1 of the oracle's and at most 1e-4 of them different at all."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    require_gpu()
    c = H.Context(0, max_codeblocks=1024)
    yield c
    c.close()


def test_generator_matches_oracle(ctx):
    print("case generator vs oracle", flush=True)
    rng = np.random.default_rng(21)
    descs, expect = [], []
    llr_off = msg_off = 0
    sizes = list(O.LIFTING_SIZES)
    for i in range(96):
        bg = 1 + (i % 2)
        Z = int(sizes[i % len(sizes)]) if i < 51 * 2 else int(rng.choice(sizes))
        Qm = int(rng.choice([2, 4, 6, 8]))
        K, N = O.K_b(bg) * Z, O.N_short(bg) * Z
        nsys = (O.K_b(bg) - 2) * Z
        poly = int(rng.choice([O.CRC24A, O.CRC24B, O.CRC16, 0])) if K > 64 else 0
        order = {O.CRC24A: 24, O.CRC24B: 24, O.CRC16: 16, 0: 0}[poly]
        F = int(rng.integers(0, min(nsys, K - order - 8) // 4)) if i % 3 == 0 else 0
        rv = int(rng.integers(0, 4))
        Nref = int(rng.integers(nsys + 1, N)) if i % 7 == 3 else 0
        E = max(Qm, int(N * rng.uniform(0.3, 1.5)) // Qm * Qm)
        n0 = float(10 ** (-rng.uniform(-2, 25) / 10)) if i % 11 else 0.0
        sm, sn = int(rng.integers(0, 2 ** 63)), int(rng.integers(0, 2 ** 63))
        descs.append(H.GenDesc(llr_off, msg_off, sm, sn, n0, E, bg, Z, rv, Qm, Nref, F, poly))
        expect.append((llr_off, msg_off, O.gen_codeblock(bg, Z, E, Qm, n0, sm, sn, rv=rv, F=F, crc_poly=poly,
                                                         Nref=Nref)))
        llr_off += E
        msg_off += (K + 7) // 8
    d_llr = H.DeviceBuffer(llr_off)
    d_msg = H.DeviceBuffer(msg_off)
    ctx.gen_codeblocks(descs, d_llr, d_msg)
    llr = d_llr.download(np.int8, llr_off)
    msg = d_msg.download(np.uint8, msg_off)
    ndiff = 0
    for i, (lo, mo, (m_exp, l_exp)) in enumerate(expect):
        assert np.array_equal(msg[mo:mo + m_exp.size], m_exp), f"cb {i}: message differs"
        d = llr[lo:lo + l_exp.size].astype(np.int16) - l_exp
        assert np.abs(d).max() <= 1, f"cb {i}: LLR off by {np.abs(d).max()}"
        ndiff += int((d != 0).sum())
    assert ndiff <= 1e-4 * llr_off, (ndiff, llr_off)
    print(f"llrs {llr_off} differing {ndiff}", flush=True)


def test_generator_rejects_bad_descriptors(ctx):
    d = H.DeviceBuffer(64)
    good = dict(rm_length=600, base_graph=1, lifting_size=16, rv=0, modulation_order=2, Nref=0,
                nof_filler_bits=0, crc_poly=O.CRC24B)
    for k, v in [("modulation_order", 1), ("lifting_size", 17), ("rm_length", 601), ("crc_poly", 0x1234),
                 ("base_graph", 3)]:
        args = dict(good, **{k: v})
        desc = H.GenDesc(0, 0, 1, 2, 0.1, *(args[f] for f in ("rm_length", "base_graph", "lifting_size", "rv",
                                                                "modulation_order", "Nref", "nof_filler_bits",
                                                                "crc_poly")))
        with pytest.raises(H.LhError):
            ctx.gen_codeblocks([desc], d, d)
