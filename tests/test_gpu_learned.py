"""Learned int8 decoder (lh_pusch_decode_batch_learned) bit-exact against the learned
oracle for random tables, and the plain decoder as its fallback: a codeblock the learned
pass fails is decoded exactly as the plain oracle decodes it."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(1024, request.param)
    yield c
    c.close()


def _batch(seed, n=24):
    rng = np.random.default_rng(seed)
    specs, parts = [], []
    cw_off = rm_off = bits_off = 0
    for i in range(n):
        bg = 1 + (i % 2)
        Z = int(rng.choice(O.LIFTING_SIZES[18:]))
        Qm = int(rng.choice([2, 4, 6, 8]))
        F = int(rng.integers(0, 3)) * 8 if Z >= 16 else 0
        N = O.N_short(bg) * Z
        E = int(N * rng.uniform(0.4, 1.2)) // Qm * Qm
        rv = 0
        snr = {2: 1.5, 4: 7.0, 6: 12.0, 8: 17.0}[Qm] + float(rng.uniform(-1, 3))
        _, le = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=1000 * seed + i, rv=rv, F=F)
        specs.append(dict(bg=bg, Z=Z, Qm=Qm, F=F, N=N, E=E, rv=rv, cw=cw_off, rm=rm_off, bits=bits_off, llr=le[0]))
        parts.append((cw_off, le[0]))
        cw_off += E
        rm_off += N
        bits_off += H.msg_bytes(bg, Z)
    cw = np.zeros(cw_off, np.int8)
    for o, a in parts:
        cw[o:o + a.size] = a
    return specs, cw, rm_off, bits_off


def _run(ctx, specs, cw, rm_bytes, bits_bytes, iters, fb_iters):
    cbs = [H.PuschCbDesc(H.RmDesc(s["cw"], s["rm"], s["E"], s["bg"], s["Z"], s["rv"], s["Qm"], 0, s["F"]),
                         s["bits"], 24, O.CRC24B) for s in specs]
    d_cw = H.DeviceBuffer.from_array(cw)
    d_rm = H.DeviceBuffer(rm_bytes)
    d_bits = H.DeviceBuffer(bits_bytes)
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * len(cbs))
    ctx.pusch_decode_batch_learned(cbs, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=iters,
                                   nof_fallback_iterations=fb_iters)
    return d_res.download(np.int32, len(cbs)), d_bits.download(np.uint8, bits_bytes)


def _random_tables(rng):
    tabs = {}
    for bg in (1, 2):
        M = O.M_rows(bg)
        a = rng.integers(150, 300, (8, M)).astype(np.uint16)
        b = rng.integers(-300, 600, (8, M)).astype(np.int16)
        tabs[bg] = (a, b)
    return tabs


@pytest.mark.parametrize("seed", [1, 2])
def test_learned_matches_oracle(ctx, seed):
    print(f"case learned vs oracle seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    tabs = _random_tables(rng)
    for bg in (1, 2):
        ctx.set_scale_table(bg, *tabs[bg])
    specs, cw, rmb, bb = _batch(seed)
    res, bits = _run(ctx, specs, cw, rmb, bb, iters=8, fb_iters=0)
    for i, s in enumerate(specs):
        rm = O.rate_dematch(s["llr"], s["bg"], s["Z"], rv=s["rv"], Qm=s["Qm"], F=s["F"])
        r, b = O.decode_learned(rm, s["bg"], s["Z"], *tabs[s["bg"]], F=s["F"], max_iter=8)
        nb = H.msg_bytes(s["bg"], s["Z"])
        assert res[i] == r, f"cb {i}: iterations {res[i]} vs {r}"
        assert np.array_equal(bits[s["bits"]:s["bits"] + nb], b), f"cb {i}: bits differ"


def test_plain_fallback(ctx):
    print("case learned + plain fallback", flush=True)
    rng = np.random.default_rng(9)
    tabs = _random_tables(rng)
    # Layers 1 and 3 of BG1 get no messages at all: the learned pass fails more often.
    tabs[1][0][:, 1] = 0
    tabs[1][0][:, 3] = 0
    for bg in (1, 2):
        ctx.set_scale_table(bg, *tabs[bg])
    specs, cw, rmb, bb = _batch(5, n=32)
    res, bits = _run(ctx, specs, cw, rmb, bb, iters=4, fb_iters=6)
    n_fb = 0
    for i, s in enumerate(specs):
        rm = O.rate_dematch(s["llr"], s["bg"], s["Z"], rv=s["rv"], Qm=s["Qm"], F=s["F"])
        r, b = O.decode_learned(rm, s["bg"], s["Z"], *tabs[s["bg"]], F=s["F"], max_iter=4)
        if r == 0:
            n_fb += 1
            r, b = O.decode(rm, s["bg"], s["Z"], F=s["F"], max_iter=6)
        nb = H.msg_bytes(s["bg"], s["Z"])
        assert res[i] == r, f"cb {i}: iterations {res[i]} vs {r}"
        assert np.array_equal(bits[s["bits"]:s["bits"] + nb], b), f"cb {i}: bits differ"
    assert n_fb > 0


def test_scale_table_validation(ctx):
    M = O.M_rows(1)
    with pytest.raises(H.LhError):
        ctx.set_scale_table(1, np.full((8, M), 600, np.uint16), np.zeros((8, M), np.int16))
    with pytest.raises(H.LhError):
        ctx.set_scale_table(1, np.zeros((9, M), np.uint16), np.zeros((9, M), np.int16))
    with pytest.raises(H.LhError):
        ctx.set_scale_table(2, np.zeros((8, M), np.uint16), np.zeros((8, M), np.int16))
