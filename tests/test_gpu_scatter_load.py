"""GPU parity of the fused PUSCH decode's load of new data, over the shapes that select
its scatter form (lifting size a multiple of 8, E no larger than the non-filler
circular buffer) and those next to it: every modulation order, every RV (wrap-around of
the circular buffer), fillers crossing the LLR ranks, N_ref, codeword and HARQ offsets
at every alignment, E = P exactly, one LLR per codeblock, and codeblocks with E > P or
Z not a multiple of 8 in the same batch.  HARQ buffers, hard bits and iteration counts
are checked against the oracle's rate dematcher and decoder."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(4096, request.param)
    yield c
    c.close()


def _n_cb(bg, Z, Nref):
    N = O.N_short(bg) * Z
    return min(Nref, N) if Nref > 0 else N


def _run(ctx, cases, seed, snr=None, iters=6):
    """cases: (bg, Z, rv, Qm, F, Nref, E, cw_pad, rm_pad); returns nothing, asserts parity."""
    rng = np.random.default_rng(seed)
    cbs, cw_parts, expect = [], [], []
    cw_off = rm_off = bits_off = 0
    for i, (bg, Z, rv, Qm, F, Nref, E, cw_pad, rm_pad) in enumerate(cases):
        cw_off += cw_pad
        rm_off += rm_pad
        N = O.N_short(bg) * Z
        if snr is None:
            le = rng.integers(-127, 128, E).astype(np.int8)
        else:
            _, lea = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=seed * 1000 + i, rv=rv, F=F, Nref=Nref)
            le = lea[0]
        cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, rv, Qm, Nref, F), bits_off, 24, O.CRC24B))
        cw_parts.append((cw_off, le))
        rm_exp = O.rate_dematch(le, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F)
        r_o, b_o = O.decode(rm_exp, bg, Z, F=F, crc_poly=O.CRC24B, max_iter=iters)
        expect.append((rm_off, N, rm_exp, bits_off, H.msg_bytes(bg, Z), r_o, b_o))
        cw_off += E
        rm_off += N
        bits_off += H.msg_bytes(bg, Z)
    cw = np.zeros(cw_off + 16, np.int8)
    for o, a in cw_parts:
        cw[o:o + a.size] = a
    d_cw = H.DeviceBuffer.from_array(cw)
    # Stale HARQ contents: new data must overwrite every position.
    d_rm = H.DeviceBuffer.from_array(rng.integers(-127, 128, rm_off + 16).astype(np.int8))
    d_bits = H.DeviceBuffer(bits_off + 16)
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * len(cbs))
    ctx.pusch_decode_batch(cbs, d_cw, d_rm, d_bits, d_res, new_data=True, use_early_stop=True,
                           nof_ldpc_iterations=iters)
    res = d_res.download(np.int32, len(cbs))
    rm = d_rm.download(np.int8, rm_off)
    bits = d_bits.download(np.uint8, bits_off)
    for i, (ro, N, rm_exp, bo, nb, r_o, b_o) in enumerate(expect):
        assert np.array_equal(rm[ro:ro + N], rm_exp), f"case {cases[i]}: HARQ buffer differs"
        assert res[i] == r_o, f"case {cases[i]}: iterations {res[i]} vs oracle {r_o}"
        assert np.array_equal(bits[bo:bo + nb], b_o), f"case {cases[i]}: bits differ"


def _cases(rng, n, zs):
    out = []
    for _ in range(n):
        bg = int(rng.integers(1, 3))
        Z = int(rng.choice(zs))
        Qm = int(rng.choice([1, 2, 4, 6, 8]))
        rv = int(rng.integers(0, 4))
        nsys = (O.K_b(bg) - 2) * Z
        F = int(rng.choice([0, 0, 8, Z // 2, Z + 8, 2 * Z])) if nsys > 2 * Z + 8 else 0
        N = O.N_short(bg) * Z
        Nref = int(rng.choice([0, 0, int(N * 0.8)]))
        if Nref and Nref <= nsys:
            Nref = 0
        P = _n_cb(bg, Z, Nref) - F
        kind = rng.uniform()
        if kind < 0.1:
            E = P // Qm * Qm  # E as close to P as the modulation allows
        elif kind < 0.15:
            E = Qm  # one symbol
        elif kind < 0.85:
            E = max(Qm, int(P * rng.uniform(0.1, 1.0)) // Qm * Qm)
        else:
            E = max(Qm, int(P * rng.uniform(1.0, 1.6)) // Qm * Qm)  # repetition: combining load
        out.append((bg, Z, rv, Qm, F, Nref, E, int(rng.integers(0, 16)), int(rng.integers(0, 16))))
    return out


Z8 = [z for z in O.LIFTING_SIZES if z % 8 == 0]


def test_scatter_load_sweep_random_llrs(ctx):
    # Random LLRs (the decoder mostly fails): the HARQ buffers and the trajectories of
    # failing codeblocks, over every load shape.
    print("case scatter sweep random", flush=True)
    rng = np.random.default_rng(11)
    _run(ctx, _cases(rng, 160, Z8 + [15, 36, 52, 104, 208]), seed=1, iters=3)


def test_scatter_load_sweep_coded(ctx):
    # Encoded codeblocks near their decoding threshold (iteration counts vary).
    print("case scatter sweep coded", flush=True)
    rng = np.random.default_rng(12)
    cases = [c for c in _cases(rng, 60, [z for z in Z8 if z >= 32]) if c[3] != 1]
    _run(ctx, cases, seed=2, snr=6.0, iters=8)


def test_scatter_load_headline_shape(ctx):
    # The slot batch shape (BG1 Z=384, 256-QAM, E=9088) at 16-byte aligned and unaligned
    # codeword offsets, next to a short BG2 codeblock.
    print("case scatter headline", flush=True)
    cases = [(1, 384, 0, 8, 0, 0, 9088, 0, 0), (1, 384, 0, 8, 0, 0, 9088, 8, 4), (1, 384, 0, 8, 0, 0, 9088, 3, 1),
             (2, 208, 0, 4, 24, 0, 2972 // 4 * 4, 5, 2), (1, 384, 2, 8, 0, 0, 9088, 0, 0)]
    _run(ctx, cases, seed=3, snr=25.0, iters=8)
