"""The one-row decoder keeps extension-layer messages in one of two forms, chosen per launch:
compressed words (two minima, argmin, signs) when the launch's items do not all fit the GPU
at once, int8 bytes per edge when they do.  The same low-rate codeblocks (BG1 Z=384 at 46
layers, BG2 Z=208 at 42 layers, near their thresholds) decoded as one launch larger than
the resident slots and in resident-sized chunks give identical results, hard bits and HARQ
buffers, plain and learned; a sample matches the oracle."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def _cases():
    cbs = []
    for k in range(320):
        cbs.append((1, 384, 25344, 2, (-2.6, -1.8, -1.0, -0.2)[k % 4]))
    for k in range(64):
        cbs.append((2, 208, 10400, 2, (-4.5, -3.5, -2.5, -1.5)[k % 4]))
    return cbs


def _decode(ctx, b, learned):
    b.decode(ctx, learned=learned, iters=6, fb_iters=0)
    return (b.d_res.download(np.int32, b.n), b.d_bits.download(np.uint8, b.msg_bytes),
            b.d_rm.download(np.int8, b.d_rm.nbytes))


@pytest.mark.parametrize("learned", [False, True])
def test_one_launch_equals_resident_chunks(learned):
    require_gpu()
    print(f"case extension-layer forms learned={learned}", flush=True)
    from ocudu_hip import nms
    ctx = H.Context(0, max_codeblocks=4096)
    ctx.tune_rows(1, 384, 0)
    ctx.tune_rows(2, 208, 0)
    if learned:
        for path in ("nms_bg1.json", "nms_bg2.json"):
            tbg, aq, bq, _ = nms.load_table(str(nms.TABLE_DIR / path))
            ctx.set_scale_table(tbg, aq, bq)
    cbs = _cases()
    seeds = np.random.default_rng(31).integers(0, 2 ** 63, (len(cbs), 2))
    whole = nms.MixedBatch(ctx, cbs, 0, seeds=seeds)
    res_w, bits_w, rm_w = _decode(ctx, whole, learned)
    assert (res_w > 0).any() and (res_w <= 0).any(), "the batch should hold passing and failing codeblocks"
    rm_off = 0
    for c0 in range(0, len(cbs), 96):
        sl = slice(c0, min(c0 + 96, len(cbs)))
        part = nms.MixedBatch(ctx, cbs[sl], 0, seeds=seeds[sl])
        res_p, bits_p, rm_p = _decode(ctx, part, learned)
        assert np.array_equal(res_p, res_w[sl]), f"chunk at {c0}: results differ"
        o0 = whole.msg_off[c0]
        assert np.array_equal(bits_p, bits_w[o0:o0 + part.msg_bytes]), f"chunk at {c0}: bits differ"
        assert np.array_equal(rm_p, rm_w[rm_off:rm_off + rm_p.size]), f"chunk at {c0}: HARQ buffers differ"
        rm_off += rm_p.size
    if not learned:
        # A sample against the oracle (dematched buffer from the device, decode on the host).
        offs = np.cumsum([0] + [O.N_short(c[0]) * c[1] for c in cbs])
        for i in (0, 3, 161, 318, 322, 383):
            bg, Z = cbs[i][0], cbs[i][1]
            r_o, b_o = O.decode(rm_w[offs[i]:offs[i + 1]], bg, Z, crc_poly=O.CRC24B, max_iter=6)
            o = whole.msg_off[i]
            assert res_w[i] == r_o, f"cb {i}: iterations {res_w[i]} vs oracle {r_o}"
            assert np.array_equal(bits_w[o:o + whole.mbs[i]], b_o), f"cb {i}: bits differ from the oracle"
    ctx.close()
