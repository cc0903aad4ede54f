"""What the trained learned decoder gains over the plain one (tables/nms_bg2.json, trained
by ocudu_hip.nms on 2^20 device-generated BG2 Z=208 codeblocks): at Es/N0 = 1.80 dB, rate
0.5 QPSK, 8192 seeded codeblocks (the device generator is deterministic), the learned
int8 decoder's BLER is below half the plain decoder's and it needs fewer iterations; with
the plain decoder as fallback the BLER does not rise."""
from pathlib import Path

import pytest

import ocudu_hip as H
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

TABLE = Path(__file__).resolve().parent.parent / "ocudu-dapp-platform_amd" / "tables" / "nms_bg2.json"


def test_learned_gain_bg2_z208():
    print("case learned gain", flush=True)
    require_gpu()
    from ocudu_hip import nms
    ctx = H.Context(0, max_codeblocks=8192)
    bg, aq, bq, _ = nms.load_table(TABLE)
    assert bg == 2
    ctx.set_scale_table(bg, aq, bq)
    E = nms.rm_length(2, 208, 0.5, 2)
    b = nms.DeviceBatch(ctx, 2, 208, [(E, 2, 1.80)] * 8192, seed=2024)
    p = nms.measure(ctx, b, False)
    l = nms.measure(ctx, b, True)
    lf = nms.measure(ctx, b, True, fb_iters=8)
    print(f"plain {p} learned {l} learned+fallback {lf}", flush=True)
    assert p["bler"] > 0.05  # a point in the waterfall
    assert l["bler"] < 0.5 * p["bler"]
    assert l["iters"] < p["iters"]
    assert lf["bler"] <= l["bler"]
    ctx.close()
