"""Uneven slot batch in one persistent launch: 16 UEs x 16 codeblocks of random BG, Z,
code rate and modulation (some with E > N, the wrap/combine branch) decode to the same
results and bits in the mix as each UE alone, with default and with autotuned launch
shapes; every codeblock of the mix (result, hard bits, HARQ buffer) matches the oracle
bit-exactly."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=ROW_MODES)
def ctx(request):
    c = make_ctx(4096, request.param)
    yield c
    c.close()


def _outputs(b):
    return b.d_res.download(np.int32, b.n), b.d_bits.download(np.uint8, b.msg_bytes)


def test_uneven_mix_equals_alone(ctx):
    print("case uneven mix vs alone", flush=True)
    from ocudu_hip import nms, workloads
    ues = workloads.ue_configs(16, seed=5)
    cbs = []
    for bg, Z, E, qm, rate in ues:
        snr = {2: 2.0, 4: 8.0, 6: 13.0, 8: 19.0}[qm] + 8.0 * (rate - 0.5)
        cbs += [(bg, Z, E, qm, snr)] * 16
    mix = nms.MixedBatch(ctx, cbs, seed=9)
    mix.decode(ctx)
    res_mix, bits_mix = _outputs(mix)
    # Each UE alone: the same generator seeds, so regenerate per UE and decode alone.
    seeds = np.random.default_rng(9).integers(0, 2 ** 63, (len(cbs), 2))
    for u in range(16):
        sl = slice(16 * u, 16 * u + 16)
        # The UE's batch with the generator seeds it had in the mix.
        sub = nms.MixedBatch(ctx, cbs[sl], 0, seeds=seeds[sl])
        sub.decode(ctx)
        res_a, bits_a = _outputs(sub)
        assert np.array_equal(res_a, res_mix[sl]), f"UE {u}: results differ"
        o0 = mix.msg_off[16 * u]
        assert np.array_equal(bits_a, bits_mix[o0:o0 + sub.msg_bytes]), f"UE {u}: bits differ"
    # Autotuned shapes (every candidate packing is legal): same results again.
    for bg, Z, *_ in ues:
        ctx.tune_set(bg, Z, 1 + (Z % 3), Z % 2)
    mix.decode(ctx)
    res_t, bits_t = _outputs(mix)
    assert np.array_equal(res_t, res_mix) and np.array_equal(bits_t, bits_mix)
    for bg, Z, *_ in ues:
        ctx.tune_set(bg, Z, 0, 0)
    # Every codeblock against the oracle: HARQ buffer, result, hard bits.
    cw = mix.d_cw.download(np.int8, mix.d_cw.nbytes)
    rm_dev = mix.d_rm.download(np.int8, mix.d_rm.nbytes)

    def oracle(i):
        bg, Z, E, qm, _ = cbs[i]
        d = mix.pusch[i]
        rm = O.rate_dematch(cw[d.rm.cw_offset:d.rm.cw_offset + E], bg, Z, Qm=qm)
        return rm, O.decode(rm, bg, Z, crc_poly=d.crc_poly, crc_bits=d.nof_crc_bits, max_iter=8)

    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(oracle, range(len(cbs))))
    for i, (rm, (r, b)) in enumerate(want):
        d = mix.pusch[i]
        assert np.array_equal(rm_dev[d.rm.rm_offset:d.rm.rm_offset + rm.size], rm), f"cb {i}: HARQ buffer differs"
        assert res_mix[i] == r, f"cb {i}: {res_mix[i]} vs oracle {r}"
        o = mix.msg_off[i]
        assert np.array_equal(bits_mix[o:o + b.size], b), f"cb {i}: bits differ from the oracle"
    assert 0 < (res_mix > 0).sum() < len(cbs)
