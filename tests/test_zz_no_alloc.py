"""The allocation-free decode entry point (lh_pusch_decode_batch_noalloc, descriptor
block in the library's static device scratch) against lh_pusch_decode_batch, called
through the C ABI: two calls in a row, the second with a smaller batch (the scratch keeps
the first call's contents), bit-exact results, hard bits and HARQ buffers; and the size
limit rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    """n mixed codeblocks (BG, Z, Qm, rate, rv, fillers) with their codeword LLRs."""
    rng = np.random.default_rng(seed)
    cbs, parts = [], []
    cw_off = rm_off = bits_off = 0
    for i in range(n):
        bg = 1 + (i % 2)
        Z = int(rng.choice(O.LIFTING_SIZES[18:]))
        Qm = int(rng.choice([2, 4, 6, 8]))
        F = int(rng.integers(0, 3)) * 8
        N = O.N_short(bg) * Z
        E = max(Qm, int(N * rng.uniform(0.35, 1.2)) // Qm * Qm)
        rv = int(rng.integers(0, 4))
        _, le = O.make_codeblocks(1, bg, Z, E, Qm, float(rng.uniform(3.0, 9.0)), seed=seed * 1000 + i, rv=rv, F=F)
        cbs.append(H.PuschCbDesc(H.RmDesc(cw_off, rm_off, E, bg, Z, rv, Qm, 0, F), bits_off, 24, O.CRC24B))
        parts.append((cw_off, le[0]))
        cw_off += E
        rm_off += N
        bits_off += H.msg_bytes(bg, Z)
    cw = np.zeros(cw_off, np.int8)
    for o, a in parts:
        cw[o:o + a.size] = a
    return cbs, cw, rm_off, bits_off


def _call(fn, ctx, cbs, d_cw, rm_bytes, bits_bytes):
    arr = (H.PuschCbDesc * len(cbs))(*cbs)
    d_rm, d_bits, d_res = H.DeviceBuffer(rm_bytes), H.DeviceBuffer(bits_bytes), H.DeviceBuffer(4 * len(cbs))
    d_bits.zero()
    rc = fn(ctx.handle, arr, len(arr), d_cw.ptr, d_rm.ptr, d_bits.ptr, d_res.ptr, 1, 1, 8, None)
    assert rc == 0, H.lib().lh_last_error().decode()
    H.hip().hipDeviceSynchronize()
    return (d_res.download(np.int32, len(cbs)), d_bits.download(np.uint8, bits_bytes),
            d_rm.download(np.int8, rm_bytes))


def test_noalloc_equals_decode_batch():
    require_gpu()
    L = H.lib()
    ctx = H.Context(0, 4096)
    for n, seed in ((96, 1), (40, 2)):  # second call smaller: stale scratch contents must not matter
        print(f"case noalloc n={n} seed={seed}", flush=True)
        cbs, cw, rm_bytes, bits_bytes = _batch(n, seed)
        d_cw = H.DeviceBuffer.from_array(cw)
        got = _call(L.lh_pusch_decode_batch_noalloc, ctx, cbs, d_cw, rm_bytes, bits_bytes)
        want = _call(L.lh_pusch_decode_batch, ctx, cbs, d_cw, rm_bytes, bits_bytes)
        assert np.array_equal(got[0], want[0]), "results differ"
        assert np.array_equal(got[1], want[1]), "hard bits differ"
        assert np.array_equal(got[2], want[2]), "HARQ buffers differ"
        assert (got[0] > 0).sum() >= 1 and (got[0] == 0).sum() >= 1
    ctx.close()


def test_noalloc_rejects_more_than_the_scratch_holds():
    require_gpu()
    L = H.lib()
    ctx = H.Context(0, 4097)
    n = 4097
    cbs = (H.PuschCbDesc * n)(*[H.PuschCbDesc(H.RmDesc(0, 0, 2, 2, 2, 0, 2, 0, 0), 0, 16, O.CRC16)] * n)
    d = H.DeviceBuffer(1 << 16)
    rc = L.lh_pusch_decode_batch_noalloc(ctx.handle, cbs, n, d.ptr, d.ptr, d.ptr, d.ptr, 1, 1, 8, None)
    assert rc == -3 and b"scratch" in L.lh_last_error()
    ctx.close()
