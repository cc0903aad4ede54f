"""Consecutive decode batches on two streams, each with its own HARQ buffers, bits and
results, overlap on the GPU (a launch starts on the compute units the other one's last
codeblocks leave idle).  Every launch carries its own descriptors and work counters, so
the overlapped launches give exactly the results of the same batch decoded alone: CRC
verdicts and iteration counts, hard bits and HARQ buffers, near the decoding threshold
(passing and failing codeblocks), in both row modes; a sample matches the oracle."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import ROW_MODES, make_ctx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows", ROW_MODES)
def test_two_streams_equal_one(rows):
    print(f"case two streams rows={rows}", flush=True)
    bg, Z, E, Qm = 1, 384, 9088, 8
    K, N = 22 * Z, 66 * Z
    mb = K // 8
    n = 1216
    ctx = make_ctx(4096, rows)
    d_cw, d_msg = H.DeviceBuffer(n * E), H.DeviceBuffer(n * mb)
    seeds = np.random.default_rng(11).integers(0, 2 ** 63, (n, 2))
    # Es/N0 25.0..25.6 dB: about half the codeblocks fail at 8 iterations.
    gen = [H.GenDesc(i * E, i * mb, int(seeds[i, 0]), int(seeds[i, 1]), float(10 ** (-(25.0 + 0.2 * (i % 4)) / 10)),
                     E, bg, Z, 0, Qm, 0, 0, H.LH_CRC24B) for i in range(n)]
    ctx.gen_codeblocks(gen, d_cw, d_msg)
    H.hip().hipDeviceSynchronize()
    cbs = (H.PuschCbDesc * n)()
    for i in range(n):
        cbs[i] = H.PuschCbDesc(H.RmDesc(i * E, i * N, E, bg, Z, 0, Qm, 0, 0), i * mb, 24, H.LH_CRC24B)
    outs = [(H.DeviceBuffer(n * N), H.DeviceBuffer(n * mb), H.DeviceBuffer(4 * n)) for _ in range(3)]
    for o in outs:
        o[1].zero()
    streams = [H.Stream(), H.Stream()]

    def get(o):
        return o[2].download(np.int32, n), o[1].download(np.uint8, n * mb), o[0].download(np.int8, n * N)

    # Alone: one launch, waited for.
    ctx.pusch_decode_batch(None, d_cw, *outs[0], nof_ldpc_iterations=8, stream=streams[0], arr=cbs)
    streams[0].sync()
    res0, bits0, rm0 = get(outs[0])
    assert (res0 > 0).any() and (res0 <= 0).any(), "the batch should hold passing and failing codeblocks"
    # Overlapped: six launches alternating over the two streams, no wait in between.
    for k in range(6):
        ctx.pusch_decode_batch(None, d_cw, *outs[1 + k % 2], nof_ldpc_iterations=8, stream=streams[k % 2], arr=cbs)
    H.hip().hipDeviceSynchronize()
    for j in (1, 2):
        res, bits, rm = get(outs[j])
        assert np.array_equal(res, res0), f"stream {j - 1}: results differ from the launch alone"
        assert np.array_equal(bits, bits0), f"stream {j - 1}: hard bits differ from the launch alone"
        assert np.array_equal(rm, rm0), f"stream {j - 1}: HARQ buffers differ from the launch alone"
    for i in (0, 1, 2, 3, 600, 1215):
        r_o, b_o = O.decode(rm0[i * N:(i + 1) * N], bg, Z, crc_poly=O.CRC24B, max_iter=8)
        assert res0[i] == r_o, f"cb {i}: iterations {res0[i]} vs oracle {r_o}"
        assert np.array_equal(bits0[i * mb:(i + 1) * mb], b_o), f"cb {i}: bits differ from the oracle"
    ctx.close()
