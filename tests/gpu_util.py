"""Helpers shared by the GPU tests: build batches, run them through the C ABI."""
import numpy as np

import ocudu_hip as H
import oracle as O


def require_gpu():
    if H.device_count() < 1:
        raise RuntimeError("GPU test run without a visible GPU")
    H.lib()  # raises when the library is not built


# Row modes of the decoder under test: "auto" (chosen from occupancy and load; the small
# test batches run one check row per thread) and "pairs" (two rows per thread forced for
# every even lifting size).
ROW_MODES = ["auto", "pairs"]


def make_ctx(max_codeblocks, rows="auto"):
    require_gpu()
    c = H.Context(0, max_codeblocks=max_codeblocks)
    if rows == "pairs":
        for bg in (1, 2):
            for Z in O.LIFTING_SIZES:
                if Z % 2 == 0:
                    c.tune_rows(bg, Z, 1)
    return c


def decode_batch_gpu(ctx, llr_list, cfgs, crc_polys, flags=0):
    """llr_list: list of int8 arrays (decoder input); cfgs: (bg, Z, F, crc_bits, max_iter)."""
    n = len(llr_list)
    offs = np.zeros(n + 1, np.int64)
    for i, a in enumerate(llr_list):
        offs[i + 1] = offs[i] + ((a.size + 15) // 16) * 16
    buf = np.zeros(int(offs[-1]) or 16, np.int8)
    for i, a in enumerate(llr_list):
        buf[offs[i]:offs[i] + a.size] = a
    bofs = np.zeros(n + 1, np.int64)
    for i, c in enumerate(cfgs):
        bofs[i + 1] = bofs[i] + ((H.msg_bytes(c[0], c[1]) + 15) // 16) * 16
    d_llr = H.DeviceBuffer.from_array(buf)
    d_bits = H.DeviceBuffer(int(bofs[-1]))
    d_bits.zero()
    d_res = H.DeviceBuffer(4 * n)
    descs = []
    for i, (bg, Z, F, cb, mi) in enumerate(cfgs):
        descs.append(H.DecodeDesc(int(offs[i]), int(bofs[i]), llr_list[i].size, crc_polys[i],
                                  H.DecoderConfig(bg, Z, F, cb, mi)))
    ctx.ldpc_decode(descs, d_llr, d_bits, d_res, flags=flags)
    res = d_res.download(np.int32, n)
    allbits = d_bits.download(np.uint8, int(bofs[-1]))
    bits = [allbits[bofs[i]:bofs[i] + H.msg_bytes(c[0], c[1])] for i, c in enumerate(cfgs)]
    return res, bits
