"""The oracle chain of one HARQ process as the reference's PUSCH decoder runs it, for the
serving-runtime tests: RX segmentation (oracle.segment_rx, ldpc_segmenter_rx_impl.cpp:33-85),
rate dematching into the HARQ soft buffer with combining on retransmissions
(ldpc_rate_dematcher_impl.cpp:29-250), decoding with CRC early stop of the codeblocks whose
CRC has not passed yet -- the others only combine their new LLRs and report the earlier
result (pusch_decoder_impl.cpp:412-431) -- and the TB verdict: codeblock CRC failures, or
the concatenation and TB CRC24A (concatenate_codeblocks, :734-781; with one codeblock its
CRC is the TB CRC), with every codeblock decoded again after a TB CRC failure
(:657-669)."""
import numpy as np

import oracle as O

TB_CRC_OK, TB_CB_FAILED, TB_CRC_FAILED = 1, 0, -1


class HarqModel:
    def __init__(self, bg, tbs_bits, qm, layers, nsym, Nref=0):
        self.bg, self.tbs, self.qm, self.Nref = bg, tbs_bits, qm, Nref
        self.cbs, self.seg = O.segment_rx(bg, tbs_bits // 8, qm, layers, nsym)
        self.C = len(self.cbs)
        self.Z = self.cbs[0]["Z"]
        self.F = self.cbs[0]["nof_filler_bits"]
        self.K = O.K_b(bg) * self.Z
        self.stride = (self.K + 7) // 8
        if self.C > 1:
            self.poly, self.crc_bits = O.CRC24B, 24
        else:
            self.poly, self.crc_bits = (O.CRC16, 16) if tbs_bits <= 3824 else (O.CRC24A, 24)
        self.soft = [None] * self.C
        self.bits = np.zeros((self.C, self.stride), np.uint8)
        self.passed = [0] * self.C

    def transmit(self, llrs, rv, new_data, iters):
        """One transmission: llrs = the codeword's LLRs (nof_ch_symbols * Qm).  Returns
        (verdict, codeblocks passed, max iterations over them, TB bytes or None)."""
        res = np.zeros(self.C, np.int32)
        for c, m in enumerate(self.cbs):
            e = np.asarray(llrs[m["cw_offset"]:m["cw_offset"] + m["rm_length"]], np.int8)
            self.soft[c] = O.rate_dematch(e, self.bg, self.Z, rv=rv, Qm=self.qm, Nref=self.Nref, F=self.F,
                                          new_data=new_data, harq=None if new_data else self.soft[c])
            if new_data:
                self.passed[c] = 0
            if self.passed[c]:
                res[c] = self.passed[c]
                continue
            r, b = O.decode(self.soft[c], self.bg, self.Z, F=self.F, crc_poly=self.poly, max_iter=iters,
                            crc_bits=self.crc_bits)
            res[c] = r
            if r:
                self.bits[c] = b
                self.passed[c] = r
        ok = int((res > 0).sum())
        its = int(res.max()) if ok else 0
        if ok < self.C:
            return TB_CB_FAILED, ok, its, None
        if self.C == 1:
            tb = np.packbits(np.unpackbits(self.bits[0])[:self.tbs])
            return TB_CRC_OK, ok, its, tb
        tb, chk = O.tb_concat_crc(self.bits.reshape(-1), self.stride, self.tbs // 8, self.seg)
        if chk != 0:
            self.passed = [0] * self.C
            return TB_CRC_FAILED, ok, its, None
        return TB_CRC_OK, ok, its, tb
