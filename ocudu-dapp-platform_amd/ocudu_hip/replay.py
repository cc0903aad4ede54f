"""Offline PUSCH burst replay on the device: one recorded burst (the equalized data
symbols and grant of the E3 pusch.symbols stream) through soft demapping, descrambling,
UL-SCH demultiplexing, rate dematching, LDPC decoding and the transport-block CRC, with
the verdict printed beside the one the gNB reported.

Mirrors apps/tools/pusch_replay/ocudu_pusch_replay.cpp:
  * load (:85-137): the burst JSON written by the SDK's export_pusch_burst.py and its
    complex64 sample file (a bare file name is taken relative to the JSON);
  * run (:139-300): UL-SCH bit accounting with get_ulsch_information at a target code rate
    of 0 (the tool's "not needed for bit accounting", which selects BG2 for the codeblock
    sizes), N_ref from the LBRM TBS, a flat noise variance 10^(-SINR/10), scrambling
    c_init = RNTI * 2^15 + n_ID, UL-SCH demultiplexing without CSI Part 2, a new-data
    decode of the grant's RV with the CRC early stop and `iterations` LDPC iterations;
  * the output lines (:301-335).

The device path: lh_pusch_demodulate -> lh_ulsch_demultiplex -> the serving runtime
(lh_server_*: segmentation, HARQ buffer, dematching, decoding, TB CRC).  The packed
store form of the same record (OTE3 header version 2 + pusch_symbols_metadata_v2 + the
92-byte grant block, include/ocudu/dapp/e3/telemetry_types.h:61-168) is read by
parse_store_record.

    python -m ocudu_hip.replay burst.json [--iterations N] [--verbose] [--code-rate R]
"""
import argparse
import json
import struct
import sys
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import (LH_TB_CRC_OK, Context, DemodDesc, DeviceBuffer, Server, TbRequest, UlschDemuxDesc, ulsch_demux_sizes,
               ulsch_information)

MAX_NSYMB_PER_SLOT = 14
MAX_CODEBLOCK_SIZE = 384 * 66


@dataclass
class Burst:
    rnti: int
    numerology: int
    sfn: int
    slot: int
    modulation_bits: int
    sinr_db: float
    nof_prb: int
    nof_layers: int
    crc_known: bool
    crc_ok_reported: bool
    symbols_truncated: bool
    grant: dict
    symbols: np.ndarray = field(repr=False)  # complex64, codeword order


@dataclass
class ReplayResult:
    crc_ok: bool
    tb: np.ndarray          # TB bytes (meaningful when crc_ok)
    nof_ul_sch_bits: int
    nof_codeblocks: int
    nof_codeblocks_ok: int
    max_iterations: int
    lines: list             # the tool's report lines


def _number_or(obj, key, fallback):
    v = obj.get(key)
    return float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else fallback


def load_json(path):
    """A burst from the exporter's JSON and its complex64 sample file."""
    path = Path(path)
    root = json.loads(path.read_text())
    slot = root["radio_slot"]
    samples = root["samples_file"]
    spath = Path(samples)
    if "/" not in samples:
        spath = path.parent / samples
    raw = spath.read_bytes()
    if len(raw) % 8 != 0:
        raise ValueError("sample file is not whole complex64 values")
    crc = root.get("crc_ok")
    trunc = root.get("symbols_truncated")
    return Burst(rnti=int(root["rnti"]), numerology=int(slot["numerology"]), sfn=int(slot["sfn"]),
                 slot=int(slot["slot"]), modulation_bits=int(root["modulation_bits"]),
                 sinr_db=_number_or(root, "sinr_db", 20.0), nof_prb=int(root["nof_prb"]),
                 nof_layers=int(root["nof_layers"]), crc_known=crc is not None, crc_ok_reported=bool(crc),
                 symbols_truncated=isinstance(trunc, bool) and trunc, grant=dict(root["grant"]),
                 symbols=np.frombuffer(raw, np.complex64).copy())


# Packed store layouts (little endian, no padding).
_HEADER = struct.Struct("<IHHQ")                     # magic, version, kind, monotonic_ns
_SYMBOLS_V1 = struct.Struct("<HHBBHBBBB7f")          # rnti, sfn, slot, numerology, nof_prb, nof_data_symbols,
                                                     # nof_layers, modulation_bits, reserved, 7 floats
_GRANT = struct.Struct("<BBBBHHHHHHHHHII10B4f35sB")  # pusch_grant_metadata_v1
OTE3_MAGIC = 0x3345544F
KIND_PUSCH_SYMBOLS = 1


def parse_store_record(meta, tensor):
    """A burst from a packed pusch.symbols store record (header version 2, 152 bytes) and
    its tensor (complex64 equalized symbols)."""
    meta = bytes(meta)
    if len(meta) < 152:
        raise ValueError("pusch.symbols record shorter than 152 bytes (version 2)")
    magic, version, kind, _ = _HEADER.unpack_from(meta, 0)
    if magic != OTE3_MAGIC or kind != KIND_PUSCH_SYMBOLS or version != 2:
        raise ValueError(f"not a version-2 pusch.symbols record (magic {magic:#x}, version {version}, kind {kind})")
    (rnti, sfn, slot, numerology, nof_prb, _nsym, nof_layers, mod_bits, _r, sinr_db, _epre, _rsrp, _evm, _ta, _cfo,
     tcr) = _SYMBOLS_V1.unpack_from(meta, 16)
    flags = meta[56]
    if not flags & 8:
        raise ValueError("record without a valid grant block")
    g = _GRANT.unpack_from(meta, 60)
    (harq_id, rv, new_data, bg, n_id, bwp_start, bwp_size, dmrs_mask, dmrs_scr, dc_pos, n_ack, n_csi1, n_rs_id,
     tbs_bytes, tbs_lbrm) = g[:15]
    (gflags, mcs_index, mcs_table, start_symbol, nof_symbols, dmrs_type, cdm, nof_rx_ports, g_layers,
     cp) = g[15:25]
    alpha, b_ack, b_csi1, b_csi2 = g[25:29]
    prb_mask = g[29]
    grant = dict(harq_id=harq_id, rv=rv, new_data=new_data, ldpc_base_graph=bg, n_id=n_id, bwp_start_rb=bwp_start,
                 bwp_size_rb=bwp_size, dmrs_symbol_mask=dmrs_mask, dmrs_scrambling_id=dmrs_scr,
                 dc_position=dc_pos if gflags & 8 else None, nof_harq_ack_bits=n_ack, nof_csi_part1_bits=n_csi1,
                 n_rs_id=n_rs_id, tbs_bytes=tbs_bytes, tbs_lbrm_bytes=tbs_lbrm,
                 mcs_index=mcs_index if gflags & 1 else None, mcs_table=mcs_table if gflags & 1 else None,
                 transform_precoding=bool(gflags & 2), n_scid=int(bool(gflags & 4)),
                 csi_part2_expected=bool(gflags & 16), start_symbol=start_symbol, nof_symbols=nof_symbols,
                 dmrs_type=dmrs_type, nof_cdm_groups_without_data=cdm, nof_rx_ports=nof_rx_ports,
                 nof_layers=g_layers, cyclic_prefix=cp, alpha_scaling=alpha, beta_offset_harq_ack=b_ack,
                 beta_offset_csi_part1=b_csi1, beta_offset_csi_part2=b_csi2, prb_mask=prb_mask.hex(),
                 target_code_rate=tcr)
    sym = np.frombuffer(bytes(tensor), np.complex64).copy() if not isinstance(tensor, np.ndarray) else \
        np.ascontiguousarray(tensor, np.complex64).reshape(-1)
    return Burst(rnti=rnti, numerology=numerology, sfn=sfn, slot=slot, modulation_bits=mod_bits, sinr_db=sinr_db,
                 nof_prb=nof_prb, nof_layers=nof_layers, crc_known=bool(flags & 1), crc_ok_reported=bool(flags & 2),
                 symbols_truncated=bool(flags & 4), grant=grant, symbols=sym)


def noise_variance(sinr_db):
    """pow(10, -SINR / 10) in float, as the tool computes it."""
    return np.power(np.float32(10.0), -np.float32(sinr_db) / np.float32(10.0), dtype=np.float32)


def _contains_dc(g):
    """The DC subcarrier's PRB is allocated (ocudu_pusch_replay.cpp:177-184): byte
    dc_prb / 8 of the hex CRB mask, bit dc_prb % 8."""
    dc = g.get("dc_position")
    if dc is None:
        return 0
    prb = int(dc) // 12
    mask = g["prb_mask"]
    return (int(mask[2 * (prb // 8):2 * (prb // 8) + 2], 16) >> (prb % 8)) & 1


def nof_codeblocks(tbs_bits, bg):
    """compute_nof_codeblocks (include/ocudu/ran/sch/sch_segmentation.h:36-47)."""
    b = tbs_bits + (16 if tbs_bits <= 3824 else 24)
    k = 8448 if bg == 1 else 3840
    return 1 if b <= k else -(-b // (k - 24))


def grant_accounting(burst, code_rate=0.0):
    """(ulsch_information dict, demultiplexer descriptor fields) of a burst's grant."""
    g = burst.grant
    info = ulsch_information(
        tbs_bits=int(g["tbs_bytes"]) * 8, target_code_rate=code_rate, modulation=burst.modulation_bits,
        nof_harq_ack_bits=int(g["nof_harq_ack_bits"]), nof_csi_part1_bits=int(g["nof_csi_part1_bits"]),
        nof_csi_part2_bits=0, alpha_scaling=float(g["alpha_scaling"]),
        beta_offset_harq_ack=float(g["beta_offset_harq_ack"]), beta_offset_csi_part1=float(g["beta_offset_csi_part1"]),
        beta_offset_csi_part2=float(g["beta_offset_csi_part2"]), nof_rb=burst.nof_prb,
        start_symbol_index=int(g["start_symbol"]), nof_symbols=int(g["nof_symbols"]),
        dmrs_type=2 if int(g["dmrs_type"]) == 2 else 1, dmrs_symbol_mask=int(g["dmrs_symbol_mask"]),
        nof_cdm_groups_without_data=int(g["nof_cdm_groups_without_data"]), nof_layers=burst.nof_layers,
        contains_dc=_contains_dc(g))
    demux = dict(modulation=burst.modulation_bits, nof_layers=burst.nof_layers, nof_prb=burst.nof_prb,
                 start_symbol_index=int(g["start_symbol"]), nof_symbols=int(g["nof_symbols"]),
                 dmrs_symbol_mask=int(g["dmrs_symbol_mask"]), dmrs_type=2 if int(g["dmrs_type"]) == 2 else 1,
                 nof_cdm_groups_without_data=int(g["nof_cdm_groups_without_data"]),
                 nof_harq_ack_rvd=info["nof_harq_ack_rvd"], nof_harq_ack_bits=int(g["nof_harq_ack_bits"]),
                 nof_enc_harq_ack_bits=info["nof_harq_ack_bits"], nof_csi_part1_bits=int(g["nof_csi_part1_bits"]),
                 nof_enc_csi_part1_bits=info["nof_csi_part1_bits"],
                 c_init=(burst.rnti << 15) + int(g["n_id"]))
    return info, demux


class Replayer:
    """Device buffers and a serving runtime reused over bursts."""

    def __init__(self, ctx=None, iterations=10, max_cbs=512):
        self.ctx = ctx if ctx is not None else Context(0, max_codeblocks=max(max_cbs, 64))
        self.own_ctx = ctx is None
        self.iterations = iterations
        # One HARQ process of up to max_cbs BG1 Z=384 codeblocks.
        self.server = Server(self.ctx, harq_budget_bytes=max_cbs * 66 * 384 + (1 << 20), max_cbs_per_slot=max_cbs,
                             nof_ldpc_iterations=iterations)
        self._bufs = {}

    def _buf(self, name, nbytes):
        b = self._bufs.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = DeviceBuffer(max(nbytes, 64))
            self._bufs[name] = b
        return b

    def run(self, burst, code_rate=0.0):
        g = burst.grant
        qm = burst.modulation_bits
        info, dm = grant_accounting(burst, code_rate)
        desc = UlschDemuxDesc(**{k: v for k, v in dm.items()})
        sizes = ulsch_demux_sizes(desc)
        nof_re_symbol = burst.nof_prb * 12
        n_sym = burst.symbols.size
        if n_sym % (nof_re_symbol * burst.nof_layers) != 0:
            raise ValueError(f"sample count {n_sym} is not a whole number of OFDM symbols")
        if n_sym * qm != sizes["nof_llrs"]:
            raise ValueError(f"{n_sym} symbols x {qm} bits do not fill the grant's {sizes['nof_llrs']} soft bits")
        bg = 2 if int(g["ldpc_base_graph"]) == 2 else 1
        tbs_bits = int(g["tbs_bytes"]) * 8
        ncb = nof_codeblocks(tbs_bits, bg)
        n_ref = min(int(g["tbs_lbrm_bytes"]) * 8 * 3 // (2 * ncb), MAX_CODEBLOCK_SIZE)
        # Demapping with a flat noise variance, then descrambling.
        d_sym = self._buf("sym", n_sym * 8)
        d_sym.upload(burst.symbols.view(np.float32))
        d_nv = self._buf("nv", n_sym * 4)
        d_nv.upload(np.full(n_sym, noise_variance(burst.sinr_db), np.float32))
        d_llr = self._buf("llr", sizes["nof_llrs"])
        self.ctx.pusch_demodulate([DemodDesc(0, 0, n_sym, qm, dm["c_init"], 0)], d_sym, d_nv, d_llr)
        # UL-SCH demultiplexing; UCI soft bits are not decoded here.
        d_sch = self._buf("sch", sizes["nof_sch"])
        d_ack = self._buf("ack", sizes["nof_harq_ack"])
        d_c1 = self._buf("c1", sizes["nof_csi_part1"])
        self.ctx.ulsch_demultiplex([desc], d_llr, d_sch, d_ack, d_c1, None)
        # Decoding through the serving runtime: one new-data transmission.
        harq = int(g.get("harq_id", 0))
        req = TbRequest(id=1, d_llrs=d_sch.ptr, harq_process=harq, new_data=1, base_graph=bg, tbs_bits=tbs_bits,
                        modulation_order=qm, nof_layers=burst.nof_layers, nof_ch_symbols=info["nof_ul_sch_bits"] // qm,
                        rv=int(g["rv"]), Nref=n_ref, slot=0)
        tb = np.zeros(tbs_bits // 8, np.uint8)
        self.server.submit(req, tb.ctypes.data)
        self.server.step()
        res = []
        while not res:
            res = self.server.poll(wait=True)
        r = res[0]
        self.server.harq_release(harq)
        ok = r.tb_crc_ok == LH_TB_CRC_OK
        lines = [
            f"rnti=0x{burst.rnti:04x} sfn={burst.sfn} slot={burst.slot} tbs={tbs_bits // 8} B "
            f"mcs={_mcs_text(g)} bg={int(g['ldpc_base_graph'])} rv={int(g['rv'])} prb={burst.nof_prb} Qm={qm} "
            f"layers={burst.nof_layers} ulsch_bits={info['nof_ul_sch_bits']} codeblocks={ncb}",
            "replay CRC: {} | gNB CRC: {}{}{}".format(
                "OK" if ok else "FAIL",
                ("OK" if burst.crc_ok_reported else "FAIL") if burst.crc_known else "unknown",
                " | capture incomplete" if burst.symbols_truncated else "",
                " | MISMATCH" if burst.crc_known and ok != burst.crc_ok_reported else "")]
        return ReplayResult(crc_ok=ok, tb=tb, nof_ul_sch_bits=info["nof_ul_sch_bits"], nof_codeblocks=r.nof_cbs,
                            nof_codeblocks_ok=r.nof_cbs_ok, max_iterations=r.max_iterations, lines=lines)

    def close(self):
        for b in self._bufs.values():
            b.free()
        self._bufs = {}
        self.server.close()
        if self.own_ctx:
            self.ctx.close()


def _mcs_text(g):
    if g.get("mcs_index") is None:
        return "unknown"
    t = g.get("mcs_table")
    return f"{int(g['mcs_index'])}/table{int(t) if isinstance(t, (int, float)) else 0}"


def main(argv=None):
    ap = argparse.ArgumentParser(description="Replays a PUSCH burst captured over E3 through the device receive "
                                             "chain and compares the transport-block CRC with the recorded verdict.")
    ap.add_argument("burst", help="burst JSON exported by export_pusch_burst.py")
    ap.add_argument("--iterations", type=int, default=10, help="LDPC decoder iteration limit (1..10000)")
    ap.add_argument("--verbose", action="store_true", help="print LDPC statistics and the first TB bytes")
    ap.add_argument("--code-rate", type=float, default=0.0,
                    help="target code rate for the UL-SCH bit accounting (0: as the reference tool)")
    a = ap.parse_args(argv)
    if not 1 <= a.iterations <= 10000:
        ap.error("--iterations out of range")
    try:
        burst = load_json(a.burst)
        rp = Replayer(iterations=a.iterations)
        try:
            r = rp.run(burst, a.code_rate)
        finally:
            rp.close()
    except Exception as e:  # the tool's "error: ..." and status 2
        print(f"error: {e}", file=sys.stderr)
        return 2
    for line in r.lines:
        print(line)
    if a.verbose:
        print(f"ldpc iterations: max {r.max_iterations} over {r.nof_codeblocks} codeblocks")
        if r.crc_ok:
            print("transport block (first 16 bytes): " + " ".join(f"{b:02x}" for b in r.tb[:16]))
    return 0 if r.crc_ok else 1


if __name__ == "__main__":
    sys.exit(main())
