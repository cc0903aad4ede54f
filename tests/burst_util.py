"""Synthetic PUSCH bursts in the E3 replay formats (test infrastructure): a UE-side
transmitter built on the oracle -- TX segmentation with TB and codeblock CRCs
(tests/tb_util.py), LDPC encoding and rate matching (oracle), the UL-SCH bit accounting
the replay tool uses (get_ulsch_information at code rate 0), UL-SCH / UCI multiplexing as
the inverse of the oracle's demultiplexer positions (a 1- or 2-bit HARQ-ACK overwriting
the bits of its REs), scrambling with c_init = RNTI * 2^15 + n_ID, Gray QAM mapping
(TS38.211 5.1), AWGN at the reported SINR -- written as the exporter's JSON + complex64
pair or as a packed version-2 pusch.symbols store record.  The oracle replay chain
(replay_oracle) is the checker of the device replay."""
import json
import struct
from pathlib import Path

import numpy as np

import oracle as O
from np_datagen import modulate
from serve_oracle import HarqModel
from tb_util import tb_codeblocks

MOD_NORM = {2: 1 / np.sqrt(2), 4: 1 / np.sqrt(10), 6: 1 / np.sqrt(42), 8: 1 / np.sqrt(170)}


def accounting(grant, qm, layers, nof_prb):
    cfg = dict(tbs_bits=grant["tbs_bytes"] * 8, target_code_rate=0.0, modulation=qm,
               nof_harq_ack_bits=grant["nof_harq_ack_bits"], nof_csi_part1_bits=grant["nof_csi_part1_bits"],
               nof_csi_part2_bits=0, alpha_scaling=grant["alpha_scaling"],
               beta_offset_harq_ack=grant["beta_offset_harq_ack"],
               beta_offset_csi_part1=grant["beta_offset_csi_part1"],
               beta_offset_csi_part2=grant["beta_offset_csi_part2"], nof_rb=nof_prb,
               start_symbol_index=grant["start_symbol"], nof_symbols=grant["nof_symbols"],
               dmrs_type=grant["dmrs_type"], dmrs_symbol_mask=grant["dmrs_symbol_mask"],
               nof_cdm_groups_without_data=grant["nof_cdm_groups_without_data"], nof_layers=layers, contains_dc=0)
    info = O.ulsch_information(cfg)
    demux = dict(modulation=qm, nof_layers=layers, nof_prb=nof_prb, start_symbol_index=grant["start_symbol"],
                 nof_symbols=grant["nof_symbols"], dmrs_symbol_mask=grant["dmrs_symbol_mask"],
                 dmrs_type=grant["dmrs_type"], nof_cdm_groups_without_data=grant["nof_cdm_groups_without_data"],
                 nof_harq_ack_rvd=info["nof_harq_ack_rvd"], nof_harq_ack_bits=grant["nof_harq_ack_bits"],
                 nof_enc_harq_ack_bits=info["nof_harq_ack_bits"], nof_csi_part1_bits=grant["nof_csi_part1_bits"],
                 nof_enc_csi_part1_bits=info["nof_csi_part1_bits"])
    return info, demux


def n_ref(grant, bg):
    return O.compute_n_ref(grant["tbs_lbrm_bytes"], O.nof_codeblocks(grant["tbs_bytes"] * 8, bg))


def make_burst(rng, grant, qm, layers, nof_prb, sinr_db, rnti=0x4601):
    """(burst dict for the JSON, complex64 symbols, TB bytes)."""
    bg = grant["ldpc_base_graph"]
    info, demux = accounting(grant, qm, layers, nof_prb)
    G = info["nof_ul_sch_bits"]
    tb = rng.integers(0, 256, grant["tbs_bytes"]).astype(np.uint8)
    cbs, seg = O.segment_rx(bg, grant["tbs_bytes"], qm, layers, G // qm)
    msgs = tb_codeblocks(tb, seg)
    nref = n_ref(grant, bg)
    sch = np.zeros(G, np.uint8)
    for c, m in enumerate(cbs):
        d = O.encode(msgs[c], bg, m["Z"])[2 * m["Z"]:]  # shortened by the 2Z punctured bits
        e = O.rate_match(d, m["rm_length"], bg, m["Z"], rv=grant["rv"], Qm=qm, Nref=nref, F=m["nof_filler_bits"])
        sch[m["cw_offset"]:m["cw_offset"] + m["rm_length"]] = e
    pos = O.ulsch_positions(demux)
    n = O.demux_sizes(demux)["nof_llrs"]
    cw = np.zeros(n, np.uint8)
    cw[pos[0]] = sch
    cw[pos[2]] = rng.integers(0, 2, pos[2].size)
    cw[pos[1]] = rng.integers(0, 2, pos[1].size)  # HARQ-ACK last: punctures UL-SCH REs of 1-2 bits
    c_init = (rnti << 15) + grant["n_id"]
    scr = cw ^ O.scrambling_sequence(c_init, n)
    iq = modulate(scr.reshape(1, -1), qm).reshape(-1, 2).astype(np.float64)
    n0 = 10.0 ** (-sinr_db / 10.0)
    iq += rng.standard_normal(iq.shape) * np.sqrt(n0 / 2)
    sym = (iq[:, 0] + 1j * iq[:, 1]).astype(np.complex64)
    burst = {"rnti": rnti, "radio_slot": {"numerology": 1, "sfn": 123, "slot": 7}, "modulation_bits": qm,
             "sinr_db": sinr_db, "nof_prb": nof_prb, "nof_layers": layers, "crc_ok": True,
             "symbols_truncated": False, "grant": grant}
    return burst, sym, tb


def write_json(directory, name, burst, sym):
    directory = Path(directory)
    (directory / f"{name}.cf32").write_bytes(sym.tobytes())
    b = dict(burst, samples_file=f"{name}.cf32")
    (directory / f"{name}.json").write_text(json.dumps(b))
    return directory / f"{name}.json"


def store_record(burst, sym, target_code_rate=0.0):
    """The packed store form: OTE3 header version 2, pusch_symbols_metadata_v2."""
    g = burst["grant"]
    head = struct.pack("<IHHQ", 0x3345544F, 2, 1, 123456789)
    nsym = g["nof_symbols"]
    v1 = struct.pack("<HHBBHBBBB7f", burst["rnti"], burst["radio_slot"]["sfn"], burst["radio_slot"]["slot"],
                     burst["radio_slot"]["numerology"], burst["nof_prb"], nsym, burst["nof_layers"],
                     burst["modulation_bits"], 0, burst["sinr_db"], -10.0, -12.0, 0.05, 0.1, 20.0, target_code_rate)
    flags = 1 | (2 if burst["crc_ok"] else 0) | (4 if burst["symbols_truncated"] else 0) | 8
    gflags = (1 if g.get("mcs_index") is not None else 0) | (8 if g.get("dc_position") is not None else 0)
    prb = bytes.fromhex(g["prb_mask"]).ljust(35, b"\0")
    grant = struct.pack("<BBBBHHHHHHHHHII10B4f35sB", g["harq_id"], g["rv"], 1, g["ldpc_base_graph"], g["n_id"], 0,
                        273, g["dmrs_symbol_mask"], 0, g.get("dc_position") or 0, g["nof_harq_ack_bits"],
                        g["nof_csi_part1_bits"], 0, g["tbs_bytes"], g["tbs_lbrm_bytes"], gflags,
                        g.get("mcs_index") or 0, g.get("mcs_table") or 0, g["start_symbol"], nsym, g["dmrs_type"],
                        g["nof_cdm_groups_without_data"], 4, burst["nof_layers"], 0, g["alpha_scaling"],
                        g["beta_offset_harq_ack"], g["beta_offset_csi_part1"], g["beta_offset_csi_part2"], prb, 0)
    meta = head + v1 + bytes([flags, 0, 0, 0]) + grant
    assert len(meta) == 152
    return meta, sym.tobytes()


def replay_oracle(burst, sym, iterations=10):
    """The replay tool's chain on the oracle: demapping + descrambling
    (orc_pusch_demodulate), demultiplexing (orc_ulsch_demultiplex), the PUSCH decoder's
    new-data decode of one transmission (HarqModel).  Returns (crc_ok, TB bytes or None,
    UL-SCH soft bits)."""
    g = burst["grant"]
    qm, layers, prb = burst["modulation_bits"], burst["nof_layers"], burst["nof_prb"]
    info, demux = accounting(g, qm, layers, prb)
    c_init = (burst["rnti"] << 15) + g["n_id"]
    nv = np.float32(10.0) ** (-np.float32(burst["sinr_db"]) / np.float32(10.0))
    iq = np.ascontiguousarray(sym, np.complex64).view(np.float32)
    llr = O.pusch_demodulate(iq, np.full(sym.size, nv, np.float32), qm, c_init)
    sch = O.ulsch_demultiplex(demux, llr, c_init)[0]
    bg = g["ldpc_base_graph"]
    hm = HarqModel(bg, g["tbs_bytes"] * 8, qm, layers, info["nof_ul_sch_bits"] // qm, Nref=n_ref(g, bg))
    verdict, _, _, tb = hm.transmit(sch, g["rv"], True, iterations)
    return verdict == 1, tb, sch


def base_grant(**kw):
    g = dict(harq_id=3, rv=0, n_id=500, ldpc_base_graph=1, tbs_bytes=1000, tbs_lbrm_bytes=100000,
             nof_harq_ack_bits=0, nof_csi_part1_bits=0, alpha_scaling=1.0, beta_offset_harq_ack=8.0,
             beta_offset_csi_part1=2.0, beta_offset_csi_part2=2.0, start_symbol=0, nof_symbols=14, dmrs_type=1,
             dmrs_symbol_mask=(1 << 2) | (1 << 11), nof_cdm_groups_without_data=2, dc_position=None,
             prb_mask="ff" * 35, mcs_index=20, mcs_table=1)
    g.update(kw)
    return g
