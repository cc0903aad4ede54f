"""The burst replay's host side without a GPU: the exporter's JSON + complex64 pair and
the packed version-2 pusch.symbols store record parse to the same burst
(ocudu_hip.replay.load_json / parse_store_record, record layout of
include/ocudu/dapp/e3/telemetry_types.h:61-168), malformed records are refused, the
grant's UL-SCH accounting through the C ABI equals the oracle's, and the oracle replay
chain (demapping, descrambling, demultiplexing, decoding, TB CRC) recovers the TB the
synthetic UE sent -- with and without UCI, BG1 and BG2, one to four layers, limited
buffer -- and fails at low SINR."""
import numpy as np
import pytest

import burst_util as B
from ocudu_hip import replay as R

CASES = [
    (dict(), 4, 1, 20, 18.0),
    (dict(nof_harq_ack_bits=1, nof_csi_part1_bits=9), 6, 2, 30, 24.0),
    (dict(ldpc_base_graph=2, tbs_bytes=300, nof_harq_ack_bits=5, rv=0), 2, 1, 10, 8.0),
    (dict(tbs_bytes=20000, nof_harq_ack_bits=2, tbs_lbrm_bytes=25000), 8, 4, 100, 30.0),
    (dict(nof_csi_part1_bits=30, nof_harq_ack_bits=20, start_symbol=2, nof_symbols=12,
          dmrs_symbol_mask=(1 << 2) | (1 << 7), dmrs_type=2, nof_cdm_groups_without_data=3), 4, 2, 16, 20.0),
]


def _burst(i, sinr_off=0.0):
    kw, qm, layers, prb, snr = CASES[i]
    rng = np.random.default_rng(100 + i)
    return B.make_burst(rng, B.base_grant(**kw), qm, layers, prb, snr + sinr_off)


def test_json_and_store_record_parse_alike(tmp_path):
    burst, sym, _ = _burst(1)
    a = R.load_json(B.write_json(tmp_path, "b1", burst, sym))
    meta, tensor = B.store_record(burst, sym)
    b = R.parse_store_record(meta, tensor)
    for k in ("rnti", "sfn", "slot", "numerology", "modulation_bits", "nof_prb", "nof_layers", "crc_known",
              "crc_ok_reported", "symbols_truncated"):
        assert getattr(a, k) == getattr(b, k), k
    assert np.float32(a.sinr_db) == np.float32(b.sinr_db)
    assert np.array_equal(a.symbols, sym) and np.array_equal(b.symbols, sym)
    for k in ("harq_id", "rv", "n_id", "ldpc_base_graph", "tbs_bytes", "tbs_lbrm_bytes", "nof_harq_ack_bits",
              "nof_csi_part1_bits", "start_symbol", "nof_symbols", "dmrs_type", "dmrs_symbol_mask",
              "nof_cdm_groups_without_data", "mcs_index", "mcs_table"):
        assert a.grant[k] == b.grant[k], k
    for k in ("alpha_scaling", "beta_offset_harq_ack", "beta_offset_csi_part1", "beta_offset_csi_part2"):
        assert np.float32(a.grant[k]) == np.float32(b.grant[k]), k


def test_store_record_refusals():
    burst, sym, _ = _burst(0)
    meta, tensor = B.store_record(burst, sym)
    with pytest.raises(ValueError, match="152"):
        R.parse_store_record(meta[:56], tensor)
    bad = bytearray(meta)
    bad[4] = 1  # version 1
    with pytest.raises(ValueError, match="version-2"):
        R.parse_store_record(bytes(bad), tensor)
    bad = bytearray(meta)
    bad[56] &= ~8  # grant not valid
    with pytest.raises(ValueError, match="grant"):
        R.parse_store_record(bytes(bad), tensor)


def test_dc_position_lookup():
    g = dict(dc_position=12 * 9 + 5, prb_mask="00" + "02" + "00" * 33)  # PRB 9: byte 1, bit 1
    assert R._contains_dc(g) == 1
    assert R._contains_dc(dict(g, prb_mask="00" * 35)) == 0
    assert R._contains_dc(dict(g, dc_position=None)) == 0


@pytest.mark.parametrize("i", range(len(CASES)))
def test_grant_accounting_matches_oracle(i, tmp_path):
    burst, sym, _ = _burst(i)
    b = R.load_json(B.write_json(tmp_path, f"b{i}", burst, sym))
    info, demux = R.grant_accounting(b)
    want_info, want_demux = B.accounting(burst["grant"], burst["modulation_bits"], burst["nof_layers"],
                                         burst["nof_prb"])
    for k, v in want_info.items():
        assert info[k] == v, k
    for k, v in want_demux.items():
        assert demux[k] == v, k


@pytest.mark.parametrize("i", range(len(CASES)))
def test_oracle_replay_recovers_the_tb(i):
    burst, sym, tb = _burst(i)
    ok, got, _ = B.replay_oracle(burst, sym)
    assert ok and np.array_equal(got, tb)


def test_oracle_replay_fails_at_low_sinr():
    burst, sym, _ = _burst(0, sinr_off=-14.0)
    ok, got, _ = B.replay_oracle(burst, sym)
    assert not ok and got is None
