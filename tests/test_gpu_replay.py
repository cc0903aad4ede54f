"""Burst replay on the device (ocudu_hip.replay: lh_pusch_demodulate ->
lh_ulsch_demultiplex -> the serving runtime) against the oracle replay chain
(tests/burst_util.replay_oracle): the UL-SCH soft bits bit-exact, the same CRC verdict
and TB bytes, which are the TB the synthetic UE sent; from the JSON + complex64 pair
and from the packed store record; a burst that fails at low SINR fails on both; the
command line prints the tool's lines and exit status."""
import numpy as np
import pytest

import burst_util as B
from gpu_util import require_gpu
from ocudu_hip import replay as R

pytestmark = pytest.mark.gpu

CASES = [
    (dict(), 4, 1, 20, 18.0),
    (dict(nof_harq_ack_bits=1, nof_csi_part1_bits=9), 6, 2, 30, 24.0),
    (dict(ldpc_base_graph=2, tbs_bytes=300, nof_harq_ack_bits=5), 2, 1, 10, 8.0),
    (dict(tbs_bytes=20000, nof_harq_ack_bits=2, tbs_lbrm_bytes=25000), 8, 4, 100, 30.0),
    (dict(nof_csi_part1_bits=30, nof_harq_ack_bits=20, start_symbol=2, nof_symbols=12,
          dmrs_symbol_mask=(1 << 2) | (1 << 7), dmrs_type=2, nof_cdm_groups_without_data=3), 4, 2, 16, 20.0),
    (dict(tbs_bytes=38000, rv=0, nof_harq_ack_bits=2, nof_csi_part1_bits=5), 8, 4, 273, 30.0),
    (dict(ldpc_base_graph=2, tbs_bytes=2000, rv=3, nof_csi_part1_bits=2), 2, 1, 100, 6.0),
    (dict(), 4, 1, 20, 4.0),  # below threshold: CRC fails
]


@pytest.fixture(scope="module")
def replayer():
    require_gpu()
    r = R.Replayer(iterations=10)
    yield r
    r.close()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_replay_matches_oracle(replayer, i, tmp_path):
    kw, qm, layers, prb, snr = CASES[i]
    rng = np.random.default_rng(200 + i)
    burst, sym, tb = B.make_burst(rng, B.base_grant(**kw), qm, layers, prb, snr)
    print(f"case replay {i}: {kw} Qm={qm} layers={layers} prb={prb} sinr={snr}", flush=True)
    ok_o, tb_o, sch_o = B.replay_oracle(burst, sym)
    b = R.load_json(B.write_json(tmp_path, f"b{i}", burst, sym))
    r = replayer.run(b)
    sch = replayer._bufs["sch"].download(np.int8, sch_o.size)
    assert np.array_equal(sch, sch_o), "UL-SCH soft bits differ from the oracle"
    assert r.crc_ok == ok_o
    assert r.nof_ul_sch_bits == sch_o.size
    if ok_o:
        assert np.array_equal(r.tb, tb_o) and np.array_equal(r.tb, tb)
    # The packed store record replays alike.
    meta, tensor = B.store_record(burst, sym)
    r2 = replayer.run(R.parse_store_record(meta, tensor))
    assert r2.crc_ok == r.crc_ok and (not r.crc_ok or np.array_equal(r2.tb, r.tb))
    assert r.lines[1].startswith("replay CRC: " + ("OK" if ok_o else "FAIL") + " | gNB CRC: OK")
    assert ("MISMATCH" in r.lines[1]) == (not ok_o)


def test_command_line(tmp_path, capsys):
    require_gpu()
    rng = np.random.default_rng(300)
    burst, sym, tb = B.make_burst(rng, B.base_grant(nof_harq_ack_bits=1), 4, 1, 20, 18.0)
    path = B.write_json(tmp_path, "cli", burst, sym)
    assert R.main([str(path), "--verbose"]) == 0
    out = capsys.readouterr().out.splitlines()
    assert out[0].startswith("rnti=0x4601 sfn=123 slot=7 tbs=1000 B mcs=20/table1 bg=1 rv=0 prb=20 Qm=4 layers=1")
    assert out[1] == "replay CRC: OK | gNB CRC: OK"
    assert out[3] == "transport block (first 16 bytes): " + " ".join(f"{v:02x}" for v in tb[:16])
    burst["nof_prb"] = 21  # samples no longer a whole number of OFDM symbols
    path = B.write_json(tmp_path, "bad", burst, sym)
    assert R.main([str(path)]) == 2
