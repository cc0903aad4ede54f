"""The oracle's UL-SCH demultiplexer (oracle/ldpc_oracle.c, orc_ulsch_demultiplex) pinned
to the reference's own ulsch_demultiplex_impl built from source (oracle/_ref, fed in
blocks of random sizes, CSI Part 2 configured before the first block), to the committed
fixture written from that build (tests/golden/ulsch_demux.npz), and to what TS38.212 /
TS38.211 fix: a hand-worked allocation, the identity without UCI, the placeholder rules
of the transmitter (TS38.211 6.3.1.1: an "x" placeholder is sent as 1, a "y" placeholder
repeats the previous scrambled bit) undone by the receiver, and the configurations the
reference rejects.  The device kernel's arithmetic (lh::demux_plan, lh::demux_classify)
is replayed on the host against the oracle by tests/cpp/demux_rank_test."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from demux_util import config_words, random_config

GOLDEN = Path(__file__).resolve().parent / "golden" / "ulsch_demux.npz"
RANK_BIN = Path(__file__).resolve().parent / "cpp" / "demux_rank_test"


def _llrs(rng, n):
    return rng.integers(-120, 121, n).astype(np.int8)


@pytest.mark.skipif(O.ref_demod() is None, reason="reference demultiplexer not built (oracle/_ref)")
def test_oracle_matches_reference_build():
    rng = np.random.default_rng(2024)
    compared = 0
    for case in range(400):
        cfg = random_config(rng, max_prb=24 if case % 4 else 275)
        sz = O.demux_sizes(cfg)
        llr = _llrs(rng, sz["nof_llrs"])
        c_init = int(rng.integers(0, 2 ** 31))
        want = O.ref_ulsch_demultiplex(cfg, llr, c_init, block=int(rng.integers(1, 6000)))
        if want is None:
            continue
        compared += 1
        got = O.ulsch_demultiplex(cfg, llr, c_init)
        for k, name in enumerate(("sch", "harq_ack", "csi_part1", "csi_part2")):
            assert np.array_equal(got[k], want[k]), f"case {case} {cfg}: {name} differs from the reference"
    assert compared > 300


def test_oracle_matches_golden_fixture():
    g = np.load(GOLDEN)
    words, offs = g["config"], g["llr_off"]
    for i in range(len(words)):
        cfg = dict(zip(O.DEMUX_FIELDS, (int(v) for v in words[i][:15])))
        llr = g["llr"][offs[i]:offs[i + 1]]
        got = O.ulsch_demultiplex(cfg, llr, int(words[i][15]))
        for k, name in enumerate(("sch", "harq_ack", "csi_part1", "csi_part2")):
            o = g[f"{name}_off"]
            assert np.array_equal(got[k], g[name][o[i]:o[i + 1]]), f"fixture case {i}: {name}"


def test_hand_worked_allocation():
    """QPSK, 1 layer, 1 PRB, symbols 0-3, DM-RS on symbol 0 with both CDM groups without
    data (no data REs there), HARQ-ACK 3 bits on 8 REs, CSI Part 1 on 10 REs.  Symbol 1:
    M_uci = 12, HARQ-ACK spacing 12 // 8 = 1 -> REs 0..7; CSI Part 1 gets the 4 left.
    Symbol 2: CSI Part 1 remainder 6 < 12, spacing 2 -> REs 0, 2, .., 10; UL-SCH the odd
    REs.  Symbol 3: all UL-SCH."""
    cfg = dict(modulation=2, nof_layers=1, nof_prb=1, start_symbol_index=0, nof_symbols=4, dmrs_symbol_mask=1,
               dmrs_type=1, nof_cdm_groups_without_data=2, nof_harq_ack_bits=3, nof_enc_harq_ack_bits=16,
               nof_csi_part1_bits=5, nof_enc_csi_part1_bits=20)
    llr = (np.arange(72) - 36).astype(np.int8)
    sch, ack, c1, c2 = O.ulsch_demultiplex(cfg, llr, 77)
    re = llr.reshape(36, 2)  # 3 symbols x 12 REs
    assert np.array_equal(ack, re[0:8].ravel())
    assert np.array_equal(c1, np.concatenate([re[8:12].ravel(), re[12:24:2].ravel()]))
    assert np.array_equal(sch, np.concatenate([re[13:24:2].ravel(), re[24:36].ravel()]))
    assert c2.size == 0


def test_hand_worked_one_bit_harq_ack():
    """QPSK, 1 layer, 1 PRB, DM-RS symbol 0 (no data REs): 6 reserved REs in symbol 1
    (spacing 2: REs 0, 2, .., 10), a 1-bit HARQ-ACK on 3 of them (every 2nd: REs 0, 4, 8).
    The UL-SCH keeps all 12 REs of symbol 1 with REs 0, 4, 8 zeroed; the HARQ-ACK soft
    bits are those REs with bit 1 negated where c(n) ^ c(n + 1) = 1."""
    cfg = dict(modulation=2, nof_layers=1, nof_prb=1, start_symbol_index=0, nof_symbols=3, dmrs_symbol_mask=1,
               dmrs_type=1, nof_cdm_groups_without_data=2, nof_harq_ack_rvd=12, nof_harq_ack_bits=1,
               nof_enc_harq_ack_bits=6)
    llr = (np.arange(48) * 5 - 110).astype(np.int8)
    c_init = 0x1234567
    sch, ack, _, _ = O.ulsch_demultiplex(cfg, llr, c_init)
    c = O.scrambling_sequence(c_init, 48)
    re = llr.reshape(24, 2).copy()
    want_ack = []
    for x in (0, 4, 8):
        n = 2 * x
        want_ack += [re[x, 0], -re[x, 1] if c[n] ^ c[n + 1] else re[x, 1]]
    assert np.array_equal(ack, np.array(want_ack, np.int8))
    re[[0, 4, 8]] = 0
    assert np.array_equal(sch, re.ravel())


def test_no_uci_is_identity():
    rng = np.random.default_rng(5)
    for _ in range(50):
        cfg = random_config(rng, uci=False)
        sz = O.demux_sizes(cfg)
        llr = _llrs(rng, sz["nof_llrs"])
        sch, ack, c1, c2 = O.ulsch_demultiplex(cfg, llr, 1)
        assert np.array_equal(sch, llr) and ack.size == c1.size == c2.size == 0


def _tx_all_placeholders(rng, cfg, n, c, payload_bits):
    """Soft bits of a codeword whose every modulation symbol is sent as a UCI symbol of a
    `payload_bits` payload (TS38.212 5.3.3.1-2: [a, y, x, ..] for one bit, [a0, a1, x, ..]
    for two), scrambled with the placeholder rules of TS38.211 6.3.1.1, then descrambled
    as the receiver does (noise-free, LLR +100 for bit 0)."""
    qm = O.bits_per_symbol(cfg["modulation"])
    b = rng.integers(0, 2, n).astype(np.uint8)
    bt = np.zeros(n, np.uint8)  # scrambled bits
    for i in range(n):
        k = i % qm
        if k >= 2:
            bt[i] = 1  # x
        elif k == 1 and payload_bits == 1:
            bt[i] = bt[i - 1]  # y
        else:
            bt[i] = b[i] ^ c[i]
    llr = np.where((bt ^ c) == 0, 100, -100).astype(np.int8)
    return b, llr


@pytest.mark.parametrize("field", ["harq_ack", "csi_part1", "csi_part2"])
@pytest.mark.parametrize("payload_bits", [1, 2])
def test_placeholders_recover_transmitter_bits(field, payload_bits):
    """Whatever REs the field gets, its soft bits after demultiplexing read: bit 0 = the
    payload bit sent, bit 1 = the same bit again (1-bit payload, y) or the second bit
    sent (2 bits), bits 2.. = 1 (x)."""
    rng = np.random.default_rng(11 + payload_bits)
    key = {"harq_ack": 1, "csi_part1": 2, "csi_part2": 3}[field]
    checked = 0
    for _ in range(60):
        cfg = random_config(rng)
        if cfg.get({"harq_ack": "nof_harq_ack_bits", "csi_part1": "nof_csi_part1_bits",
                    "csi_part2": "nof_csi_part2_bits"}[field], 0) != payload_bits:
            continue
        qm = O.bits_per_symbol(cfg["modulation"])
        sz = O.demux_sizes(cfg)
        c_init = int(rng.integers(0, 2 ** 31))
        c = O.scrambling_sequence(c_init, sz["nof_llrs"])
        _, llr = _tx_all_placeholders(rng, cfg, sz["nof_llrs"], c, payload_bits)
        out = O.ulsch_demultiplex(cfg, llr, c_init)[key].reshape(-1, qm)
        if qm >= 2:
            if payload_bits == 1:
                assert np.array_equal(np.sign(out[:, 1]), np.sign(out[:, 0]))
            assert np.all(out[:, 2:] <= 0)
            assert np.all(np.abs(out[:, 2:]) == 100) or field == "csi_part2"
        checked += 1
    assert checked > 3


def test_rejected_configurations():
    base = dict(modulation=4, nof_layers=2, nof_prb=4, start_symbol_index=0, nof_symbols=14,
                dmrs_symbol_mask=(1 << 2) | (1 << 11), dmrs_type=1, nof_cdm_groups_without_data=2)
    assert O.demux_sizes(base) is not None
    assert O.demux_sizes(dict(base, dmrs_symbol_mask=0)) is None
    assert O.demux_sizes(dict(base, nof_cdm_groups_without_data=3)) is None
    assert O.demux_sizes(dict(base, nof_harq_ack_bits=4, nof_enc_harq_ack_bits=12)) is None  # not a multiple of 8
    assert O.demux_sizes(dict(base, nof_harq_ack_bits=4)) is None  # payload without encoded bits
    assert O.demux_sizes(dict(base, nof_enc_csi_part1_bits=8)) is None  # encoded bits without payload
    # More HARQ-ACK REs than the UCI REs from l1 on (11 symbols x 48 REs).
    assert O.demux_sizes(dict(base, nof_harq_ack_bits=4, nof_enc_harq_ack_bits=8 * 48 * 11)) is None
    assert O.demux_sizes(dict(base, nof_harq_ack_bits=4, nof_enc_harq_ack_bits=8 * 48 * 10)) is not None


def test_abi_sizes_match_oracle():
    """lh_ulsch_demux_get_sizes (host only) accepts and rejects what the oracle does and
    gives its sizes."""
    import ocudu_hip as H
    rng = np.random.default_rng(7)
    bad = 0
    for case in range(400):
        cfg = random_config(rng, max_prb=40)
        if case % 3 == 0:  # perturb into configurations the reference may reject
            k = ("nof_enc_harq_ack_bits", "nof_enc_csi_part1_bits", "nof_csi_part2_bits", "dmrs_symbol_mask",
                 "nof_cdm_groups_without_data")[case % 5]
            cfg[k] = int(cfg.get(k, 0)) + int(rng.integers(1, 4))
        want = O.demux_sizes(cfg)
        d = H.UlschDemuxDesc(**{f: int(cfg.get(f, 0)) for f in H.DEMUX_FIELDS if f != "c_init"})
        try:
            got = H.ulsch_demux_sizes(d)
        except RuntimeError:
            got = None
        assert got == want, f"case {case} {cfg}"
        bad += want is None
    assert 20 < bad < 300


@pytest.mark.skipif(not RANK_BIN.exists(), reason="tests/cpp/demux_rank_test not built")
def test_device_arithmetic_replayed_on_host():
    rng = np.random.default_rng(99)
    lines = []
    for case in range(600):
        cfg = random_config(rng, max_prb=24 if case % 5 else 275, uci=case % 10 != 0)
        lines.append(" ".join(str(v) for v in config_words(cfg, int(rng.integers(0, 2 ** 31)))))
    # Also configurations the plan must reject like the oracle.
    lines.append("4 2 4 0 14 0 1 2 0 0 0 0 0 0 0 5")
    lines.append("4 2 4 0 14 2052 1 2 0 4 12 0 0 0 0 5")
    out = subprocess.run([str(RANK_BIN)], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == f"ok {len(lines)}"


def _random_ulsch_config(rng):
    while True:
        mod = int(rng.choice([0, 1, 2, 4, 6, 8]))
        start = int(rng.integers(0, 4))
        nsym = int(rng.integers(4, 15 - start))
        from demux_util import _dmrs_mask
        cfg = dict(tbs_bits=int(rng.choice([0, 24, 200, 300, 1000, 3824, 3840, 8000, 40000, 150000])) // 8 * 8,
                   target_code_rate=float(rng.choice([0.0, 0.12, 0.3, 0.5, 0.66, 0.7, 0.9, rng.uniform(0.05, 0.95)])),
                   modulation=mod, nof_harq_ack_bits=int(rng.choice([0, 1, 2, 3, 11, 12, 19, 20, 40])),
                   nof_csi_part1_bits=int(rng.choice([0, 0, 5, 15, 30])), nof_csi_part2_bits=0,
                   alpha_scaling=float(rng.choice([0.5, 0.65, 0.8, 1.0])),
                   beta_offset_harq_ack=float(rng.choice([2.0, 5.0, 8.0, 20.0, 126.0])),
                   beta_offset_csi_part1=float(rng.choice([1.125, 2.0, 6.25])),
                   beta_offset_csi_part2=float(rng.choice([1.125, 2.0, 6.25])), nof_rb=int(rng.integers(1, 276)),
                   start_symbol_index=start, nof_symbols=nsym, dmrs_type=int(rng.integers(1, 3)),
                   dmrs_symbol_mask=_dmrs_mask(rng, start, nsym), nof_layers=1 if mod == 0 else int(rng.integers(1, 5)),
                   contains_dc=int(rng.integers(0, 2)))
        cfg["nof_cdm_groups_without_data"] = int(rng.integers(1, 3 if cfg["dmrs_type"] == 1 else 4))
        if cfg["nof_csi_part1_bits"] and rng.random() < 0.5:
            cfg["nof_csi_part2_bits"] = int(rng.choice([3, 20, 60]))
        if cfg["tbs_bits"] == 0 and cfg["target_code_rate"] == 0.0:
            continue  # a UCI-only grant divides by the code rate
        return cfg


@pytest.mark.skipif(O.ref_demod() is None, reason="reference build absent (oracle/_ref)")
def test_ulsch_information_matches_reference_build():
    """orc_ulsch_information against the reference's get_ulsch_information
    (lib/ran/pusch/ulsch_info.cpp, lib/ran/sch/sch_segmentation.cpp built from source),
    every field, over grants with and without a TB, every UCI size class (CRC 0/6/11
    bits), reservation for 0..2 HARQ-ACK bits, CSI Part 2, code rates including 0."""
    rng = np.random.default_rng(44)
    for case in range(3000):
        cfg = _random_ulsch_config(rng)
        want = O.ref_ulsch_information(cfg)
        got = O.ulsch_information(cfg)
        assert got == want, f"case {case} {cfg}"


def test_ulsch_information_hand_worked():
    """TBS 1000 bits at R 0.5: BG2 (1000 <= 3824, R <= 0.67), one codeblock, K_b 10
    (1016 > 640), Z = 104 (smallest with 10 Z >= 1016), 1040 bits per codeblock.  16-QAM,
    1 layer, 10 PRBs, symbols 0..13, DM-RS on 2 and 11 (type 1, 2 CDM groups: no data
    there): 12 x 120 = 1440 UCI REs, 11 x 120 = 1320 after the first DM-RS.  HARQ-ACK 3
    bits, beta 8: ceil(3 * 8 * 1440 / 1040) = 34 REs (alpha 1 bound 1320); no reservation.
    CSI Part 1 5 bits, beta 2: ceil(5 * 2 * 1440 / 1040) = 14.  UL-SCH: 1440 - 34 - 14 REs."""
    cfg = dict(tbs_bits=1000, target_code_rate=0.5, modulation=4, nof_harq_ack_bits=3, nof_csi_part1_bits=5,
               alpha_scaling=1.0, beta_offset_harq_ack=8.0, beta_offset_csi_part1=2.0, beta_offset_csi_part2=1.0,
               nof_rb=10, start_symbol_index=0, nof_symbols=14, dmrs_type=1, dmrs_symbol_mask=(1 << 2) | (1 << 11),
               nof_cdm_groups_without_data=2, nof_layers=1)
    i = O.ulsch_information(cfg)
    assert (i["sch_base_graph"], i["sch_nof_cb"], i["sch_lifting_size"], i["sch_nof_bits_per_cb"]) == (2, 1, 104, 1040)
    assert (i["nof_harq_ack_re"], i["nof_harq_ack_rvd"], i["nof_csi_part1_re"]) == (34, 0, 14)
    assert i["nof_ul_sch_bits"] == (1440 - 34 - 14) * 4
    # One HARQ-ACK bit: the reservation is sized for two bits (ceil(2 * 8 * 1440 / 1040) = 23),
    # the HARQ-ACK itself for one (ceil(8 * 1440 / 1040) = 12), and neither leaves UL-SCH.
    i1 = O.ulsch_information(dict(cfg, nof_harq_ack_bits=1))
    assert (i1["nof_harq_ack_re"], i1["nof_harq_ack_rvd"]) == (12, 23 * 4)
    assert i1["nof_csi_part1_re"] == 14 and i1["nof_ul_sch_bits"] == (1440 - 14) * 4


def test_abi_ulsch_information_matches_oracle():
    """lh_ulsch_information (host only) gives the oracle's accounting, field by field."""
    import ocudu_hip as H
    rng = np.random.default_rng(45)
    for case in range(3000):
        cfg = _random_ulsch_config(rng)
        want = O.ulsch_information(cfg)
        got = H.ulsch_information(**cfg)
        for k, v in want.items():
            assert got[k] == v, f"case {case} {cfg}: {k} {got[k]} vs oracle {v}"
