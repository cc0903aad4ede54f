"""UL-SCH demultiplexing on the device (lh_ulsch_demultiplex) against the oracle
(orc_ulsch_demultiplex, pinned to the reference's own ulsch_demultiplex_impl in
test_oracle_demux.py), bit-exact in every field: random configurations (every modulation
and layer count, HARQ-ACK of 0..2 bits over reserved REs or more bits, CSI Part 1 and 2
of 1, 2 or more bits, DM-RS symbols without data REs), whole batches in one call with
output offsets of every alignment (1-, 2- and 4-byte units), 273-PRB 4-layer 256-QAM
codewords, and the bytes between the outputs left untouched."""
import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from demux_util import random_config
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

FIELDS = ("sch", "harq_ack", "csi_part1", "csi_part2")
SIZE_KEYS = ("nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")
OFF_KEYS = ("sch_offset", "harq_ack_offset", "csi_part1_offset", "csi_part2_offset")


def _run_batch(ctx, cases, rng, gap_max=9):
    """cases: (cfg, llr, c_init).  One lh_ulsch_demultiplex call; returns the outputs per
    case and checks that the gaps between outputs keep their fill byte."""
    descs, lo, outs_off, total = [], 0, [], [0, 0, 0, 0]
    llr_blob = []
    for cfg, llr, ci in cases:
        sz = O.demux_sizes(cfg)
        gap = int(rng.integers(0, gap_max))
        lo += gap
        llr_blob.append((lo, llr))
        offs = []
        for k, key in enumerate(SIZE_KEYS):
            total[k] += int(rng.integers(0, gap_max))
            offs.append(total[k])
            total[k] += sz[key]
        outs_off.append(offs)
        d = H.UlschDemuxDesc(**{f: int(cfg.get(f, 0)) for f in H.DEMUX_FIELDS if f != "c_init"})
        d.c_init = ci
        d.llr_offset = lo
        for k, key in enumerate(OFF_KEYS):
            setattr(d, key, offs[k])
        descs.append(d)
        lo += llr.size
    host_in = np.zeros(lo + 8, np.int8)
    for off, llr in llr_blob:
        host_in[off:off + llr.size] = llr
    d_in = H.DeviceBuffer.from_array(host_in)
    fill = np.int8(0x55)
    bufs = [H.DeviceBuffer.from_array(np.full(t + 8, fill, np.int8)) for t in total]
    ctx.ulsch_demultiplex(descs, d_in, *bufs)
    got = [b.download(np.int8, t + 8) for b, t in zip(bufs, total)]
    results, written = [], [np.zeros(t + 8, bool) for t in total]
    for (cfg, llr, ci), offs in zip(cases, outs_off):
        sz = O.demux_sizes(cfg)
        res = []
        for k, key in enumerate(SIZE_KEYS):
            res.append(got[k][offs[k]:offs[k] + sz[key]])
            written[k][offs[k]:offs[k] + sz[key]] = True
        results.append(res)
    for k in range(4):
        assert np.all(got[k][~written[k]] == fill), f"{FIELDS[k]}: bytes outside the outputs were written"
    return results


def _check(cases, results):
    for i, ((cfg, llr, ci), got) in enumerate(zip(cases, results)):
        want = O.ulsch_demultiplex(cfg, llr, ci)
        for k in range(4):
            assert np.array_equal(got[k], want[k]), (
                f"codeword {i} {cfg}: {FIELDS[k]} differs at {np.flatnonzero(got[k] != want[k])[:8]}")


@pytest.fixture(scope="module")
def ctx():
    require_gpu()
    c = H.Context(0, max_codeblocks=4096)
    yield c
    c.close()


def test_random_batch_matches_oracle(ctx):
    rng = np.random.default_rng(31)
    cases = []
    for i in range(300):
        cfg = random_config(rng, max_prb=32, uci=i % 8 != 0)
        n = O.demux_sizes(cfg)["nof_llrs"]
        cases.append((cfg, rng.integers(-120, 121, n).astype(np.int8), int(rng.integers(0, 2 ** 31))))
    print(f"case demux random batch, {len(cases)} codewords", flush=True)
    _check(cases, _run_batch(ctx, cases, rng))


def test_aligned_units_match_oracle(ctx):
    """Offsets all multiples of 4 (4-byte units where the bits per RE allow) and of 2."""
    rng = np.random.default_rng(32)
    for gap_unit in (4, 2):
        cases = []
        for i in range(120):
            cfg = random_config(rng, max_prb=32)
            n = O.demux_sizes(cfg)["nof_llrs"]
            cases.append((cfg, rng.integers(-120, 121, n).astype(np.int8), int(rng.integers(0, 2 ** 31))))
        print(f"case demux aligned batch, unit {gap_unit}", flush=True)

        class Aligned:
            def integers(self, lo, hi):
                return gap_unit * int(rng.integers(lo, hi))

        _check(cases, _run_batch(ctx, cases, Aligned()))


@pytest.mark.parametrize("qm_layers", [(8, 4), (6, 2), (2, 1), (4, 3)])
def test_full_carrier_codewords(ctx, qm_layers):
    """273 PRBs, 14 symbols: the largest OFDM symbols (3276 REs, up to 32 soft bits each)."""
    rng = np.random.default_rng(33 + qm_layers[0])
    cases = []
    while len(cases) < 6:
        cfg = random_config(rng, max_prb=275, qm_layers=qm_layers)
        cfg.update(nof_prb=273, start_symbol_index=0, nof_symbols=14)
        if O.demux_sizes(cfg) is None:
            continue
        n = O.demux_sizes(cfg)["nof_llrs"]
        cases.append((cfg, rng.integers(-120, 121, n).astype(np.int8), int(rng.integers(0, 2 ** 31))))
    print(f"case demux full carrier {qm_layers}", flush=True)
    _check(cases, _run_batch(ctx, cases, rng))


def test_rejects_invalid_and_null_outputs(ctx):
    cfg = dict(modulation=4, nof_layers=1, nof_prb=4, start_symbol_index=0, nof_symbols=14,
               dmrs_symbol_mask=1 << 2, dmrs_type=1, nof_cdm_groups_without_data=2, nof_harq_ack_bits=5,
               nof_enc_harq_ack_bits=40)
    d = H.UlschDemuxDesc(**cfg)
    n = O.demux_sizes(cfg)["nof_llrs"]
    d_in = H.DeviceBuffer(n)
    d_sch = H.DeviceBuffer(n)
    with pytest.raises(H.LhError, match="NULL"):
        ctx.ulsch_demultiplex([d], d_in, d_sch)
    bad = H.UlschDemuxDesc(**dict(cfg, nof_enc_harq_ack_bits=42))
    with pytest.raises(H.LhError, match="multiple"):
        ctx.ulsch_demultiplex([bad], d_in, d_sch, d_sch)
