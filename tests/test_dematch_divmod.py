"""The fused dematcher's quotient i div Kq (deinterleaving, ldpc_rate_dematcher_impl.cpp
:192-202): the device takes trunc((float)i * rcp_kq) with rcp_kq = float((1 + 2^-21)/Kq)
and no correction (fast new-data path).  Checked here with IEEE float32 arithmetic (the
same round-to-nearest conversion and product, truncating conversion) against integer
division, exhaustively over i < E for every E = Kq * Qm that a sample of Kq values
allows, every Kq up to 4096 and the extremes up to the largest rate-matched length."""
import numpy as np

MAX_E = 22 * 384 * 35  # largest rate-matched length (MAX_CODEBLOCK_RM_SIZE)


def rcp(kq):
    return np.float32((1.0 + 1.0 / 2097152.0) / kq)


def check(kq, e):
    i = np.arange(e, dtype=np.uint32)
    q = (i.astype(np.float32) * rcp(kq)).astype(np.uint32)
    bad = np.nonzero(q != i // kq)[0]
    assert bad.size == 0, f"Kq={kq}: i={int(bad[0])} gives {int(q[bad[0]])}, want {int(bad[0]) // kq}"


def test_small_kq_exhaustive():
    for kq in range(1, 4097):
        check(kq, min(8 * kq, MAX_E))


def test_large_kq_sample():
    rng = np.random.default_rng(0)
    kqs = set(rng.integers(4097, MAX_E + 1, 150).tolist()) | {MAX_E, MAX_E // 2, MAX_E // 8, MAX_E // 6, 36960}
    for kq in sorted(kqs):
        qm = max(1, min(8, MAX_E // kq))
        check(kq, kq * qm)
