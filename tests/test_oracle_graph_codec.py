"""Base-graph tables, encoder, rate matcher and the numpy data generator, pinned by
the standard's invariants (nnz, row degrees, H*c = 0) and by the reference build."""
import numpy as np
import pytest

import oracle as O
import np_datagen as D

# TS38.212 Table 5.3.2-2/-3 structure
NNZ = {1: 316, 2: 197}
CORE_DEG = {1: [19, 19, 19, 19], 2: [8, 10, 8, 10]}


@pytest.mark.parametrize("bg", [1, 2])
def test_table_invariants(bg):
    rs, cols, v = D.base_graph(bg)
    K, M = D.k_b(bg), D.m_rows(bg)
    assert rs[-1] == NNZ[bg] and len(rs) == M + 1
    deg = np.diff(rs)
    assert list(deg[:4]) == CORE_DEG[bg]
    for i in range(M):
        c = cols[rs[i]:rs[i + 1]]
        assert list(c) == sorted(c)
        ext = [x for x in c if x >= K + 4]
        assert (i < 4 and not ext) or (i >= 4 and ext == [K + i])
    assert v.shape == (8, NNZ[bg]) and v.min() >= 0 and v.max() < 384


@pytest.mark.parametrize("bg", [1, 2])
def test_shifts_against_reference_build(bg):
    R = O.ref()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not available")
    K, M = O.K_b(bg), O.M_rows(bg)
    for Z in O.LIFTING_SIZES:
        for m in range(M):
            for n in range(K + M):
                assert O.lib().orc_graph_shift(bg, Z, m, n) == R.ref_graph_shift(bg, Z, m, n), (bg, Z, m, n)


@pytest.mark.parametrize("bg", [1, 2])
def test_encoder_parity_checks(bg):
    rng = np.random.default_rng(bg)
    for Z in O.LIFTING_SIZES[::5] + [384]:
        K = O.K_b(bg) * Z
        m = rng.integers(0, 2, (2, K), dtype=np.uint8)
        cw = D.encode(m, bg, Z)
        for i in range(2):
            assert O.syndrome_weight(cw[i], bg, Z) == 0
            assert np.array_equal(cw[i], O.encode(m[i], bg, Z))
        # a single flipped bit violates the parity checks
        bad = cw[0].copy()
        bad[3] ^= 1
        assert O.syndrome_weight(bad, bg, Z) > 0


def test_encoder_rate_matcher_against_reference_build():
    R = O.ref()
    if R is None:
        pytest.skip("reference build (oracle/_ref) not available")
    rng = np.random.default_rng(5)
    for bg in (1, 2):
        for Z in (2, 9, 40, 176, 384):
            K, N = O.K_b(bg) * Z, O.N_short(bg) * Z
            nsys = (O.K_b(bg) - 2) * Z
            for rv in range(4):
                for Qm in (1, 2, 4, 6, 8):
                    F = int(rng.choice([0, 4, 8])) if nsys > 8 else 0
                    m = rng.integers(0, 2, K, dtype=np.uint8)
                    m[K - F:] = 0
                    E = max(Qm, int(N * rng.uniform(0.3, 1.6)) // Qm * Qm)
                    cw_ref = np.zeros(N, np.uint8)
                    rm_ref = np.zeros(E, np.uint8)
                    R.ref_encode_rate_match(O._p(m, O._u8p), bg, Z, O._p(cw_ref, O._u8p), O._p(rm_ref, O._u8p), E,
                                            rv, Qm, 0, F)
                    cw = O.encode(m, bg, Z)[2 * Z:]
                    assert np.array_equal(cw, cw_ref)
                    assert np.array_equal(O.rate_match(cw, E, bg, Z, rv=rv, Qm=Qm, F=F), rm_ref)
                    idx = D.rate_match_indices(bg, Z, E, rv, Qm, 0, F)
                    assert np.array_equal(cw[idx], rm_ref)


def test_datagen_mapper_demapper_against_oracle():
    import ctypes as C
    rng = np.random.default_rng(8)
    for Qm, snr in ((2, 2.0), (4, 10.0), (6, 16.0), (8, 24.0)):
        bits = rng.integers(0, 2, (1, Qm * 500), dtype=np.uint8)
        iq = D.modulate(bits, Qm)
        iq2 = np.zeros(1000, np.float32)
        L = O.lib()
        L.orc_modulate.argtypes = [C.POINTER(C.c_float), O._u8p, C.c_uint, C.c_uint]
        L.orc_demap.argtypes = [O._i8p, C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_float]
        b0 = np.ascontiguousarray(bits[0])
        L.orc_modulate(iq2.ctypes.data_as(C.POINTER(C.c_float)), O._p(b0, O._u8p), b0.size, Qm)
        assert np.allclose(iq.reshape(-1), iq2)
        # unit average symbol energy
        assert abs(float((iq ** 2).sum(-1).mean()) - 1.0) < 0.15
        n0 = 10 ** (-snr / 10)
        y = (iq + rng.normal(0, np.sqrt(n0 / 2), iq.shape)).astype(np.float32)
        l1 = D.demap(y, Qm, n0)
        l2 = np.zeros(bits.size, np.int8)
        yf = np.ascontiguousarray(y.reshape(-1))
        L.orc_demap(O._p(l2, O._i8p), yf.ctypes.data_as(C.POINTER(C.c_float)), 500, Qm, n0)
        assert np.array_equal(l1[0], l2)
        # noiseless: every LLR sign gives the transmitted bit
        l0 = D.demap(iq, Qm, 1e-3)
        assert np.array_equal((l0[0] <= 0).astype(np.uint8), bits[0])


def test_segmentation_kats():
    # pusch_decoder_capacity_test.cpp:9-25 geometry: 4 PRB x 1 layer -> 2 CBs;
    # 51 PRB x 1 layer -> 18 CBs; closed forms of TS38.212 5.2.2.
    assert O.lib().orc_nof_codeblocks(8424, 1) == 1
    assert O.lib().orc_nof_codeblocks(8425, 1) == 2
    assert O.lib().orc_nof_codeblocks(3816, 2) == 1
    for tbs in (24, 100, 3824, 3825, 8424, 20000, 100000, 500000):
        for bg in (1, 2):
            C = O.lib().orc_nof_codeblocks(tbs, bg)
            B = tbs + (16 if tbs <= 3824 else 24)
            kcb = 8448 if bg == 1 else 3840
            assert C == (1 if B <= kcb else -(-B // (kcb - 24)))
            Z = O.lib().orc_compute_lifting_size(tbs, bg, C)
            kb = 22 if bg == 1 else (10 if B > 640 else 9 if B > 560 else 8 if B > 192 else 6)
            Bp = B + (24 * C if C > 1 else 0)
            assert Z == min(z for z in O.LIFTING_SIZES if z * kb * C >= Bp)
