"""Rate dematcher oracle, pinned by the reference test's own golden restatement of
TS38.212 5.4.2 (tests/unittests/phy/upper/channel_coding/ldpc/ldpc_rate_dematcher_test.cpp:175-248,
restated here independently in Python) over that test's parameter sweep (:290-386),
and by the reference build."""
import numpy as np
import pytest

import oracle as O

K_B = {1: 22, 2: 10}


def _sat_add(a, b):
    if a == -b:
        return 0
    if abs(a) > 120:
        return a
    if abs(b) > 120:
        return b
    return max(-120, min(120, a + b))


def golden(rm_in, llrs, new_data, bg, Z, rv, Qm, Nref, F):
    full = O.N_short(bg) * Z
    size = min(Nref if Nref > 0 else 384 * 66, full)
    f = (0, 17, 33, 56) if bg == 1 else (0, 13, 25, 43)
    k0 = int(np.floor((f[rv] * size) / full)) * Z
    nsys = (K_B[bg] - 2) * Z
    info = nsys - F
    E = llrs.size
    Kq = E // Qm
    de = np.zeros(E, np.int64)
    idx = 0
    for i in range(Kq):
        for j in range(Qm):
            de[j * Kq + i] = llrs[idx]
            idx += 1
    out = [0] * full if new_data else [int(x) for x in rm_in]
    for p in range(info, nsys):
        out[p] = 127
    pos = k0
    for i in range(E):
        if info <= pos < nsys:
            pos = nsys
        out[pos] = _sat_add(out[pos], int(de[i]))
        pos = (pos + 1) % size
    return np.array(out, np.int8)


def cases():
    rng = np.random.default_rng(0)
    out = []
    for bg in (1, 2):
        for Z in (2, 16, 32, 64, 96, 160, 256, 384):
            for rv in range(4):
                for ratio in (0.0, 0.75):
                    for new_data in (True, False):
                        Qm = int(rng.choice([1, 2, 4, 6, 8]))
                        F = int(rng.choice([0, 2, 4, 8]))
                        rate = float(rng.uniform(0.1, 0.95))
                        nsys = (K_B[bg] - 2) * Z
                        Nref = int(ratio * O.N_short(bg) * Z)
                        E = int(np.floor((nsys - F) / rate / Qm)) * Qm
                        if E == 0 or F >= nsys:
                            continue
                        out.append((bg, Z, rv, Qm, Nref, F, E, new_data))
    return out


def test_oracle_matches_reference_golden_restatement():
    rng = np.random.default_rng(1)
    for (bg, Z, rv, Qm, Nref, F, E, new_data) in cases():
        llrs = rng.integers(-120, 121, E).astype(np.int8)
        rm_in = rng.integers(-120, 121, O.N_short(bg) * Z).astype(np.int8)
        exp = golden(rm_in, llrs, new_data, bg, Z, rv, Qm, Nref, F)
        got = O.rate_dematch(llrs, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F, new_data=new_data, harq=rm_in)
        assert np.array_equal(got, exp), (bg, Z, rv, Qm, Nref, F, E, new_data)
        if O.ref() is not None:
            ref = O.ref_rate_dematch(llrs, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F, new_data=new_data, harq=rm_in)
            assert np.array_equal(got, ref)


def test_infinity_combining():
    # HARQ combining keeps the +-127 rules of the LLR type (log_likelihood_ratio.cpp:22-55).
    bg, Z = 2, 8
    N = O.N_short(bg) * Z
    llrs = np.array([127, -127, 120, -120] * 40, np.int8)
    rm_in = np.array(([127, -127, 5, -5] * N)[:N], np.int8)
    got = O.rate_dematch(llrs, bg, Z, Qm=2, new_data=False, harq=rm_in)
    exp = golden(rm_in, llrs, False, bg, Z, 0, 2, 0, 0)
    assert np.array_equal(got, exp)
