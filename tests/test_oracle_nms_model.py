"""The learned decoder's floating-point training model in the oracle (oracle.nms_forward)
pinned by an independent numpy restatement (vectorised over check rows) on small lifting
sizes, by the closed form at alpha = 0 (no check messages: the loss is the channel's BCE
after every iteration) and by float/double agreement."""
import numpy as np
import pytest

import oracle as O
import np_datagen as D


def _np_forward(llrs, msg, bg, Z, L, alpha, beta, w, tau, F=0):
    rs, cols, v = D.base_graph(bg)
    sh = v[D.lifting_index(Z)] % Z
    K = O.K_b(bg)
    soft = np.zeros((K + L) * Z)
    n = min(llrs.size, (K + L - 2) * Z)
    soft[2 * Z:2 * Z + n] = np.clip(llrs[:n].astype(np.float64), -64, 64)
    c2v = {}
    bits = np.unpackbits(np.asarray(msg, np.uint8))[:K * Z - F].astype(np.float64)
    sg = 1 - 2 * bits
    loss = 0.0
    r = np.arange(Z)
    for t in range(alpha.shape[0]):
        for l in range(L):
            es = range(rs[l], rs[l + 1])
            addr = np.stack([cols[e] * Z + (r + sh[e]) % Z for e in es])  # (deg, Z)
            old = np.stack([c2v.get((l, e), np.zeros(Z)) for e in es])
            vv = soft[addr] - old
            mag = np.abs(vv)
            order = np.argsort(mag, axis=0, kind="stable")  # ties: first edge first
            a1 = order[0]
            m1 = mag[a1, r]
            m2 = mag[order[1], r]
            neg = (vv < 0).sum(axis=0) % 2
            for k, e in enumerate(es):
                mg = np.where(a1 == k, m2, m1)
                c = np.maximum(alpha[t, l] * mg - beta[t, l], 0.0)
                c = np.where((neg ^ (vv[k] < 0)) == 1, -c, c)
                c2v[(l, e)] = c
                soft[addr[k]] = vv[k] + c
        z = sg * soft[:K * Z - F] * tau
        loss += w[t] * np.mean(np.maximum(-z, 0) + np.log1p(np.exp(-np.abs(z))))
    return loss


@pytest.mark.parametrize("bg,Z,L,E,Qm,snr", [(2, 2, 6, 60, 2, 1.0), (2, 5, 9, 120, 2, 2.0), (1, 3, 7, 152, 4, 6.0),
                                          (1, 8, 12, 400, 2, 0.0), (2, 13, 20, 400, 2, -1.0)])
def test_model_matches_numpy_restatement(bg, Z, L, E, Qm, snr):
    rng = np.random.default_rng(Z * 7 + L)
    msgs, llrs = O.make_codeblocks(2, bg, Z, E, Qm, snr, seed=Z, crc_poly=O.CRC16, crc_bits=16)
    M = O.M_rows(bg)
    T = 4
    alpha = rng.uniform(0.5, 1.2, (T, M))
    beta = rng.uniform(-1.0, 3.0, (T, M))
    w = rng.uniform(0.5, 1.0, T)
    for i in range(2):
        rm = O.rate_dematch(llrs[i], bg, Z, Qm=Qm)
        ref = _np_forward(rm, msgs[i], bg, Z, L, alpha, beta, w, 0.125)
        got = O.nms_forward(rm, msgs[i], bg, Z, L, alpha, beta, w, 0.125)
        assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref)), (got, ref)
        # Float arithmetic may break a near tie of two magnitudes the other way: loose bound.
        gotf = O.nms_forward(rm, msgs[i], bg, Z, L, alpha, beta, w, 0.125, double=False)
        assert abs(gotf - ref) <= 1e-2 * max(1.0, abs(ref))


def test_model_without_messages_is_channel_bce():
    bg, Z, L = 2, 24, 8
    msgs, llrs = O.make_codeblocks(1, bg, Z, 800, 2, 1.0, seed=4, crc_poly=O.CRC16, crc_bits=16)
    rm = O.rate_dematch(llrs[0], bg, Z, Qm=2)
    M = O.M_rows(bg)
    T = 3
    alpha = np.zeros((T, M))
    beta = np.ones((T, M))  # alpha * m - beta < 0: every message is 0
    w = np.array([1.0, 0.5, 0.25])
    K = O.K_b(bg) * Z
    soft = np.concatenate([np.zeros(2 * Z), np.clip(rm.astype(np.float64), -64, 64)])[:K]
    z = (1 - 2 * np.unpackbits(msgs[0])[:K].astype(np.float64)) * soft * 0.125
    bce = np.mean(np.maximum(-z, 0) + np.log1p(np.exp(-np.abs(z))))
    got = O.nms_forward(rm, msgs[0], bg, Z, L, alpha, beta, w, 0.125)
    assert abs(got - w.sum() * bce) < 1e-12
