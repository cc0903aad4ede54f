"""FP32 training kernels of the learned decoder (lh_nms_train_step, lh_nms_adam): the
GPU forward loss equals the oracle's float model, the GPU backward gradients equal central
finite differences of the oracle's double-precision model, and Adam steps reduce the loss.

Tolerances: loss within 1e-5 (relative; float sums in another order).  Gradients: the
finite differences see the piecewise-linear model (min, relu) in double precision while
the kernels differentiate the float trajectory, so a near tie may resolve differently;
every component within 2% of the largest gradient, and 95% of them within 2% of their own
value (or 1e-4 of the largest)."""
import ctypes as C

import numpy as np
import pytest

import ocudu_hip as H
import oracle as O
from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


class _View:
    def __init__(self, buf, off_bytes, nbytes):
        self.ptr = buf.ptr + off_bytes
        self.nbytes = nbytes


@pytest.fixture(scope="module")
def ctx():
    require_gpu()
    c = H.Context(0, max_codeblocks=1024)
    yield c
    c.close()


def _setup(bg, specs, seed):
    """specs: (Z, L, E, Qm, snr); returns host data and TrainDescs."""
    llr_parts, msg_parts, descs, host = [], [], [], []
    lo = mo = 0
    for i, (Z, L, E, Qm, snr) in enumerate(specs):
        msgs, llrs = O.make_codeblocks(1, bg, Z, E, Qm, snr, seed=seed + i)
        rm = O.rate_dematch(llrs[0], bg, Z, Qm=Qm)
        descs.append(H.TrainDesc(lo, mo, rm.size, Z, 0, L))
        host.append((rm, msgs[0], Z, L))
        llr_parts.append(rm)
        msg_parts.append(msgs[0])
        lo += rm.size
        mo += msgs[0].size
    return np.concatenate(llr_parts), np.concatenate(msg_parts), descs, host


def _gpu_step(ctx, bg, descs, llr, msg, alpha, beta, w, tau, backward=True):
    T, M = alpha.shape
    d_llr = H.DeviceBuffer.from_array(llr)
    d_msg = H.DeviceBuffer.from_array(msg)
    d_par = H.DeviceBuffer.from_array(np.concatenate([alpha.ravel(), beta.ravel()]).astype(np.float32))
    d_grad = H.DeviceBuffer(8 * T * M)
    d_grad.zero()
    d_ws = H.DeviceBuffer(H.Context.nms_workspace_bytes(bg, descs, T))
    d_loss = H.DeviceBuffer(4 * len(descs))
    ctx.nms_train_step(bg, descs, T, _View(d_par, 0, 4 * T * M), _View(d_par, 4 * T * M, 4 * T * M), w, tau, d_llr,
                       d_msg, d_ws, d_loss, _View(d_grad, 0, 4 * T * M), _View(d_grad, 4 * T * M, 4 * T * M),
                       backward=backward)
    loss = d_loss.download(np.float32, len(descs))
    g = d_grad.download(np.float32, 2 * T * M)
    return loss, g[:T * M].reshape(T, M), g[T * M:].reshape(T, M)


@pytest.mark.parametrize("bg,specs", [
    (2, [(16, 10, 400, 2, 1.0), (52, 7, 1200, 2, 2.0), (208, 12, 4160, 4, 6.5)]),
    (1, [(24, 8, 1200, 2, 1.5), (384, 4, 9088, 8, 21.0)]),
])
def test_forward_and_gradients_against_oracle(ctx, bg, specs):
    print(f"case nms gradients bg{bg}", flush=True)
    rng = np.random.default_rng(bg)
    M = O.M_rows(bg)
    T = 4
    alpha = rng.uniform(0.6, 1.1, (T, M)).astype(np.float32).astype(np.float64)
    # Non-negative offsets, as training keeps them: with a negative offset the model jumps
    # where a v crosses zero, and rounding decides the sign of values that are exactly zero.
    beta = rng.uniform(0.0, 2.0, (T, M)).astype(np.float32).astype(np.float64)
    w = np.array([0.4, 0.6, 0.8, 1.0])
    tau = 0.125
    llr, msg, descs, host = _setup(bg, specs, seed=40 + bg)
    loss, ga, gb = _gpu_step(ctx, bg, descs, llr, msg, alpha, beta, w, tau)
    for i, (rm, m, Z, L) in enumerate(host):
        ref = O.nms_forward(rm, m, bg, Z, L, alpha, beta, w, tau, double=False)
        assert abs(loss[i] - ref) <= 1e-5 * max(1.0, abs(ref)), (i, loss[i], ref)

    def total(a, b):
        return sum(O.nms_forward(rm, m, bg, Z, L, a, b, w, tau) for rm, m, Z, L in host)

    Lmax = max(L for *_, L in host)
    # Small steps: the model is piecewise linear in the messages and larger steps cross
    # the kinks (a minimum or a relu changing side) of large codeblocks.
    eps = 1e-6
    fd, an = [], []
    for which, g in ((0, ga), (1, gb)):
        for t in range(T):
            for l in range(Lmax):
                a1, b1, a2, b2 = alpha.copy(), beta.copy(), alpha.copy(), beta.copy()
                if which == 0:
                    a1[t, l] += eps
                    a2[t, l] -= eps
                else:
                    b1[t, l] += eps
                    b2[t, l] -= eps
                fd.append((total(a1, b1) - total(a2, b2)) / (2 * eps))
                an.append(float(g[t, l]))
    fd, an = np.array(fd), np.array(an)
    scale = np.abs(fd).max()
    assert scale > 0
    err = np.abs(an - fd)
    assert (err <= 0.02 * scale).all(), (err.max() / scale)
    close = err <= np.maximum(0.02 * np.abs(fd), 1e-4 * scale)
    assert close.mean() >= 0.95, close.mean()
    # Layers beyond every codeblock's L get no gradient.
    assert not ga[:, Lmax:].any() and not gb[:, Lmax:].any()


def test_adam_reduces_loss(ctx):
    print("case nms adam", flush=True)
    bg, M, T = 2, 42, 4
    specs = [(52, 12, 1040, 2, 0.5 + 0.1 * i) for i in range(16)]
    llr, msg, descs, host = _setup(bg, specs, seed=7)
    alpha = np.full((T, M), 1.0)
    beta = np.zeros((T, M))
    w = np.ones(T)
    d_llr = H.DeviceBuffer.from_array(llr)
    d_msg = H.DeviceBuffer.from_array(msg)
    d_par = H.DeviceBuffer.from_array(np.concatenate([alpha.ravel(), beta.ravel()]).astype(np.float32))
    n = 2 * T * M
    d_grad, d_m, d_v = H.DeviceBuffer(4 * n), H.DeviceBuffer(4 * n), H.DeviceBuffer(4 * n)
    d_m.zero()
    d_v.zero()
    d_ws = H.DeviceBuffer(H.Context.nms_workspace_bytes(bg, descs, T))
    d_loss = H.DeviceBuffer(4 * len(descs))
    losses = []
    for step in range(1, 31):
        d_grad.zero()
        ctx.nms_train_step(bg, descs, T, _View(d_par, 0, 4 * T * M), _View(d_par, 4 * T * M, 4 * T * M), w, 0.125,
                           d_llr, d_msg, d_ws, d_loss, _View(d_grad, 0, 4 * T * M), _View(d_grad, 4 * T * M, 4 * T * M))
        losses.append(float(d_loss.download(np.float32, len(descs)).sum()))
        ctx.nms_adam(d_par, d_grad, d_m, d_v, n, T * M, 0.02, step)
    p = d_par.download(np.float32, n)
    assert losses[-1] < 0.9 * losses[0], losses
    assert (p[:T * M] >= 0).all() and (p[:T * M] <= 511 / 256).all()
