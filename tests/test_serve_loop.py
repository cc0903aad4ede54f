"""Host logic of the slot-deadline serving loop (`ocudu_hip.serve.serve`), on the CPU with
a stand-in server: every transmission completes exactly once, a failed first
transmission is retransmitted (rv2) in the slot HARQ_RTT_SLOTS after its own (or later)
and never past the run, failed retransmissions and late failures release their HARQ
process, and the deterministic mode releases slot k only after the slots whose
retransmissions fall in it have completed."""
import ctypes as C
import types

import pytest

from ocudu_hip import LH_TB_CRC_FAILED, LH_TB_CRC_OK, TbRequest
from ocudu_hip import serve as S


class FakeServer:
    """Completes whole launches in order; the first transmission of TB i fails when
    i % 3 == 0, retransmissions always pass."""

    def __init__(self, *a, **kw):
        self.ready, self.flight, self.log, self.released = [], [], [], []
        self.submitted = {}
        self.step_no = 0
        self.slots = []

    def submit_batch(self, reqs, outs):
        assert len(reqs) == len(outs)
        for q, o in zip(reqs, outs):
            assert o, "a transmission without an output buffer"
            assert q.id not in self.submitted, "transmission submitted twice"
            assert q.slot == self.slots[-1], "a request carries the slot it is sent in"
            self.submitted[q.id] = self.step_no
            self.ready.append(q)

    def step(self):
        self.step_no += 1
        if self.ready and len(self.flight) < 3:
            self.flight.append(self.ready)
            self.ready = []

    def poll(self, wait=False):
        if not self.flight:
            return []
        out = []
        for q in self.flight.pop(0):
            i, tx = divmod(q.id, 2)
            ok = tx == 1 or i % 3 != 0
            out.append(types.SimpleNamespace(id=q.id, tb_crc_ok=LH_TB_CRC_OK if ok else LH_TB_CRC_FAILED,
                                             latency_us=1.0))
        return out

    def harq_release(self, i):
        self.released.append(i)

    def run_slot(self, k):
        assert not self.slots or k > self.slots[-1], "slots run out of order"
        self.slots.append(k)

    def stats(self):
        return types.SimpleNamespace(slots_in_flight=len(self.flight), ready=len(self.ready))

    def close(self):
        pass


class FakeTrace:
    def __init__(self, n, per_slot=3):
        self.tbs = [dict(slot=i // per_slot, tbs=64) for i in range(n)]

    def request(self, i, tx, slot=None):
        return TbRequest(id=2 * i + tx, d_llrs=None, harq_process=i, new_data=int(tx == 0), base_graph=2,
                         tbs_bits=64, modulation_order=2, nof_layers=1, nof_ch_symbols=576, rv=2 * tx, Nref=0,
                         slot=self.tbs[i]["slot"] if slot is None else slot)


@pytest.mark.parametrize("n", [1, 7, 30])
def test_serving_loop_deterministic(monkeypatch, n):
    servers = []

    def make(*a, **kw):
        servers.append(FakeServer())
        return servers[-1]

    monkeypatch.setattr(S, "Server", make)
    trace = FakeTrace(n)
    results, slot_lat, wall, bufs = S.serve(None, trace, realtime=False, timeout_s=30)
    srv = servers[0]
    nslots = max(t["slot"] for t in trace.tbs) + 1
    last_slot = nslots + S.HARQ_RTT_SLOTS + 2
    for i, t in enumerate(trace.tbs):
        assert 2 * i in results, f"TB {i}: first transmission never completed"
        failed = i % 3 == 0
        if failed and t["slot"] + S.HARQ_RTT_SLOTS < last_slot:
            assert 2 * i + 1 in results, f"TB {i}: failed first transmission not retransmitted"
            assert results[2 * i + 1].tb_crc_ok == LH_TB_CRC_OK
            assert 2 * i + 1 in bufs
        elif not failed:
            assert 2 * i + 1 not in results
    # Every submitted transmission completed once; nothing was left in flight.
    assert set(srv.submitted) == set(results)
    assert not srv.flight and not srv.ready
    assert len(slot_lat) > 0 and all(x >= 0 for x in slot_lat)
    # Retransmissions pass here, so no HARQ process is released for a failed one.
    assert all(i % 3 == 0 for i in srv.released)


def test_retransmission_waits_for_the_round_trip(monkeypatch):
    """The retransmission of a TB first sent in slot s is submitted no earlier than the
    submission of slot s + HARQ_RTT_SLOTS."""
    servers = []

    def make(*a, **kw):
        servers.append(FakeServer())
        return servers[-1]

    monkeypatch.setattr(S, "Server", make)
    trace = FakeTrace(30)
    S.serve(None, trace, realtime=False, timeout_s=30)
    srv = servers[0]
    first_of_slot = {}
    for i, t in enumerate(trace.tbs):
        first_of_slot.setdefault(t["slot"], srv.submitted[2 * i])
    for i, t in enumerate(trace.tbs):
        if 2 * i + 1 in srv.submitted:
            due = t["slot"] + S.HARQ_RTT_SLOTS
            if due in first_of_slot:
                assert srv.submitted[2 * i + 1] >= first_of_slot[due]
