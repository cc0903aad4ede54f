"""Slot-deadline serving loop on the lh_server runtime.

Workload: 0.5 ms slots; Poisson(mean 24) UEs per slot, each with one uplink transport block:
MCS uniform over the 256-QAM table (TS38.214 Table 5.1.3.1-2), PRBs uniform in [4, 273],
1..4 layers, 12 data symbols per slot (144 REs per PRB), TBS per TS38.214 5.1.3.2, base
graph per TS38.212 7.2.2; Es/N0 near the plain decoder's 10% BLER point of its MCS.  A TB
that fails is retransmitted with rv2 four slots later and soft-combined in its HARQ
buffers on the device; the HARQ process is released after the retransmission.  All
transmissions' LLRs are generated on the device before the loop (TB generator + codeblock
generator), so the loop itself only serves.  The decoders run at most 6 LDPC iterations,
the reference gNB's PUSCH default (apps/units/flexible_o_du/o_du_low/du_low_config.h:30,
pusch_decoder.h:55), and the per-MCS operating points are found at that count.

    python -m ocudu_hip.serve --slots 2000 [--learned]
"""
import argparse
import ctypes as C
import gc
import json
import math
import sys
import time

import numpy as np

from . import (LH_CRC16, LH_CRC24A, LH_CRC24B, LH_GEN_MSG_INPUT, LH_TB_CRC_OK, LH_TB_NOT_SENT, Context, DeviceBuffer,
               GenDesc, GenTbDesc, ServeTb, Server, TbRequest, msg_bytes, segment_rx)

# TS38.214 Table 5.1.3.1-2 (MCS index table 2, 256-QAM): (Qm, R x 1024).
MCS_256QAM = [(2, 120), (2, 193), (2, 308), (2, 449), (2, 602), (4, 378), (4, 434), (4, 490), (4, 553), (4, 616),
              (4, 658), (6, 466), (6, 517), (6, 567), (6, 616), (6, 666), (6, 719), (6, 772), (6, 822), (6, 873),
              (8, 682.5), (8, 711), (8, 754), (8, 797), (8, 841), (8, 885), (8, 916.5), (8, 948)]
# TS38.214 Table 5.1.3.2-1: TBS for N_info <= 3824.
TBS_TABLE = [24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144, 152, 160, 168, 176, 184, 192, 208,
             224, 240, 256, 272, 288, 304, 320, 336, 352, 368, 384, 408, 432, 456, 480, 504, 528, 552, 576, 608, 640,
             672, 704, 736, 768, 808, 848, 888, 928, 984, 1032, 1064, 1128, 1160, 1192, 1224, 1256, 1288, 1320, 1352,
             1416, 1480, 1544, 1608, 1672, 1736, 1800, 1864, 1928, 2024, 2088, 2152, 2216, 2280, 2408, 2472, 2536, 2600,
             2664, 2728, 2792, 2856, 2976, 3104, 3240, 3368, 3496, 3624, 3752, 3824]
RE_PER_PRB = 144
HARQ_RTT_SLOTS = 4
# HARQ process lifetime after its last transmission (the reference's PUSCH tests use 10 slots).
HARQ_EXPIRE_SLOTS = 10
# Maximum LDPC iterations of the PUSCH decoder (the reference's default).
PUSCH_LDPC_ITERATIONS = 6


def tbs_of(n_prb, qm, r1024, layers):
    """TS38.214 5.1.3.2 with 144 REs per PRB."""
    n_re = min(156, RE_PER_PRB) * n_prb
    r = r1024 / 1024.0
    n_info = n_re * r * qm * layers
    if n_info <= 3824:
        n = max(3, int(math.floor(math.log2(n_info))) - 6)
        ni = max(24, (2 ** n) * int(math.floor(n_info / 2 ** n)))
        return next(t for t in TBS_TABLE if t >= ni)
    n = int(math.floor(math.log2(n_info - 24))) - 5
    ni = max(3840, (2 ** n) * int(round((n_info - 24) / 2 ** n)))
    if r <= 0.25:
        c = math.ceil((ni + 24) / 3816)
        return 8 * c * math.ceil((ni + 24) / (8 * c)) - 24
    if ni > 8424:
        c = math.ceil((ni + 24) / 8424)
        return 8 * c * math.ceil((ni + 24) / (8 * c)) - 24
    return 8 * math.ceil((ni + 24) / 8) - 24


def base_graph_of(tbs, r):
    """TS38.212 7.2.2."""
    return 2 if (tbs <= 292 or (tbs <= 3824 and r <= 0.67) or r <= 0.25) else 1


def mcs_thresholds(ctx, n=128, log=None, iters=PUSCH_LDPC_ITERATIONS):
    """Es/N0 where the plain decoder's BLER is 10% for one codeblock of each MCS (Z = 384,
    or the BG2 equivalent at low rate), at `iters` LDPC iterations."""
    from .nms import rm_length
    from .workloads import snr_for_bler
    out = []
    for i, (qm, r1024) in enumerate(MCS_256QAM):
        r = r1024 / 1024.0
        bg = 2 if r <= 0.25 else 1
        Z = 384
        E = rm_length(bg, Z, r, qm)
        s = snr_for_bler(ctx, bg, Z, E, qm, 0.10, n=n, iters=iters)
        out.append(s)
        if log:
            log(f"MCS {i}: Qm {qm} R {r:.3f} BG{bg}: 10% BLER at {s:.2f} dB")
    return out


class Trace:
    """The TBs of a run with all their transmissions' LLRs on the device."""

    def __init__(self, ctx, slots, mean_ues, thresholds, seed=0, snr_lo=-0.2, snr_hi=1.0, prb_max=273,
                 max_cbs_per_tb=512, carrier_prbs=0, layers=(1, 4)):
        """carrier_prbs > 0: the slot's UEs share one carrier of that many PRBs (a random
        split, at least 4 PRBs each, so fewer UEs when the carrier is full); otherwise each
        UE draws its PRBs in [4, prb_max].  Layers uniform in the `layers` range."""
        rng = np.random.default_rng(seed)
        self.tbs = []  # dicts
        for k in range(slots):
            n_ue = int(rng.poisson(mean_ues))
            if carrier_prbs:
                n_ue = max(1, min(n_ue, carrier_prbs // 4))
                cuts = np.sort(rng.choice(np.arange(1, carrier_prbs // 4), n_ue - 1, replace=False)) if n_ue > 1 else []
                edges = np.concatenate([[0], cuts, [carrier_prbs // 4]]).astype(int)
                prbs = [4 * int(b - a) for a, b in zip(edges[:-1], edges[1:])]
                prbs[-1] += carrier_prbs % 4
            for u in range(n_ue):
                mcs = int(rng.integers(0, len(MCS_256QAM)))
                qm, r1024 = MCS_256QAM[mcs]
                prb = prbs[u] if carrier_prbs else int(rng.integers(4, prb_max + 1))
                nl = int(rng.integers(layers[0], layers[1] + 1))
                tbs = tbs_of(prb, qm, r1024, nl)
                bg = base_graph_of(tbs, r1024 / 1024.0)
                nsym = RE_PER_PRB * prb * nl
                seg, rm, off = segment_rx(bg, tbs, qm, nl, nsym, max_cbs_per_tb)
                snr = thresholds[mcs] + float(rng.uniform(snr_lo, snr_hi))
                self.tbs.append(dict(slot=k, mcs=mcs, qm=qm, prb=prb, layers=nl, tbs=tbs, bg=bg, nsym=nsym,
                                     seg=seg, rm=rm, off=off, snr=snr, seed=int(rng.integers(1, 2 ** 62)),
                                     noise=[int(rng.integers(1, 2 ** 62)) for _ in range(2)]))
        self._generate(ctx)

    def _generate(self, ctx):
        # Offsets: TB bytes, codeblock messages, and the LLRs of both transmissions.
        tb_off = msg_off = llr_off = 0
        for t in self.tbs:
            seg = t["seg"]
            t["tb_off"] = tb_off
            t["msg_off"] = msg_off
            t["stride"] = msg_bytes(t["bg"], seg.lifting_size)
            t["llr_off"] = [llr_off, llr_off + ((t["nsym"] * t["qm"] + 255) // 256) * 256]
            tb_off += t["tbs"] // 8
            msg_off += seg.nof_cbs * t["stride"]
            llr_off = t["llr_off"][1] + ((t["nsym"] * t["qm"] + 255) // 256) * 256
        self.d_tb = DeviceBuffer(tb_off)
        self.d_llr = DeviceBuffer(llr_off)
        d_msg = DeviceBuffer(msg_off)
        gt = [GenTbDesc(t["tb_off"], t["msg_off"], t["seed"], t["stride"], t["tbs"], t["seg"].nof_cbs,
                        t["seg"].segment_length, t["seg"].nof_filler_bits) for t in self.tbs]
        for c0 in range(0, len(gt), 4096):
            ctx.gen_transport_blocks(gt[c0:c0 + 4096], self.d_tb, d_msg)
        gens = []
        for t in self.tbs:
            seg = t["seg"]
            C = seg.nof_cbs
            poly = LH_CRC24B if C > 1 else (LH_CRC16 if t["tbs"] <= 3824 else LH_CRC24A)
            n0 = float(10.0 ** (-t["snr"] / 10.0))
            for tx, rv in enumerate((0, 2)):
                for c in range(C):
                    gens.append(GenDesc(t["llr_off"][tx] + t["off"][c], t["msg_off"] + c * t["stride"], 0,
                                        t["noise"][tx] + c, n0, t["rm"][c], t["bg"], seg.lifting_size, rv, t["qm"], 0,
                                        seg.nof_filler_bits, poly, LH_GEN_MSG_INPUT, 0))
        for c0 in range(0, len(gens), 4096):
            ctx.gen_codeblocks(gens[c0:c0 + 4096], self.d_llr, d_msg)
        from . import hip
        hip().hipDeviceSynchronize()
        d_msg.free()
        self.tb_host = self.d_tb.download(np.uint8, max(1, tb_off))

    def request(self, i, tx, slot=None):
        """Transmission tx of TB i, sent in `slot` (default: the TB's own slot)."""
        t = self.tbs[i]
        return TbRequest(id=2 * i + tx, d_llrs=self.d_llr.ptr + t["llr_off"][tx], harq_process=i, new_data=int(tx == 0),
                         base_graph=t["bg"], tbs_bits=t["tbs"], modulation_order=t["qm"], nof_layers=t["layers"],
                         nof_ch_symbols=t["nsym"], rv=2 * tx, Nref=0, slot=t["slot"] if slot is None else slot)


def serve(ctx, trace, learned=False, slot_us=500.0, harq_budget=8 << 30, max_cbs_per_slot=4096, slots_in_flight=3,
          iters=PUSCH_LDPC_ITERATIONS, realtime=True, log=None, timeout_s=600.0, expire_slots=HARQ_EXPIRE_SLOTS,
          native=False):
    """Runs the trace through a server: returns per-transmission results and per-slot
    latencies.  realtime: slots are released every slot_us; else as fast as possible.
    HARQ processes expire expire_slots after their last transmission's slot (0: never).
    native: the slot loop runs in the library (lh_server_serve) instead of here."""
    srv = Server(ctx, harq_budget, max_cbs_per_slot, slots_in_flight, iters, learned=learned,
                 nof_fallback_iterations=iters, expire_timeout_slots=expire_slots)
    if native:
        try:
            return _serve_native(srv, trace, slot_us, realtime, timeout_s)
        finally:
            srv.close()
    by_slot = {}
    for i, t in enumerate(trace.tbs):
        by_slot.setdefault(t["slot"], []).append(i)
    nslots = max(by_slot) + 1 if by_slot else 0
    tx_slot = {}
    out_bufs = {}
    # New transmissions' requests and TB buffers prepared ahead (the loop only submits).
    prepared = {}
    for k, idx in by_slot.items():
        for i in idx:
            out_bufs[2 * i] = np.zeros(trace.tbs[i]["tbs"] // 8, np.uint8)
            tx_slot[2 * i] = k
        prepared[k] = ((TbRequest * len(idx))(*[trace.request(i, 0) for i in idx]),
                       (C.c_void_p * len(idx))(*[out_bufs[2 * i].ctypes.data for i in idx]))
    # No garbage-collector pauses inside the slot loop: the objects made so far (the trace)
    # are frozen out of the collector's scans and collection resumes after the loop.
    gc.collect()
    gc.freeze()
    gc_was_enabled = gc.isenabled()
    gc.disable()
    try:
        results, slot_done, slot_start, wall = _serve_loop(srv, trace, prepared, by_slot, out_bufs, tx_slot, slot_us,
                                                           realtime, timeout_s, nslots)
    finally:
        if gc_was_enabled:
            gc.enable()
        gc.unfreeze()
    srv.close()
    slot_lat = [(slot_done[s] - slot_start[s]) * 1e6 for s in slot_done if s in slot_start]
    return results, slot_lat, wall, out_bufs


def _serve_native(srv, trace, slot_us, realtime, timeout_s):
    """The same loop in the library: every transmission's request and TB buffer prepared
    ahead, one call."""
    n = len(trace.tbs)
    bufs = [(np.zeros(t["tbs"] // 8, np.uint8), np.zeros(t["tbs"] // 8, np.uint8)) for t in trace.tbs]
    arr = (ServeTb * n)()
    for i, t in enumerate(trace.tbs):
        arr[i].tx[0] = trace.request(i, 0)
        arr[i].tx[1] = trace.request(i, 1)
        arr[i].tb_out[0] = bufs[i][0].ctypes.data
        arr[i].tb_out[1] = bufs[i][1].ctypes.data
        arr[i].slot = t["slot"]
    nslots = max(t["slot"] for t in trace.tbs) + 1 if n else 1
    res, lat, wall = srv.serve(arr, nslots, HARQ_RTT_SLOTS, slot_us=slot_us, realtime=realtime, timeout_s=timeout_s)
    results, out_bufs = {}, {}
    for k, r in enumerate(res):
        if r.tb_crc_ok != LH_TB_NOT_SENT:
            results[k] = r
            out_bufs[k] = bufs[k // 2][k % 2]
    return results, [x for x in lat if x >= 0], wall, out_bufs


def _serve_loop(srv, trace, prepared, by_slot, out_bufs, tx_slot, slot_us, realtime, timeout_s, nslots):
    retx = {}  # slot -> [(TB index, request)] of the retransmissions due in it (built ahead)
    results = {}
    slot_done = {}
    slot_start = {}
    outstanding = {}  # slot -> number of transmissions not completed (open slots only)
    t0 = time.perf_counter()
    k = 0
    last_slot = nslots + HARQ_RTT_SLOTS + 2

    def complete(r):
        results[r.id] = r
        s = tx_slot[r.id]
        outstanding[s] -= 1
        if outstanding[s] == 0:
            del outstanding[s]
            slot_done[s] = time.perf_counter()
        i, tx = divmod(r.id, 2)
        if tx == 0 and r.tb_crc_ok != LH_TB_CRC_OK:
            due_slot = max(s + HARQ_RTT_SLOTS, k)
            if due_slot < last_slot:
                # The retransmission's request and TB buffer are made now, not at its slot.
                out_bufs[2 * i + 1] = np.zeros(trace.tbs[i]["tbs"] // 8, np.uint8)
                retx.setdefault(due_slot, []).append((i, trace.request(i, 1, slot=due_slot)))
            else:
                srv.harq_release(i)  # past the run: no retransmission
        elif tx == 1 and r.tb_crc_ok != LH_TB_CRC_OK:
            srv.harq_release(i)

    while k < last_slot or outstanding:
        if realtime:
            # Completions first: a slot boundary's submission does not delay them.
            for r in srv.poll():
                complete(r)
        now = time.perf_counter()
        # Without the clock, slot k is released once the slots whose retransmissions are due
        # in it have completed (the HARQ round trip), so the run is deterministic.
        due = realtime or not outstanding or min(outstanding) >= k - HARQ_RTT_SLOTS + 1
        if not realtime and not due:
            # Requests waiting for HARQ memory with nothing in flight: only time frees it
            # (a failed TB holds its buffers until its retransmission), so release a slot.
            st = srv.stats()
            due = st.slots_in_flight == 0 and st.ready == 0
        if now - t0 > timeout_s:
            srv.close()
            raise RuntimeError(f"serving loop exceeded {timeout_s} s at slot {k}")
        if k < last_slot and ((realtime and now >= t0 + k * slot_us * 1e-6) or (not realtime and due)):
            srv.run_slot(k)  # HARQ expiry at the slot boundary
            new = prepared.pop(k, None)
            again = retx.pop(k, [])
            if new is not None or again:
                n = 0
                if new is not None:
                    reqs, outs = new
                    n += len(reqs)
                if again:
                    rreqs = (TbRequest * len(again))(*[q for _, q in again])
                    routs = (C.c_void_p * len(again))(*[out_bufs[2 * i + 1].ctypes.data for i, _ in again])
                    for i, _ in again:
                        tx_slot[2 * i + 1] = k
                    n += len(again)
                slot_start[k] = time.perf_counter()
                if new is not None:
                    srv.submit_batch(reqs, outs)
                if again:
                    srv.submit_batch(rreqs, routs)
                outstanding[k] = outstanding.get(k, 0) + n
            k += 1
        srv.step()
        for r in srv.poll(wait=not realtime and (k >= last_slot or not due)):
            complete(r)
    wall = time.perf_counter() - t0
    return results, slot_done, slot_start, wall


def summarize(trace, results, slot_lat, wall, out_bufs):
    lat = np.array([r.latency_us for r in results.values()])
    ok_bits = 0
    first_fail = final_fail = 0
    wrong = 0
    for i, t in enumerate(trace.tbs):
        r0 = results.get(2 * i)
        r1 = results.get(2 * i + 1)
        final = r1 if r1 is not None else r0
        if r0 is not None and r0.tb_crc_ok != LH_TB_CRC_OK:
            first_fail += 1
        if final is None or final.tb_crc_ok != LH_TB_CRC_OK:
            final_fail += 1
            continue
        got = out_bufs[final.id]
        exp = trace.tb_host[t["tb_off"]:t["tb_off"] + t["tbs"] // 8]
        if not np.array_equal(got, exp):
            wrong += 1
        else:
            ok_bits += t["tbs"]
    sl = np.array(slot_lat) if slot_lat else np.zeros(1)
    n = len(trace.tbs)
    return {"tbs": n, "transmissions": len(results), "initial_bler": round(first_fail / max(n, 1), 4),
            "residual_bler": round(final_fail / max(n, 1), 4), "wrong_tbs": wrong,
            "info_gbps": round(ok_bits / wall / 1e9, 3), "wall_s": round(wall, 3),
            "tb_latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                              "p99": round(float(np.percentile(lat, 99)), 1)},
            "slot_latency_us": {"p50": round(float(np.percentile(sl, 50)), 1),
                                "p99": round(float(np.percentile(sl, 99)), 1), "budget_us": 500}}


def _main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--slots", type=int, default=2000)
    ap.add_argument("--ues", type=float, default=24.0)
    ap.add_argument("--learned", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--prb-max", type=int, default=273, help="PRBs per UE uniform in [4, prb-max]")
    ap.add_argument("--in-flight", type=int, default=3)
    ap.add_argument("--carrier-prbs", type=int, default=0,
                    help="> 0: the slot's UEs share one carrier of that many PRBs (instead of --prb-max each)")
    ap.add_argument("--layers", default="1,4", help="layers per UE, uniform in this range")
    ap.add_argument("--native", action="store_true", help="the slot loop in the library (lh_server_serve)")
    args = ap.parse_args(argv)
    layers = tuple(int(x) for x in args.layers.split(","))
    ctx = Context(0, max_codeblocks=8192)

    def log(m):
        print(m, flush=True)

    thr = mcs_thresholds(ctx, log=log)
    t0 = time.time()
    trace = Trace(ctx, args.slots, args.ues, thr, seed=args.seed, prb_max=args.prb_max, carrier_prbs=args.carrier_prbs,
                  layers=layers)
    log(f"trace generated in {time.time() - t0:.1f} s")
    log(f"{len(trace.tbs)} TBs, {sum(t['seg'].nof_cbs for t in trace.tbs)} codeblocks")
    if args.learned:
        from pathlib import Path
        from .nms import load_table
        tables = Path(__file__).resolve().parent.parent / "tables"
        for bg in (1, 2):
            _, aq, bq, _ = load_table(tables / f"nms_bg{bg}.json")
            ctx.set_scale_table(bg, aq, bq)
    res = serve(ctx, trace, learned=args.learned, slots_in_flight=args.in_flight, log=log, native=args.native)
    print(json.dumps(summarize(trace, *res)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())
