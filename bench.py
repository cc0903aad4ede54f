#!/usr/bin/env python3
"""Throughput / latency benchmark of the HIP PUSCH LDPC receive chain on one MI355X.

Workload (one step): a homogeneous batch of 16 slots x 152 codeblocks, BG1 Z=384,
256-QAM, E = 9088 (code rate ~0.93), new data rv0, int8 LLRs resident in HBM, max 8
LDPC iterations with CRC24B early stop: rate dematching into the HARQ buffers followed by
layered min-sum decoding (lh_pusch_decode_batch).  AWGN Es/N0 is swept; the headline
value is the info Gbit/s (CRC-passing codeblocks x 8424 info bits / step time) at the
lowest swept Es/N0 whose BLER is <= 10%.

Prints one JSON line (see the fields below).  --gpus N with N > 1 exits: the build
targets a single GPU.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "ocudu-dapp-platform_amd"))

METRIC = "info Gbit/s @BLER≤10% (plain vs learned min-sum) + p99 slot decode latency"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--slots", type=int, default=16)
    p.add_argument("--cbs-per-slot", type=int, default=152)
    p.add_argument("--snr", type=str, default="24,25,26", help="Es/N0 points (dB); the full sweep is 20..26")
    p.add_argument("--max-iter", type=int, default=8)
    p.add_argument("--latency-slots", type=int, default=2000)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU work of the baseline sample")
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--copy-probe", type=int, default=0,
                   help="MiB copied device-to-device once (calibrates profiler byte counters)")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import ocudu_hip as H
    from ocudu_hip import datagen as D

    bg, Z, E, Qm = 1, 384, 9088, 8
    K = 22 * Z
    info_bits = K - 24
    N = 66 * Z
    n_cb = a.slots * a.cbs_per_slot
    msg_b = (K + 7) // 8
    snrs = [float(s) for s in a.snr.split(",")]

    ctx = H.Context(0, max(n_cb, 4096))
    stream = H.Stream()
    d_rm = H.DeviceBuffer(n_cb * N)
    d_bits = H.DeviceBuffer(n_cb * msg_b)
    d_res = H.DeviceBuffer(4 * n_cb)
    cbs = (H.PuschCbDesc * n_cb)()
    for i in range(n_cb):
        cbs[i] = H.PuschCbDesc(H.RmDesc(i * E, i * N, E, bg, Z, 0, Qm, 0, 0), i * msg_b, 24, H.LH_CRC24B)
    ev0, ev1 = H.Event(), H.Event()

    if a.copy_probe:
        nb = a.copy_probe << 20
        src, dst = H.DeviceBuffer(nb), H.DeviceBuffer(nb)
        src.zero()
        H._chk(H.hip().hipMemcpy(dst.ptr, src.ptr, nb, H.D2D), "probe copy")
        H.hip().hipDeviceSynchronize()
        src.free()
        dst.free()

    sweep = []
    data = {}
    for snr in snrs:
        t0 = time.time()
        _, llr = D.make_batch(n_cb, bg, Z, E, Qm, snr, seed=a.seed + int(snr * 10), n_unique=a.cbs_per_slot)
        d_cw = H.DeviceBuffer.from_array(llr.reshape(-1))
        gen_s = time.time() - t0
        for _ in range(a.warmup):
            ctx.pusch_decode_batch(None, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=a.max_iter, stream=stream,
                                   arr=cbs)
        stream.sync()
        ev0.record(stream)
        for _ in range(a.steps):
            ctx.pusch_decode_batch(None, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=a.max_iter, stream=stream,
                                   arr=cbs)
        ev1.record(stream)
        ev1.sync()
        ms = ev0.elapsed_ms(ev1) / a.steps
        res = d_res.download(np.int32, n_cb)
        ok = int((res > 0).sum())
        bler = 1.0 - ok / n_cb
        gbps = ok * info_bits / (ms * 1e-3) / 1e9
        mean_it = float(res[res > 0].mean()) if ok else 0.0
        sweep.append({"esn0_db": snr, "bler": round(bler, 5), "info_gbps": round(gbps, 3), "ms_per_step": round(ms, 4),
                      "mean_iters": round(mean_it, 3), "gen_s": round(gen_s, 1)})
        data[snr] = (d_cw, llr, res)
        print(f"# Es/N0 {snr:5.1f} dB  BLER {bler:.4f}  {gbps:9.2f} info Gbit/s  {ms:.3f} ms/step  "
              f"mean iters {mean_it:.2f}", file=sys.stderr, flush=True)

    qual = [s for s in sweep if s["bler"] <= 0.10]
    head = min(qual, key=lambda s: s["esn0_db"]) if qual else max(sweep, key=lambda s: s["esn0_db"])
    snr_h = head["esn0_db"]

    # Roofline of the dominant kernel (the fused dematch + decode launch of one step):
    # algorithmic bytes per codeblock = E LLRs read + N HARQ soft bits written + packed
    # message bits written + 4-byte result; duration = the HIP-event time of the step.
    d_cw, llr_h, _ = data[snr_h]
    dec_ms = head["ms_per_step"]
    dec_bytes = n_cb * (E + N + msg_b + 4)
    achieved = dec_bytes / (dec_ms * 1e-3) / 1e9
    traffic = measured_traffic()

    # Slot latency: one slot (152 CBs) per call, launch -> host-visible CRC verdicts.
    slot_arrs = []
    for s in range(a.slots):
        arr = (H.PuschCbDesc * a.cbs_per_slot)()
        for i in range(a.cbs_per_slot):
            arr[i] = cbs[s * a.cbs_per_slot + i]
        slot_arrs.append(arr)
    lat = []
    res_host = np.empty(a.cbs_per_slot, np.int32)
    hipm = H.hip()
    for k in range(a.latency_slots):
        t0 = time.perf_counter()
        ctx.pusch_decode_batch(None, d_cw, d_rm, d_bits, d_res, nof_ldpc_iterations=a.max_iter, stream=stream,
                               arr=slot_arrs[k % a.slots])
        hipm.hipMemcpyAsync(res_host.ctypes.data, d_res.ptr, res_host.nbytes, H.D2H, stream.handle)
        stream.sync()
        lat.append((time.perf_counter() - t0) * 1e6)
    lat = np.array(lat)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    print(f"# slot latency p50 {p50:.1f} us  p99 {p99:.1f} us over {a.latency_slots} slots", file=sys.stderr,
          flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(d_rm, n_cb, N, info_bits, a.max_iter, a.cpu_seconds, snr_h)

    out = {
        "metric": METRIC,
        "value": round(head["info_gbps"], 3),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"homogeneous slot batch: {a.slots} slots x {a.cbs_per_slot} CBs, BG1 Z=384, "
                               f"256QAM, E={E} (R~0.93), rv0 new data, max {a.max_iter} iters, CRC24B early stop",
                   "global_batch": n_cb, "esn0_db": snr_h, "decoder": "plain (bit-exact generic min-sum)"},
        "bler": head["bler"],
        "mean_iters": head["mean_iters"],
        "slot_latency_us": {"p50": round(p50, 1), "p99": round(p99, 1), "slots": a.latency_slots,
                            "cbs_per_slot": a.cbs_per_slot, "budget_us": 500},
        "snr_sweep": sweep,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_ldpc_decode<1> (fused dematch + decode)", "kernel_ms": round(dec_ms, 4),
                     "bytes_per_launch": dec_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def measured_traffic():
    """HBM bytes per launch of the step's kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy kernel); None when absent."""
    path = ROOT / "profiles" / "k_ldpc_decode_bg1_z384_slot_batch.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1].strip())
    return None


def cpu_baseline(d_rm, n_cb, N, info_bits, max_iter, cpu_seconds, snr):
    """The reference decoder (AVX-512 / AVX2, built from its sources into oracle/_ref)
    on the host cores, over the same dematched int8 soft bits; the plain-C oracle when
    the reference build is absent."""
    import ctypes as C

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    sample = min(n_cb, 608)
    rm = d_rm.download(np.int8, sample * N).reshape(sample, N)
    R = O.ref()
    threads = max(1, min(16, os.cpu_count() or 1))
    if R is not None:
        kind = R.ref_decoder_kind(2)
        secs_wall = max(1.0, cpu_seconds / threads)
        done, ok, its = C.c_ulong(), C.c_ulong(), C.c_ulong()
        secs = R.ref_decode_throughput(2, threads, O._p(rm, O._i8p), sample, N, 1, 384, 0, max_iter, O.CRC24B,
                                       secs_wall, C.byref(done), C.byref(ok), C.byref(its))
        gbps = ok.value * info_bits / secs / 1e9
        name = {0: "generic", 1: "avx2", 2: "avx512"}[kind]
        return {"value": round(gbps, 3), "unit": "Gbit/s", "cores": threads, "kind": "reference",
                "sample": f"{done.value} decodes of {sample} distinct CBs (Es/N0 {snr} dB) in {secs:.1f} s by "
                          f"the reference ldpc_decoder_{name}, one decoder per thread",
                "us_per_cb_per_core": round(secs * threads / max(done.value, 1) * 1e6, 2)}
    t0 = time.time()
    done = ok = 0
    while time.time() - t0 < cpu_seconds:
        r, _ = O.decode(rm[done % sample], 1, 384, max_iter=max_iter)
        ok += r > 0
        done += 1
    secs = time.time() - t0
    return {"value": round(ok * info_bits / secs / 1e9, 4), "unit": "Gbit/s", "cores": 1, "kind": "port",
            "sample": f"{done} decodes by the C oracle in {secs:.1f} s"}


if __name__ == "__main__":
    main()
