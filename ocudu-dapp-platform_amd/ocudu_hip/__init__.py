"""Python access to the HIP PUSCH LDPC receive chain (C ABI of include/ocudu_hip_ldpc.h).

The library is the product; this module is a thin ctypes binding used by tests and
bench.py.  Device memory, streams and events come straight from the HIP runtime
(libamdhip64) -- no torch.  Loading fails loudly when the built library is missing.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "lib" / "libocudu_hip_ldpc.so"

LH_CRC24A, LH_CRC24B, LH_CRC16 = 0x1864CFB, 0x1800063, 0x11021
LH_FORCE_DECODING, LH_EARLY_STOP_SYNDROME = 1, 2

_hip = None
_lib = None


class HipError(RuntimeError):
    pass


class LhError(RuntimeError):
    pass


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL(os.environ.get("LH_HIP_RUNTIME", "libamdhip64.so"))
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamWaitEvent.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        h.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
        h.hipSetDevice.argtypes = [C.c_int]
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostFree.argtypes = [C.c_void_p]
        _hip = h
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise HipError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


H2D, D2H, D2D = 1, 2, 3


def device_count():
    n = C.c_int(0)
    try:
        rc = hip().hipGetDeviceCount(C.byref(n))
    except OSError:
        return 0
    return n.value if rc == 0 else 0


class DeviceBuffer:
    """A hipMalloc'd byte buffer."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        _chk(hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _chk(hip().hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, count, offset=0):
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _chk(hip().hipMemcpy(out.ctypes.data, self.ptr + offset, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        _chk(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _chk(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def sync(self):
        _chk(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def wait(self, event):
        """Work queued on this stream after the call waits for `event`."""
        _chk(hip().hipStreamWaitEvent(self.handle, event.handle, 0), "hipStreamWaitEvent")


class Event:
    def __init__(self):
        e = C.c_void_p()
        _chk(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _chk(hip().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def sync(self):
        _chk(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = C.c_float()
        _chk(hip().hipEventElapsedTime(C.byref(ms), self.handle, later.handle), "hipEventElapsedTime")
        return ms.value


# ---- C ABI structures ------------------------------------------------------------
class DecoderConfig(C.Structure):
    _fields_ = [("base_graph", C.c_uint32), ("lifting_size", C.c_uint32), ("nof_filler_bits", C.c_uint32),
                ("nof_crc_bits", C.c_uint32), ("max_iterations", C.c_uint32)]


class DecodeDesc(C.Structure):
    _fields_ = [("llr_offset", C.c_uint64), ("bits_offset", C.c_uint64), ("nof_llrs", C.c_uint32),
                ("crc_poly", C.c_uint32), ("cfg", DecoderConfig)]


class RmDesc(C.Structure):
    _fields_ = [("cw_offset", C.c_uint64), ("rm_offset", C.c_uint64), ("rm_length", C.c_uint32),
                ("base_graph", C.c_uint32), ("lifting_size", C.c_uint32), ("rv", C.c_uint32),
                ("modulation_order", C.c_uint32), ("Nref", C.c_uint32), ("nof_filler_bits", C.c_uint32)]


class PuschCbDesc(C.Structure):
    _fields_ = [("rm", RmDesc), ("bits_offset", C.c_uint64), ("nof_crc_bits", C.c_uint32),
                ("crc_poly", C.c_uint32)]


class TbDesc(C.Structure):
    _fields_ = [("tb_offset", C.c_uint64), ("cb_bits_offset", C.c_uint64), ("cb_bits_stride", C.c_uint32),
                ("tbs_bits", C.c_uint32), ("first_cb", C.c_uint32), ("nof_cbs", C.c_uint32),
                ("cb_data_bits", C.c_uint32), ("pad", C.c_uint32)]


class GenDesc(C.Structure):
    _fields_ = [("llr_offset", C.c_uint64), ("msg_offset", C.c_uint64), ("seed_msg", C.c_uint64),
                ("seed_noise", C.c_uint64), ("n0", C.c_double), ("rm_length", C.c_uint32), ("base_graph", C.c_uint32),
                ("lifting_size", C.c_uint32), ("rv", C.c_uint32), ("modulation_order", C.c_uint32),
                ("Nref", C.c_uint32), ("nof_filler_bits", C.c_uint32), ("crc_poly", C.c_uint32),
                ("flags", C.c_uint32), ("pad", C.c_uint32)]


class EncodeDesc(C.Structure):
    _fields_ = [("msg_offset", C.c_uint64), ("out_offset", C.c_uint64), ("cw_offset", C.c_uint64),
                ("rm_length", C.c_uint32), ("base_graph", C.c_uint32), ("lifting_size", C.c_uint32),
                ("rv", C.c_uint32), ("modulation_order", C.c_uint32), ("Nref", C.c_uint32),
                ("nof_filler_bits", C.c_uint32), ("pad", C.c_uint32)]


class DemodDesc(C.Structure):
    """One block of lh_pusch_demodulate (lh_demod_desc)."""
    _fields_ = [("symbol_offset", C.c_uint64), ("llr_offset", C.c_uint64), ("nof_symbols", C.c_uint32),
                ("modulation", C.c_uint32), ("c_init", C.c_uint32), ("seq_offset", C.c_uint32)]


DEMUX_FIELDS = ("modulation", "nof_layers", "nof_prb", "start_symbol_index", "nof_symbols", "dmrs_symbol_mask",
                "dmrs_type", "nof_cdm_groups_without_data", "nof_harq_ack_rvd", "nof_harq_ack_bits",
                "nof_enc_harq_ack_bits", "nof_csi_part1_bits", "nof_enc_csi_part1_bits", "nof_csi_part2_bits",
                "nof_enc_csi_part2_bits", "c_init")


class UlschDemuxDesc(C.Structure):
    """One codeword of lh_ulsch_demultiplex (lh_ulsch_demux_desc)."""
    _fields_ = [(f, C.c_uint64) for f in ("llr_offset", "sch_offset", "harq_ack_offset", "csi_part1_offset",
                                           "csi_part2_offset")] + [(f, C.c_uint32) for f in DEMUX_FIELDS]


class UlschDemuxSizes(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("nof_llrs", "nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")]


class UlschConfig(C.Structure):
    """lh_ulsch_config (ulsch_configuration)."""
    _fields_ = [("tbs_bits", C.c_uint32), ("target_code_rate", C.c_float), ("modulation", C.c_uint32),
                ("nof_harq_ack_bits", C.c_uint32), ("nof_csi_part1_bits", C.c_uint32),
                ("nof_csi_part2_bits", C.c_uint32), ("alpha_scaling", C.c_float), ("beta_offset_harq_ack", C.c_float),
                ("beta_offset_csi_part1", C.c_float), ("beta_offset_csi_part2", C.c_float), ("nof_rb", C.c_uint32),
                ("start_symbol_index", C.c_uint32), ("nof_symbols", C.c_uint32), ("dmrs_type", C.c_uint32),
                ("dmrs_symbol_mask", C.c_uint32), ("nof_cdm_groups_without_data", C.c_uint32),
                ("nof_layers", C.c_uint32), ("contains_dc", C.c_int32)]


class UlschInfo(C.Structure):
    """lh_ulsch_info (ulsch_information)."""
    _fields_ = [(f, C.c_uint32) for f in (
        "sch_base_graph", "sch_tb_crc_size", "sch_nof_cb", "sch_lifting_size", "sch_nof_bits_per_cb",
        "sch_nof_filler_bits_per_cb", "nof_ul_sch_bits", "nof_harq_ack_bits", "nof_harq_ack_rvd",
        "nof_csi_part1_bits", "nof_csi_part2_bits", "nof_harq_ack_re", "nof_csi_part1_re", "nof_csi_part2_re",
        "nof_dc_overlap_bits", "pad")]


def ulsch_information(**cfg):
    """get_ulsch_information of a grant (lh_ulsch_information), as a dict."""
    info = UlschInfo()
    _lh(lib().lh_ulsch_information(C.byref(UlschConfig(**cfg)), C.byref(info)), "lh_ulsch_information")
    return {f: getattr(info, f) for f, _ in UlschInfo._fields_ if f != "pad"}


def ulsch_demux_sizes(desc):
    """Soft-bit counts of one codeword's demultiplexed fields (lh_ulsch_demux_get_sizes)."""
    sz = UlschDemuxSizes()
    _lh(lib().lh_ulsch_demux_get_sizes(C.byref(desc), C.byref(sz)), "lh_ulsch_demux_get_sizes")
    return {f: getattr(sz, f) for f, _ in UlschDemuxSizes._fields_}


LH_MOD_PI_2_BPSK, LH_MOD_BPSK, LH_MOD_QPSK, LH_MOD_QAM16, LH_MOD_QAM64, LH_MOD_QAM256 = 0, 1, 2, 4, 6, 8

LH_GEN_MSG_INPUT = 1


class GenTbDesc(C.Structure):
    _fields_ = [("tb_offset", C.c_uint64), ("msg_offset", C.c_uint64), ("seed", C.c_uint64),
                ("msg_stride", C.c_uint32), ("tbs_bits", C.c_uint32), ("nof_cbs", C.c_uint32),
                ("segment_length", C.c_uint32), ("nof_filler_bits", C.c_uint32), ("pad", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("harq_budget_bytes", C.c_size_t), ("max_cbs_per_slot", C.c_uint), ("max_slots_in_flight", C.c_uint),
                ("nof_ldpc_iterations", C.c_uint), ("nof_fallback_iterations", C.c_uint), ("learned", C.c_int),
                ("expire_timeout_slots", C.c_uint)]


class TbRequest(C.Structure):
    _fields_ = [("id", C.c_uint64), ("d_llrs", C.c_void_p), ("harq_process", C.c_uint32), ("new_data", C.c_int32),
                ("base_graph", C.c_uint32), ("tbs_bits", C.c_uint32), ("modulation_order", C.c_uint32),
                ("nof_layers", C.c_uint32), ("nof_ch_symbols", C.c_uint32), ("rv", C.c_uint32), ("Nref", C.c_uint32),
                ("slot", C.c_uint32)]


class TbResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("harq_process", C.c_uint32), ("tb_crc_ok", C.c_int32), ("nof_cbs", C.c_uint32),
                ("nof_cbs_ok", C.c_uint32), ("max_iterations", C.c_uint32), ("pad", C.c_uint32),
                ("latency_us", C.c_double)]


class ServeTb(C.Structure):
    _fields_ = [("tx", TbRequest * 2), ("tb_out", C.c_void_p * 2), ("slot", C.c_uint32), ("pad", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("slot_us", C.c_double), ("timeout_s", C.c_double), ("nof_slots", C.c_uint),
                ("harq_rtt_slots", C.c_uint), ("realtime", C.c_int), ("pad", C.c_uint)]


class ServerStats(C.Structure):
    _fields_ = [("harq_free_bytes", C.c_uint64), ("harq_processes", C.c_uint32), ("pending", C.c_uint32),
                ("ready", C.c_uint32), ("slots_in_flight", C.c_uint32)]


LH_TB_CRC_OK, LH_TB_CB_FAILED, LH_TB_CRC_FAILED, LH_TB_REJECTED, LH_TB_NOT_SENT = 1, 0, -1, -2, -3

# The 51 NR lifting sizes (TS38.212 Table 5.3.2-1).
LIFTING_SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40,
                 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240,
                 256, 288, 320, 352, 384]


class Segmentation(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("nof_cbs", "lifting_size", "segment_length", "nof_filler_bits",
                                           "nof_tb_crc_bits", "nof_cb_crc_bits", "cb_data_bits", "zero_pad")]


class TrainDesc(C.Structure):
    _fields_ = [("llr_offset", C.c_uint64), ("msg_offset", C.c_uint64), ("nof_llrs", C.c_uint32),
                ("lifting_size", C.c_uint32), ("nof_filler_bits", C.c_uint32), ("nof_layers", C.c_uint32)]


EXPORTED = ["lh_ctx_create", "lh_ctx_destroy", "lh_last_error", "lh_ldpc_decode", "lh_rate_dematch",
            "lh_pusch_decode_batch", "lh_crc_calculate", "lh_tb_finalize", "lh_gen_codeblocks",
            "lh_ldpc_encode_rate_match",
            "lh_pusch_demodulate",
            "lh_set_scale_table", "lh_pusch_decode_batch_learned", "lh_nms_workspace_bytes", "lh_nms_train_step",
            "lh_nms_adam", "lh_slot_graph_build", "lh_slot_graph_launch", "lh_slot_graph_destroy", "lh_tune_set",
            "lh_tune_set_rows", "lh_gen_transport_blocks", "lh_segment_rx", "lh_server_create", "lh_server_destroy", "lh_server_submit",
            "lh_server_step", "lh_server_poll", "lh_server_harq_release", "lh_server_stats", "lh_server_submit_batch",
            "lh_pusch_decode_batch_noalloc", "lh_server_run_slot", "lh_server_serve", "lh_ulsch_demux_get_sizes",
            "lh_ulsch_demultiplex", "lh_ulsch_information"]


def lib():
    """Loads the built HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LhError(f"{LIB_PATH} is missing: build it with ocudu-dapp-platform_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        L.lh_last_error.restype = C.c_char_p
        L.lh_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint]
        L.lh_ctx_destroy.argtypes = [C.c_void_p]
        L.lh_ldpc_decode.argtypes = [C.c_void_p, C.POINTER(DecodeDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_uint32, C.c_void_p]
        L.lh_rate_dematch.argtypes = [C.c_void_p, C.POINTER(RmDesc), C.c_uint, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_void_p]
        L.lh_pusch_decode_batch.argtypes = [C.c_void_p, C.POINTER(PuschCbDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint, C.c_void_p]
        L.lh_pusch_decode_batch_noalloc.argtypes = L.lh_pusch_decode_batch.argtypes
        L.lh_tb_finalize.argtypes = [C.c_void_p, C.POINTER(TbDesc), C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
        L.lh_set_scale_table.argtypes = [C.c_void_p, C.c_uint, C.POINTER(C.c_uint16), C.POINTER(C.c_int16),
                                         C.c_uint, C.c_uint]
        L.lh_pusch_decode_batch_learned.argtypes = [C.c_void_p, C.POINTER(PuschCbDesc), C.c_uint, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint,
                                                    C.c_uint, C.c_void_p]
        L.lh_nms_workspace_bytes.restype = C.c_size_t
        L.lh_nms_workspace_bytes.argtypes = [C.c_uint, C.POINTER(TrainDesc), C.c_uint, C.c_uint]
        L.lh_nms_train_step.argtypes = [C.c_void_p, C.c_uint, C.POINTER(TrainDesc), C.c_uint, C.c_uint, C.c_void_p,
                                        C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.c_void_p]
        L.lh_nms_adam.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint, C.c_void_p]
        L.lh_slot_graph_build.argtypes = [C.c_void_p, C.POINTER(PuschCbDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint, C.c_int, C.c_uint,
                                          C.POINTER(TbDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_void_p)]
        L.lh_slot_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
        L.lh_slot_graph_destroy.argtypes = [C.c_void_p]
        L.lh_tune_set.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_int]
        L.lh_tune_set_rows.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_int]
        L.lh_gen_transport_blocks.argtypes = [C.c_void_p, C.POINTER(GenTbDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L.lh_segment_rx.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.POINTER(Segmentation),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint]
        L.lh_server_create.argtypes = [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
        L.lh_server_destroy.argtypes = [C.c_void_p]
        L.lh_server_submit.argtypes = [C.c_void_p, C.POINTER(TbRequest), C.c_void_p]
        L.lh_server_step.argtypes = [C.c_void_p]
        L.lh_server_submit_batch.argtypes = [C.c_void_p, C.POINTER(TbRequest), C.c_uint, C.POINTER(C.c_void_p)]
        L.lh_server_poll.argtypes = [C.c_void_p, C.POINTER(TbResult), C.c_uint, C.POINTER(C.c_uint), C.c_int]
        L.lh_server_harq_release.argtypes = [C.c_void_p, C.c_uint32]
        L.lh_server_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
        L.lh_server_run_slot.argtypes = [C.c_void_p, C.c_uint32]
        L.lh_server_serve.argtypes = [C.c_void_p, C.POINTER(ServeTb), C.c_uint, C.POINTER(ServeConfig),
                                      C.POINTER(TbResult), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.lh_pusch_demodulate.argtypes = [C.c_void_p, C.POINTER(DemodDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
        L.lh_ulsch_demux_get_sizes.argtypes = [C.POINTER(UlschDemuxDesc), C.POINTER(UlschDemuxSizes)]
        L.lh_ulsch_information.argtypes = [C.POINTER(UlschConfig), C.POINTER(UlschInfo)]
        L.lh_ulsch_demultiplex.argtypes = [C.c_void_p, C.POINTER(UlschDemuxDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lh_ldpc_encode_rate_match.argtypes = [C.c_void_p, C.POINTER(EncodeDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]
        L.lh_gen_codeblocks.argtypes = [C.c_void_p, C.POINTER(GenDesc), C.c_uint, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.lh_crc_calculate.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                       C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _lh(rc, what):
    if rc != 0:
        raise LhError(f"{what} failed ({rc}): {lib().lh_last_error().decode()}")


class Context:
    def __init__(self, device=0, max_codeblocks=4096):
        h = C.c_void_p()
        _lh(lib().lh_ctx_create(C.byref(h), device, max_codeblocks), "lh_ctx_create")
        self.handle = h.value

    def close(self):
        if self.handle:
            lib().lh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ldpc_decode(self, descs, d_llrs, d_bits, d_results, flags=0, stream=None):
        arr = (DecodeDesc * len(descs))(*descs)
        _lh(lib().lh_ldpc_decode(self.handle, arr, len(descs), d_llrs.ptr, d_bits.ptr, d_results.ptr, flags,
                                 stream.handle if stream else None), "lh_ldpc_decode")

    def rate_dematch(self, descs, d_cw, d_rm, new_data=True, stream=None):
        arr = (RmDesc * len(descs))(*descs)
        _lh(lib().lh_rate_dematch(self.handle, arr, len(descs), d_cw.ptr, d_rm.ptr, int(new_data),
                                  stream.handle if stream else None), "lh_rate_dematch")

    def pusch_decode_batch(self, cbs, d_cw, d_rm, d_bits, d_results, new_data=True, use_early_stop=True,
                           nof_ldpc_iterations=6, stream=None, arr=None):
        if arr is None:
            arr = (PuschCbDesc * len(cbs))(*cbs)
        _lh(lib().lh_pusch_decode_batch(self.handle, arr, len(arr), d_cw.ptr, d_rm.ptr, d_bits.ptr, d_results.ptr,
                                        int(new_data), int(use_early_stop), nof_ldpc_iterations,
                                        stream.handle if stream else None), "lh_pusch_decode_batch")

    def set_scale_table(self, bg, alpha_q, beta_q):
        """alpha_q, beta_q: (iters, layers) arrays in 1/256 units."""
        a = np.ascontiguousarray(alpha_q, dtype=np.uint16)
        b = np.ascontiguousarray(beta_q, dtype=np.int16)
        assert a.shape == b.shape and a.ndim == 2
        _lh(lib().lh_set_scale_table(self.handle, bg, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                     b.ctypes.data_as(C.POINTER(C.c_int16)), a.shape[0], a.shape[1]),
            "lh_set_scale_table")

    def pusch_decode_batch_learned(self, cbs, d_cw, d_rm, d_bits, d_results, new_data=True, use_early_stop=True,
                                   nof_ldpc_iterations=6, nof_fallback_iterations=6, stream=None, arr=None):
        if arr is None:
            arr = (PuschCbDesc * len(cbs))(*cbs)
        _lh(lib().lh_pusch_decode_batch_learned(self.handle, arr, len(arr), d_cw.ptr, d_rm.ptr, d_bits.ptr,
                                                d_results.ptr, int(new_data), int(use_early_stop),
                                                nof_ldpc_iterations, nof_fallback_iterations,
                                                stream.handle if stream else None), "lh_pusch_decode_batch_learned")

    def tb_finalize(self, tbs, d_cb_bits, d_cb_results, d_tb, d_status, stream=None):
        arr = (TbDesc * len(tbs))(*tbs)
        _lh(lib().lh_tb_finalize(self.handle, arr, len(tbs), d_cb_bits.ptr, d_cb_results.ptr, d_tb.ptr, d_status.ptr,
                                 stream.handle if stream else None), "lh_tb_finalize")

    def encode_rate_match(self, descs, d_msgs, d_out, d_codewords=None, stream=None, arr=None):
        """LDPC encoding + rate matching (lh_ldpc_encode_rate_match); `arr`: the descriptors
        already as a ctypes array (no per-call conversion)."""
        if arr is None:
            arr = (EncodeDesc * len(descs))(*descs)
        _lh(lib().lh_ldpc_encode_rate_match(self.handle, arr, len(arr), d_msgs.ptr, d_out.ptr,
                                            d_codewords.ptr if d_codewords else None,
                                            stream.handle if stream else None), "lh_ldpc_encode_rate_match")

    def pusch_demodulate(self, descs, d_symbols, d_noise_vars, d_llrs, stream=None):
        """Soft demodulation + descrambling of blocks of equalized symbols (lh_pusch_demodulate)."""
        arr = (DemodDesc * len(descs))(*descs)
        _lh(lib().lh_pusch_demodulate(self.handle, arr, len(descs), d_symbols.ptr, d_noise_vars.ptr, d_llrs.ptr,
                                      stream.handle if stream else None), "lh_pusch_demodulate")

    def ulsch_demultiplex(self, descs, d_llrs, d_sch=None, d_harq_ack=None, d_csi_part1=None, d_csi_part2=None,
                          stream=None):
        """UL-SCH demultiplexing of whole descrambled codewords (lh_ulsch_demultiplex)."""
        arr = (UlschDemuxDesc * len(descs))(*descs)

        def ptr(b):
            return b.ptr if b is not None else None

        _lh(lib().lh_ulsch_demultiplex(self.handle, arr, len(descs), d_llrs.ptr, ptr(d_sch), ptr(d_harq_ack),
                                       ptr(d_csi_part1), ptr(d_csi_part2), stream.handle if stream else None),
            "lh_ulsch_demultiplex")

    def gen_codeblocks(self, descs, d_llrs, d_msgs=None, stream=None):
        arr = (GenDesc * len(descs))(*descs)
        _lh(lib().lh_gen_codeblocks(self.handle, arr, len(descs), d_llrs.ptr, d_msgs.ptr if d_msgs else None,
                                    stream.handle if stream else None), "lh_gen_codeblocks")

    @staticmethod
    def nms_workspace_bytes(bg, descs, iters):
        arr = (TrainDesc * len(descs))(*descs)
        return int(lib().lh_nms_workspace_bytes(bg, arr, len(descs), iters))

    def nms_train_step(self, bg, descs, iters, d_alpha, d_beta, w, tau, d_llrs, d_msgs, d_ws, d_loss,
                       d_galpha=None, d_gbeta=None, backward=True, stream=None, arr=None):
        if arr is None:
            arr = (TrainDesc * len(descs))(*descs)
        wv = (C.c_float * iters)(*[float(x) for x in w])
        _lh(lib().lh_nms_train_step(self.handle, bg, arr, len(arr), iters, d_alpha.ptr, d_beta.ptr, wv, tau,
                                    d_llrs.ptr, d_msgs.ptr, d_ws.ptr, d_ws.nbytes, d_loss.ptr,
                                    d_galpha.ptr if d_galpha else None, d_gbeta.ptr if d_gbeta else None,
                                    int(backward), stream.handle if stream else None), "lh_nms_train_step")

    def nms_adam(self, d_params, d_grads, d_m, d_v, n, n_alpha, lr, step, beta1=0.9, beta2=0.999, eps=1e-8,
                 stream=None, offset=0):
        _lh(lib().lh_nms_adam(self.handle, d_params.ptr + offset, d_grads.ptr + offset, d_m.ptr + offset,
                              d_v.ptr + offset, n, n_alpha, lr, beta1, beta2, eps, step,
                              stream.handle if stream else None), "lh_nms_adam")

    def slot_graph(self, cbs, d_cw, d_rm, d_bits, d_results, new_data=True, use_early_stop=True,
                   nof_ldpc_iterations=8, learned=False, nof_fallback_iterations=0, tbs=(), d_tb=None,
                   d_tb_status=None):
        return SlotGraph(self, cbs, d_cw, d_rm, d_bits, d_results, new_data, use_early_stop, nof_ldpc_iterations,
                         learned, nof_fallback_iterations, tbs, d_tb, d_tb_status)

    def tune_set(self, bg, Z, groups, stage_llrs=False):
        _lh(lib().lh_tune_set(self.handle, bg, Z, groups, int(stage_llrs)), "lh_tune_set")

    def tune_rows(self, bg, Z, row_pairs):
        """Row mode of (bg, Z): -1 chosen per launch from occupancy and load, 0 one check row
        per thread, 1 two rows per thread (even Z)."""
        _lh(lib().lh_tune_set_rows(self.handle, bg, Z, int(row_pairs)), "lh_tune_set_rows")

    def gen_transport_blocks(self, descs, d_tb, d_msgs, stream=None):
        arr = (GenTbDesc * len(descs))(*descs)
        _lh(lib().lh_gen_transport_blocks(self.handle, arr, len(descs), d_tb.ptr, d_msgs.ptr,
                                          stream.handle if stream else None), "lh_gen_transport_blocks")

    def crc_calculate(self, poly, offsets, nbits, d_bits, d_out, stream=None):
        n = len(offsets)
        offs = (C.c_uint64 * n)(*offsets)
        nb = (C.c_uint32 * n)(*nbits)
        _lh(lib().lh_crc_calculate(self.handle, poly, offs, nb, n, d_bits.ptr, d_out.ptr,
                                   stream.handle if stream else None), "lh_crc_calculate")


class SlotGraph:
    """A slot's batch captured as a HIP graph (lh_slot_graph_*)."""

    def __init__(self, ctx, cbs, d_cw, d_rm, d_bits, d_results, new_data, use_early_stop, iters, learned, fb_iters,
                 tbs, d_tb, d_tb_status):
        arr = (PuschCbDesc * len(cbs))(*cbs)
        tarr = (TbDesc * max(1, len(tbs)))(*tbs) if len(tbs) else None
        h = C.c_void_p()
        _lh(lib().lh_slot_graph_build(ctx.handle, arr, len(cbs), d_cw.ptr, d_rm.ptr, d_bits.ptr, d_results.ptr,
                                      int(new_data), int(use_early_stop), iters, int(learned), fb_iters, tarr,
                                      len(tbs), d_tb.ptr if d_tb else None, d_tb_status.ptr if d_tb_status else None,
                                      C.byref(h)), "lh_slot_graph_build")
        self.handle = h.value

    def launch(self, stream=None):
        _lh(lib().lh_slot_graph_launch(self.handle, stream.handle if stream else None), "lh_slot_graph_launch")

    def close(self):
        if self.handle:
            lib().lh_slot_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The slot-deadline serving runtime (lh_server_*)."""

    def __init__(self, ctx, harq_budget_bytes, max_cbs_per_slot=1024, max_slots_in_flight=2, nof_ldpc_iterations=8,
                 learned=False, nof_fallback_iterations=8, expire_timeout_slots=0):
        cfg = ServerConfig(harq_budget_bytes, max_cbs_per_slot, max_slots_in_flight, nof_ldpc_iterations,
                           nof_fallback_iterations if learned else 0, int(learned), expire_timeout_slots)
        h = C.c_void_p()
        _lh(lib().lh_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "lh_server_create")
        self.handle = h.value
        self._buf = (TbResult * 4096)()

    def submit(self, req, tb_out=None):
        _lh(lib().lh_server_submit(self.handle, C.byref(req), tb_out), "lh_server_submit")

    def submit_batch(self, reqs, outs=None):
        """reqs: a ctypes array of TbRequest; outs: a ctypes array of host pointers or None."""
        _lh(lib().lh_server_submit_batch(self.handle, reqs, len(reqs), outs), "lh_server_submit_batch")

    def step(self):
        _lh(lib().lh_server_step(self.handle), "lh_server_step")

    def poll(self, wait=False):
        n = C.c_uint()
        _lh(lib().lh_server_poll(self.handle, self._buf, len(self._buf), C.byref(n), int(wait)), "lh_server_poll")
        return [TbResult.from_buffer_copy(self._buf[i]) for i in range(n.value)]

    def harq_release(self, process):
        _lh(lib().lh_server_harq_release(self.handle, process), "lh_server_harq_release")

    def run_slot(self, slot):
        _lh(lib().lh_server_run_slot(self.handle, slot), "lh_server_run_slot")

    def serve(self, tbs, nof_slots, harq_rtt_slots, slot_us=500.0, realtime=True, timeout_s=600.0):
        """The native slot loop over a ctypes array of ServeTb: (results, slot latencies in
        us, wall seconds); results[2 i + tx] per transmission."""
        cfg = ServeConfig(slot_us, timeout_s, nof_slots, harq_rtt_slots, int(realtime), 0)
        res = (TbResult * (2 * len(tbs)))()
        lat = (C.c_double * (nof_slots + harq_rtt_slots + 2))()
        wall = C.c_double()
        _lh(lib().lh_server_serve(self.handle, tbs, len(tbs), C.byref(cfg), res, lat, C.byref(wall)),
            "lh_server_serve")
        return list(res), list(lat), wall.value

    def stats(self):
        st = ServerStats()
        _lh(lib().lh_server_stats(self.handle, C.byref(st)), "lh_server_stats")
        return st

    def close(self):
        if self.handle:
            lib().lh_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def segment_rx(bg, tbs_bits, qm, nof_layers, nof_ch_symbols, max_cbs=512):
    """(Segmentation, rm lengths, cw offsets) of a received transport block (lh_segment_rx)."""
    seg = Segmentation()
    rm = (C.c_uint32 * max_cbs)()
    off = (C.c_uint32 * max_cbs)()
    _lh(lib().lh_segment_rx(bg, tbs_bits, qm, nof_layers, nof_ch_symbols, C.byref(seg), rm, off, max_cbs),
        "lh_segment_rx")
    return seg, list(rm[:seg.nof_cbs]), list(off[:seg.nof_cbs])


def k_b(bg):
    return 22 if bg == 1 else 10


def n_short(bg):
    return 66 if bg == 1 else 50


def msg_bytes(bg, Z):
    return (k_b(bg) * Z + 7) // 8
