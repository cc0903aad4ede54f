"""ctypes access to the CPU oracle (liboracle.so) and to the reference build
(_ref/libref_ldpc.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py as the
checker.  The HIP product never imports this module.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
CRC24A, CRC24B, CRC16 = 0x1864CFB, 0x1800063, 0x11021
LIFTING_SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40,
                 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240,
                 256, 288, 320, 352, 384]

_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)


def _p(a, t):
    return a.ctypes.data_as(t)


class _DecCfg(C.Structure):
    _fields_ = [("bg", C.c_int), ("Z", C.c_uint), ("nof_filler_bits", C.c_uint), ("nof_crc_bits", C.c_uint),
                ("max_iterations", C.c_uint)]


class _RmCfg(C.Structure):
    _fields_ = [("rv", C.c_uint), ("Qm", C.c_uint), ("Nref", C.c_uint), ("nof_filler_bits", C.c_uint)]


class _Seg(C.Structure):
    _fields_ = [("nof_segments", C.c_uint), ("Z", C.c_uint), ("segment_length", C.c_uint),
                ("nof_filler_bits", C.c_uint), ("nof_crc_bits", C.c_uint), ("nof_tb_crc_bits", C.c_uint),
                ("full_length", C.c_uint), ("rm_length", C.c_uint * 256), ("cw_offset", C.c_uint * 256)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built (make -f oracle/Makefile)")
        L = C.CDLL(str(path))
        L.orc_llr_add.restype = C.c_int8
        _fp = C.POINTER(C.c_float)
        L.orc_demodulate_soft.argtypes = [_i8p, _fp, _fp, C.c_uint, C.c_int]
        L.orc_scrambling_sequence.argtypes = [_u8p, C.c_uint32, C.c_uint, C.c_uint]
        L.orc_pusch_demodulate.argtypes = [_i8p, _fp, _fp, C.c_uint, C.c_int, C.c_uint32, C.c_uint]
        L.orc_llr_add.argtypes = [C.c_int8, C.c_int8]
        L.orc_llr_promotion_sum.restype = C.c_int8
        L.orc_llr_promotion_sum.argtypes = [C.c_int8, C.c_int8]
        L.orc_llr_quantize.restype = C.c_int8
        L.orc_llr_quantize.argtypes = [C.c_float, C.c_float]
        L.orc_crc_packed.restype = C.c_uint32
        L.orc_crc_packed.argtypes = [C.c_uint32, _u8p, C.c_uint]
        L.orc_crc_unpacked.restype = C.c_uint32
        L.orc_crc_unpacked.argtypes = [C.c_uint32, _u8p, C.c_uint]
        L.orc_graph_shift.restype = C.c_int
        L.orc_graph_shift.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint]
        L.orc_lifting_index.argtypes = [C.c_uint]
        L.orc_ldpc_decode.restype = C.c_uint
        L.orc_ldpc_decode.argtypes = [C.POINTER(_DecCfg), C.c_int, C.c_int, C.c_uint32, _i8p, C.c_uint, _u8p]
        L.orc_rate_dematch.argtypes = [_i8p, C.c_uint, _i8p, C.c_uint, C.c_int, C.POINTER(_RmCfg)]
        L.orc_segment_rx.argtypes = [C.POINTER(_Seg), C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint]
        L.orc_ldpc_encode.argtypes = [C.c_int, C.c_uint, _u8p, _u8p]
        L.orc_syndrome_weight.restype = C.c_uint
        L.orc_syndrome_weight.argtypes = [C.c_int, C.c_uint, _u8p]
        L.orc_rate_match.argtypes = [_u8p, C.c_uint, _u8p, C.c_uint, C.c_int, C.c_uint, C.c_uint, C.c_uint,
                                     C.c_uint, C.c_uint]
        L.orc_modulate_awgn_demap.argtypes = [_i8p, _u8p, C.c_uint, C.c_uint, C.c_double, C.c_uint64]
        L.orc_ldpc_decode_learned.restype = C.c_uint
        L.orc_ldpc_decode_learned.argtypes = [C.POINTER(_DecCfg), C.POINTER(C.c_uint32), C.c_uint32, _i8p, C.c_uint,
                                              _u8p]
        for f in (L.orc_nms_forward_f, L.orc_nms_forward_d):
            f.restype = C.c_double
            f.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, _i8p, C.c_uint, _u8p,
                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double]
        L.orc_gen_codeblock.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint32, C.c_uint, C.c_uint, C.c_uint,
                                        C.c_uint, C.c_double, C.c_uint64, C.c_uint64, _u8p, _i8p]
        L.orc_tb_concat_crc.restype = C.c_uint32
        L.orc_tb_concat_crc.argtypes = [_u8p, C.c_uint, _u8p, C.c_uint, C.POINTER(_Seg)]
        _lib = L
    return _lib


def ref():
    """The reference's own CPU code built from source, or None when not built."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libref_ldpc.so"
        if not path.exists():
            return None
        R = C.CDLL(str(path))
        R.ref_decoder_create.restype = C.c_void_p
        R.ref_decoder_create.argtypes = [C.c_int, C.c_int, C.c_int]
        R.ref_decoder_destroy.argtypes = [C.c_void_p]
        R.ref_decoder_kind.argtypes = [C.c_int]
        R.ref_decoder_decode.restype = C.c_uint
        R.ref_decoder_decode.argtypes = [C.c_void_p, _i8p, C.c_uint, C.c_int, C.c_uint, C.c_uint, C.c_uint,
                                         C.c_uint, C.c_uint32, _u8p]
        R.ref_rate_dematch.argtypes = [_i8p, C.c_uint, _i8p, C.c_uint, C.c_int, C.c_int, C.c_uint, C.c_uint,
                                       C.c_uint, C.c_uint, C.c_uint]
        R.ref_encode_rate_match.argtypes = [_u8p, C.c_int, C.c_uint, _u8p, _u8p, C.c_uint, C.c_uint, C.c_uint,
                                            C.c_uint, C.c_uint]
        R.ref_crc.restype = C.c_uint32
        R.ref_crc.argtypes = [C.c_uint32, _u8p, C.c_uint]
        R.ref_graph_shift.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint]
        R.ref_llr_add.restype = C.c_int8
        R.ref_llr_add.argtypes = [C.c_int8, C.c_int8]
        R.ref_llr_promotion_sum.restype = C.c_int8
        R.ref_llr_promotion_sum.argtypes = [C.c_int8, C.c_int8]
        R.ref_llr_quantize.restype = C.c_int8
        R.ref_llr_quantize.argtypes = [C.c_float, C.c_float]
        R.ref_segment_tx.argtypes = [C.c_int, _u8p, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                     C.POINTER(C.c_uint), _u8p, C.c_uint]
        R.ref_segment_rx.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                     C.POINTER(C.c_uint)]
        R.ref_decode_throughput.restype = C.c_double
        R.ref_decode_throughput.argtypes = [C.c_int, C.c_int, _i8p, C.c_uint, C.c_uint, C.c_int, C.c_uint,
                                            C.c_uint, C.c_uint, C.c_uint32, C.c_double,
                                            C.POINTER(C.c_ulong), C.POINTER(C.c_ulong), C.POINTER(C.c_ulong)]
        _ref = R
    return _ref


def K_b(bg):
    return 22 if bg == 1 else 10


def N_short(bg):
    return 66 if bg == 1 else 50


def M_rows(bg):
    return 46 if bg == 1 else 42


# ---------------------------------------------------------------------------------
def decode(llrs, bg, Z, F=0, crc_poly=CRC24B, max_iter=6, crc_bits=24, force=False, es_syn=False, out=None):
    """Oracle decoder; returns (iterations or 0, packed bits)."""
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    nbytes = (K_b(bg) * Z + 7) // 8
    if out is None:
        out = np.zeros(nbytes, np.uint8)
    cfg = _DecCfg(bg, Z, F, crc_bits, max_iter)
    r = lib().orc_ldpc_decode(C.byref(cfg), int(force), int(es_syn), crc_poly, _p(llrs, _i8p), llrs.size,
                              _p(out, _u8p))
    return int(r), out


def ref_decode(llrs, bg, Z, F=0, crc_poly=CRC24B, max_iter=6, crc_bits=24, kind=0, force=False, es_syn=False,
               out=None, dec=None):
    R = ref()
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    nbytes = (K_b(bg) * Z + 7) // 8
    if out is None:
        out = np.zeros(nbytes, np.uint8)
    own = dec is None
    if own:
        dec = R.ref_decoder_create(kind, int(force), int(es_syn))
    try:
        r = R.ref_decoder_decode(dec, _p(llrs, _i8p), llrs.size, bg, Z, F, crc_bits, max_iter, crc_poly,
                                 _p(out, _u8p))
    finally:
        if own:
            R.ref_decoder_destroy(dec)
    return int(r), out


def rate_dematch(llrs_e, bg, Z, rv=0, Qm=2, Nref=0, F=0, new_data=True, harq=None):
    N = N_short(bg) * Z
    out = np.zeros(N, np.int8) if harq is None else np.array(harq, dtype=np.int8, copy=True)
    llrs_e = np.ascontiguousarray(llrs_e, dtype=np.int8)
    cfg = _RmCfg(rv, Qm, Nref, F)
    rc = lib().orc_rate_dematch(_p(out, _i8p), N, _p(llrs_e, _i8p), llrs_e.size, int(new_data), C.byref(cfg))
    if rc != 0:
        raise ValueError("invalid dematcher configuration")
    return out


def ref_rate_dematch(llrs_e, bg, Z, rv=0, Qm=2, Nref=0, F=0, new_data=True, harq=None):
    N = N_short(bg) * Z
    out = np.zeros(N, np.int8) if harq is None else np.array(harq, dtype=np.int8, copy=True)
    llrs_e = np.ascontiguousarray(llrs_e, dtype=np.int8)
    ref().ref_rate_dematch(_p(out, _i8p), N, _p(llrs_e, _i8p), llrs_e.size, int(new_data), bg, Z, rv, Qm, Nref, F)
    return out


def crc(poly, packed, nbits):
    packed = np.ascontiguousarray(packed, dtype=np.uint8)
    return int(lib().orc_crc_packed(poly, _p(packed, _u8p), nbits))


def encode(msg_bits, bg, Z):
    msg_bits = np.ascontiguousarray(msg_bits, dtype=np.uint8)
    cw = np.zeros((K_b(bg) + M_rows(bg)) * Z, np.uint8)
    rc = lib().orc_ldpc_encode(bg, Z, _p(msg_bits, _u8p), _p(cw, _u8p))
    if rc != 0:
        raise RuntimeError(f"encode failed ({rc})")
    return cw


def syndrome_weight(cw, bg, Z):
    cw = np.ascontiguousarray(cw, dtype=np.uint8)
    return int(lib().orc_syndrome_weight(bg, Z, _p(cw, _u8p)))


def rate_match(d, E, bg, Z, rv=0, Qm=2, Nref=0, F=0):
    d = np.ascontiguousarray(d, dtype=np.uint8)
    out = np.zeros(E, np.uint8)
    rc = lib().orc_rate_match(_p(out, _u8p), E, _p(d, _u8p), d.size, bg, Z, rv, Qm, Nref, F)
    if rc != 0:
        raise ValueError("invalid rate matcher configuration")
    return out


def awgn_llrs(bits, Qm, esn0_db, seed):
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    out = np.zeros(bits.size, np.int8)
    n0 = 10.0 ** (-esn0_db / 10.0)
    lib().orc_modulate_awgn_demap(_p(out, _i8p), _p(bits, _u8p), bits.size, Qm, n0, seed)
    return out


def pack_bits(bits):
    return np.packbits(np.asarray(bits, dtype=np.uint8))


def make_codeblocks(n_cb, bg, Z, E, Qm, esn0_db, seed, rv=0, F=0, crc_poly=CRC24B, crc_bits=24, Nref=0):
    """Seeded synthetic codeblocks: random info bits + CB CRC (+ F zero fillers), LDPC
    encoded, rate matched to E, mapped, AWGN at Es/N0, max-log demapped to int8.
    Returns (info_bits_packed[n_cb, K_b*Z/8], llrs_e[n_cb, E])."""
    rng = np.random.default_rng(seed)
    K = K_b(bg) * Z
    k_data = K - F - crc_bits
    msgs = np.zeros((n_cb, (K + 7) // 8), np.uint8)
    llrs = np.zeros((n_cb, E), np.int8)
    for c in range(n_cb):
        m = np.zeros(K, np.uint8)
        m[:k_data] = rng.integers(0, 2, k_data, dtype=np.uint8)
        pk = pack_bits(m[:k_data])
        r = crc(crc_poly, pk, k_data)
        for i in range(crc_bits):
            m[k_data + i] = (r >> (crc_bits - 1 - i)) & 1
        cw = encode(m, bg, Z)
        d = cw[2 * Z: (K_b(bg) + M_rows(bg)) * Z]
        e = rate_match(d, E, bg, Z, rv=rv, Qm=Qm, Nref=Nref, F=F)
        llrs[c] = awgn_llrs(e, Qm, esn0_db, int(rng.integers(0, 2 ** 63)))
        msgs[c] = pack_bits(m)
    return msgs, llrs


SEG_META = ("Z", "segment_length", "full_length", "rm_length", "nof_filler_bits", "cw_offset", "nof_crc_bits",
            "info_bits")


def segment_rx(bg, tbs_bytes, Qm, nof_layers, nof_ch_symbols):
    """Oracle RX segmentation: list of per-codeblock dicts with the SEG_META keys (info_bits 0)."""
    seg = _Seg()
    rc = lib().orc_segment_rx(C.byref(seg), bg, 8 * tbs_bytes, nof_ch_symbols, Qm, nof_layers)
    if rc != 0:
        raise ValueError("invalid segmentation")
    out = []
    for c in range(seg.nof_segments):
        out.append(dict(Z=seg.Z, segment_length=seg.segment_length, full_length=seg.full_length,
                        rm_length=seg.rm_length[c], nof_filler_bits=seg.nof_filler_bits, cw_offset=seg.cw_offset[c],
                        nof_crc_bits=seg.nof_crc_bits, info_bits=0))
    return out, seg


def tb_concat_crc(cb_bits_packed, stride, tbs_bytes, seg):
    """Oracle TB reassembly: (TB bytes, received checksum XOR computed CRC24A; 0 = TB CRC ok)."""
    cb = np.ascontiguousarray(cb_bits_packed, dtype=np.uint8)
    tb = np.zeros(tbs_bytes, np.uint8)
    r = lib().orc_tb_concat_crc(_p(tb, _u8p), 8 * tbs_bytes, _p(cb, _u8p), stride, C.byref(seg))
    return tb, int(r)


def _meta_dicts(meta, n):
    m = np.ctypeslib.as_array(meta)[: 8 * n].reshape(n, 8)
    return [dict(zip(SEG_META, (int(v) for v in row))) for row in m]


def ref_segment_tx(bg, tb_bytes, Qm, nof_layers, nof_ch_symbols, rv=0, with_bits=True):
    """Reference TX segmentation of a transport block: (metadata dicts, codeblock bits (C, seg_len) uint8)."""
    R = ref()
    tb = np.ascontiguousarray(tb_bytes, dtype=np.uint8)
    meta = (C.c_uint * (8 * 256))()
    kmax = 22 * 384
    bits = np.zeros((256, kmax), np.uint8) if with_bits else None
    n = R.ref_segment_tx(bg, _p(tb, _u8p), tb.size, Qm, nof_layers, nof_ch_symbols, rv, 256, meta,
                         _p(bits, _u8p) if with_bits else None, kmax)
    if n < 0:
        raise ValueError("too many codeblocks")
    md = _meta_dicts(meta, n)
    if with_bits:
        bits = bits[:n, : md[0]["segment_length"]]
    return md, bits


def ref_segment_rx(bg, tbs_bytes, Qm, nof_layers, nof_ch_symbols, rv=0):
    """Reference RX segmentation metadata (info_bits 0)."""
    meta = (C.c_uint * (8 * 256))()
    n = ref().ref_segment_rx(bg, tbs_bytes, Qm, nof_layers, nof_ch_symbols, rv, 256, meta)
    if n < 0:
        raise ValueError("too many codeblocks")
    return _meta_dicts(meta, n)


def gen_codeblock(bg, Z, E, Qm, n0, seed_msg, seed_noise, rv=0, F=0, crc_poly=CRC24B, Nref=0):
    """The device generator's statement: (packed message, E int8 LLRs)."""
    msg = np.zeros((K_b(bg) * Z + 7) // 8, np.uint8)
    llr = np.zeros(E, np.int8)
    rc = lib().orc_gen_codeblock(bg, Z, F, crc_poly, E, rv, Qm, Nref, n0, seed_msg, seed_noise, _p(msg, _u8p),
                                 _p(llr, _i8p))
    if rc != 0:
        raise ValueError(f"generator failed ({rc})")
    return msg, llr


def scale_table_words(alpha_q, beta_q):
    """(iters, M) alpha/beta in 1/256 units -> the 8 x M words A | B << 16 (last row repeated)."""
    a = np.asarray(alpha_q, np.int64)
    b = np.asarray(beta_q, np.int64)
    rows = [np.minimum(t, a.shape[0] - 1) for t in range(8)]
    w = (a[rows] & 0xffff) | ((b[rows] & 0xffff) << 16)
    return np.ascontiguousarray(w.astype(np.uint32))


def decode_learned(llrs, bg, Z, alpha_q, beta_q, F=0, crc_poly=CRC24B, max_iter=6, crc_bits=24):
    """Learned min-sum decode (CRC early stop): (iterations or 0, packed bits)."""
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    tab = scale_table_words(alpha_q, beta_q)
    out = np.zeros((K_b(bg) * Z + 7) // 8, np.uint8)
    cfg = _DecCfg(bg, Z, F, crc_bits, max_iter)
    r = lib().orc_ldpc_decode_learned(C.byref(cfg), tab.ctypes.data_as(C.POINTER(C.c_uint32)), crc_poly,
                                      _p(llrs, _i8p), llrs.size, _p(out, _u8p))
    return int(r), out


def nms_forward(llrs, msg_packed, bg, Z, L, alpha, beta, w, tau, F=0, double=True):
    """Loss of the learned decoder's floating-point training model; alpha, beta: (T, M)."""
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    msg = np.ascontiguousarray(msg_packed, dtype=np.uint8)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    wv = np.ascontiguousarray(w, dtype=np.float64)
    T = a.shape[0]
    dp = C.POINTER(C.c_double)
    f = lib().orc_nms_forward_d if double else lib().orc_nms_forward_f
    return float(f(bg, Z, F, L, T, _p(llrs, _i8p), llrs.size, _p(msg, _u8p), a.ctypes.data_as(dp),
                   b.ctypes.data_as(dp), wv.ctypes.data_as(dp), tau))


# --- f1: PUSCH soft demodulation and descrambling --------------------------------------

_ref_demod = None


def ref_demod():
    """The reference's scalar demappers, LUT modulation mapper and pseudo-random generator
    built from source (_ref/libref_demod.so), or None when not built."""
    global _ref_demod
    if _ref_demod is None:
        path = HERE / "_ref" / "libref_demod.so"
        if not path.exists():
            return None
        R = C.CDLL(str(path))
        _fp = C.POINTER(C.c_float)
        R.ref_demodulate_soft.argtypes = [_i8p, _fp, _fp, C.c_uint, C.c_int]
        R.ref_modulate.argtypes = [_fp, _u8p, C.c_uint, C.c_int]
        R.ref_scrambling_sequence.argtypes = [_u8p, C.c_uint32, C.c_uint, C.c_uint]
        R.ref_ulsch_information.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_float), C.POINTER(C.c_uint)]
        R.ref_ulsch_demultiplex.argtypes = [C.POINTER(C.c_uint), _i8p, C.c_uint, C.c_uint32, C.c_uint, _i8p, _i8p,
                                            _i8p, _i8p, C.POINTER(C.c_uint)]
        _ref_demod = R
    return _ref_demod


def _demod_args(L, iq, nv):
    iq = np.ascontiguousarray(iq, dtype=np.float32).reshape(-1)
    nv = np.ascontiguousarray(nv, dtype=np.float32).reshape(-1)
    assert iq.size == 2 * nv.size
    return iq, nv


def bits_per_symbol(modulation):
    return 1 if modulation == 0 else modulation


def demodulate_soft(iq, nv, modulation, use_ref=False):
    """int8 LLRs of symbols iq (re, im interleaved float32) with noise variances nv."""
    L = ref_demod() if use_ref else lib()
    iq, nv = _demod_args(L, iq, nv)
    out = np.zeros(nv.size * bits_per_symbol(modulation), np.int8)
    fn = L.ref_demodulate_soft if use_ref else L.orc_demodulate_soft
    fn(_p(out, _i8p), _p(iq, C.POINTER(C.c_float)), _p(nv, C.POINTER(C.c_float)), nv.size, modulation)
    return out


def scrambling_sequence(c_init, n, offset=0, use_ref=False):
    """c(offset .. offset + n - 1) of TS38.211 5.2.1, one bit per byte."""
    out = np.zeros(max(n, 1), np.uint8)
    if use_ref:
        ref_demod().ref_scrambling_sequence(_p(out, _u8p), c_init, offset, n)
    else:
        lib().orc_scrambling_sequence(_p(out, _u8p), c_init, offset, n)
    return out[:n]


def pusch_demodulate(iq, nv, modulation, c_init, seq_offset=0):
    """Demodulated and descrambled LLRs of one block (the PUSCH demodulator's order)."""
    L = lib()
    iq, nv = _demod_args(L, iq, nv)
    out = np.zeros(nv.size * bits_per_symbol(modulation), np.int8)
    L.orc_pusch_demodulate(_p(out, _i8p), _p(iq, C.POINTER(C.c_float)), _p(nv, C.POINTER(C.c_float)), nv.size,
                           modulation, c_init, seq_offset)
    return out


# --- f1: UL-SCH demultiplexing ---------------------------------------------------------

DEMUX_FIELDS = ("modulation", "nof_layers", "nof_prb", "start_symbol_index", "nof_symbols", "dmrs_symbol_mask",
                "dmrs_type", "nof_cdm_groups_without_data", "nof_harq_ack_rvd", "nof_harq_ack_bits",
                "nof_enc_harq_ack_bits", "nof_csi_part1_bits", "nof_enc_csi_part1_bits", "nof_csi_part2_bits",
                "nof_enc_csi_part2_bits")


class _DemuxCfg(C.Structure):
    _fields_ = [("modulation", C.c_int)] + [(f, C.c_uint) for f in DEMUX_FIELDS[1:]]


class _DemuxSizes(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("nof_llrs", "nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")]


def _demux_cfg(cfg):
    return _DemuxCfg(**{f: int(cfg.get(f, 0)) for f in DEMUX_FIELDS})


def demux_sizes(cfg):
    """Soft-bit counts {nof_llrs, nof_sch, nof_harq_ack, nof_csi_part1, nof_csi_part2} of a
    demultiplexer configuration, or None when the configuration is invalid."""
    L = lib()
    sz = _DemuxSizes()
    if L.orc_ulsch_demultiplex(C.byref(_demux_cfg(cfg)), None, 0, None, None, None, None, C.byref(sz)) != 0:
        return None
    return {f: getattr(sz, f) for f, _ in _DemuxSizes._fields_}


def ulsch_demultiplex(cfg, llrs, c_init):
    """(sch, harq_ack, csi_part1, csi_part2) soft bits of one descrambled codeword."""
    L = lib()
    sz = demux_sizes(cfg)
    if sz is None:
        raise ValueError("invalid demultiplexer configuration")
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    assert llrs.size == sz["nof_llrs"]
    outs = [np.zeros(max(sz[k], 1), np.int8) for k in ("nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")]
    rc = L.orc_ulsch_demultiplex(C.byref(_demux_cfg(cfg)), _p(llrs, _i8p), c_init, *[_p(o, _i8p) for o in outs],
                                 None)
    assert rc == 0
    return tuple(o[:sz[k]] for o, k in zip(outs, ("nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")))


def ulsch_positions(cfg):
    """(sch, harq_ack, csi_part1, csi_part2): the codeword position of each field's soft
    bits (int32 arrays)."""
    sz = demux_sizes(cfg)
    if sz is None:
        raise ValueError("invalid demultiplexer configuration")
    keys = ("nof_sch", "nof_harq_ack", "nof_csi_part1", "nof_csi_part2")
    outs = [np.zeros(max(sz[k], 1), np.int32) for k in keys]
    i32p = C.POINTER(C.c_int32)
    assert lib().orc_ulsch_positions(C.byref(_demux_cfg(cfg)), *[_p(o, i32p) for o in outs]) == 0
    return tuple(o[:sz[k]] for o, k in zip(outs, keys))


def ref_ulsch_demultiplex(cfg, llrs, c_init, block=1000):
    """The reference's ulsch_demultiplex_impl built from source, fed in blocks of `block`
    soft bits, CSI Part 2 configured before the first block; None for a configuration
    whose first symbol has no data REs (the reference's streaming never leaves it)."""
    R = ref_demod()
    words = (C.c_uint * 15)(*[int(cfg.get(f, 0)) for f in DEMUX_FIELDS])
    llrs = np.ascontiguousarray(llrs, dtype=np.int8)
    n = llrs.size
    outs = [np.zeros(max(n, 1), np.int8) for _ in range(4)]
    sizes = (C.c_uint * 4)()
    rc = R.ref_ulsch_demultiplex(words, _p(llrs, _i8p), n, c_init, block, *[_p(o, _i8p) for o in outs], sizes)
    if rc == -2:
        return None  # a first symbol without data REs: the reference's streaming never advances
    if rc != 0:
        raise RuntimeError("a decoder buffer did not receive its end")
    return tuple(o[:sizes[k]] for k, o in enumerate(outs))


# --- UL-SCH bit accounting (get_ulsch_information) --------------------------------------

ULSCH_INT_FIELDS = ("tbs_bits", "modulation", "nof_harq_ack_bits", "nof_csi_part1_bits", "nof_csi_part2_bits",
                    "nof_rb", "start_symbol_index", "nof_symbols", "dmrs_type", "dmrs_symbol_mask",
                    "nof_cdm_groups_without_data", "nof_layers", "contains_dc")
ULSCH_FLOAT_FIELDS = ("target_code_rate", "alpha_scaling", "beta_offset_harq_ack", "beta_offset_csi_part1",
                      "beta_offset_csi_part2")
ULSCH_INFO_FIELDS = ("nof_ul_sch_bits", "nof_harq_ack_bits", "nof_harq_ack_rvd", "nof_csi_part1_bits",
                     "nof_csi_part2_bits", "nof_harq_ack_re", "nof_csi_part1_re", "nof_csi_part2_re",
                     "nof_dc_overlap_bits")
SCH_INFO_FIELDS = ("base_graph", "nof_cb", "nof_bits_per_cb", "lifting_size", "nof_filler_bits_per_cb",
                   "tb_crc_size")


class _SchInfo(C.Structure):
    _fields_ = [(f, C.c_uint) for f in ("base_graph", "tb_crc_size", "nof_cb", "lifting_size", "nof_bits_per_cb",
                                         "nof_filler_bits_per_cb")]


class _UlschCfg(C.Structure):
    _fields_ = [("tbs_bits", C.c_uint), ("target_code_rate", C.c_float), ("modulation", C.c_int),
                ("nof_harq_ack_bits", C.c_uint), ("nof_csi_part1_bits", C.c_uint), ("nof_csi_part2_bits", C.c_uint),
                ("alpha_scaling", C.c_float), ("beta_offset_harq_ack", C.c_float),
                ("beta_offset_csi_part1", C.c_float), ("beta_offset_csi_part2", C.c_float), ("nof_rb", C.c_uint),
                ("start_symbol_index", C.c_uint), ("nof_symbols", C.c_uint), ("dmrs_type", C.c_uint),
                ("dmrs_symbol_mask", C.c_uint), ("nof_cdm_groups_without_data", C.c_uint), ("nof_layers", C.c_uint),
                ("contains_dc", C.c_int)]


class _UlschInfo(C.Structure):
    _fields_ = [("sch", _SchInfo)] + [(f, C.c_uint) for f in ULSCH_INFO_FIELDS]


def ulsch_information(cfg):
    """get_ulsch_information of a configuration dict (ULSCH_*_FIELDS); None if invalid."""
    c = _UlschCfg(**{f: cfg.get(f, 0) for f in ULSCH_INT_FIELDS + ULSCH_FLOAT_FIELDS})
    info = _UlschInfo()
    if lib().orc_ulsch_information(C.byref(c), C.byref(info)) != 0:
        return None
    out = {f: getattr(info, f) for f in ULSCH_INFO_FIELDS}
    if cfg.get("tbs_bits", 0) > 0:
        out.update({"sch_" + f: getattr(info.sch, f) for f in SCH_INFO_FIELDS})
    return out


def ref_ulsch_information(cfg):
    """The reference's get_ulsch_information built from source (no assertions, as its
    Release build)."""
    w = (C.c_uint * 13)(*[int(cfg.get(f, 0)) for f in ULSCH_INT_FIELDS])
    fl = (C.c_float * 5)(*[float(cfg.get(f, 0.0)) for f in ULSCH_FLOAT_FIELDS])
    out = (C.c_uint * 15)()
    ref_demod().ref_ulsch_information(w, fl, out)
    res = dict(zip(ULSCH_INFO_FIELDS, out[:9]))
    if cfg.get("tbs_bits", 0) > 0:
        res.update({"sch_" + f: v for f, v in zip(SCH_INFO_FIELDS, out[9:15])})
    return res


def compute_n_ref(tbs_lbrm_bytes, nof_codeblocks):
    """ldpc::compute_N_ref (include/ocudu/phy/upper/channel_coding/ldpc/ldpc.h:119-122)."""
    return min(tbs_lbrm_bytes * 8 * 3 // (2 * nof_codeblocks), 384 * 66)  # MAX_CODEBLOCK_SIZE


def nof_codeblocks(tbs_bits, bg):
    """compute_nof_codeblocks (include/ocudu/ran/sch/sch_segmentation.h:36-47)."""
    B = tbs_bits + (16 if tbs_bits <= 3824 else 24)
    K = 8448 if bg == 1 else 3840
    return 1 if B <= K else -(-B // (K - 24))
