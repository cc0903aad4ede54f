"""The C-ABI library loads and exports every entry point declared in
include/ocudu_hip_ldpc.h (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import ocudu_hip as H

HEADER = Path(__file__).resolve().parent.parent / "include" / "ocudu_hip_ldpc.h"


def declared():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^(?:int|const char\*|void|size_t)\s+(lh_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "lh_pusch_decode_batch" in names and "lh_ldpc_decode" in names
    assert set(names) == set(H.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(H.LIB_PATH))
    for name in declared():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    # sizes of the ctypes mirrors equal the C layout (natural alignment)
    assert ctypes.sizeof(H.DecoderConfig) == 20
    assert ctypes.sizeof(H.DecodeDesc) == 8 + 8 + 4 + 4 + 20 + 4  # padded to 8
    assert ctypes.sizeof(H.RmDesc) == 8 + 8 + 7 * 4 + 4
    assert ctypes.sizeof(H.PuschCbDesc) == ctypes.sizeof(H.RmDesc) + 8 + 8
    assert ctypes.sizeof(H.TbDesc) == 40
    assert ctypes.sizeof(H.GenDesc) == 5 * 8 + 10 * 4
    assert ctypes.sizeof(H.GenTbDesc) == 48
    assert ctypes.sizeof(H.EncodeDesc) == 3 * 8 + 8 * 4
    assert ctypes.sizeof(H.DemodDesc) == 2 * 8 + 4 * 4
    assert ctypes.sizeof(H.Segmentation) == 32
    assert ctypes.sizeof(H.TbRequest) == 8 + 8 + 10 * 4
    assert ctypes.sizeof(H.TbResult) == 8 + 6 * 4 + 8
    assert ctypes.sizeof(H.ServerConfig) == 8 + 4 * 4 + 4 + 4
    assert ctypes.sizeof(H.TrainDesc) == 2 * 8 + 4 * 4
    assert ctypes.sizeof(H.UlschDemuxDesc) == 5 * 8 + 16 * 4
    assert ctypes.sizeof(H.UlschDemuxSizes) == 5 * 4
    assert ctypes.sizeof(H.UlschConfig) == 18 * 4
    assert ctypes.sizeof(H.UlschInfo) == 16 * 4


def test_invalid_arguments_fail_without_gpu():
    lib = H.lib()
    assert lib.lh_ctx_create(None, 0, 16) != 0
    assert b"invalid" in lib.lh_last_error()
