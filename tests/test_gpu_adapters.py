"""The C++ adapters (ldpc_decoder / ldpc_rate_dematcher factory type "hip" and the
pusch_resident_codeblock_decoder batch backend) driven through the reference's class
interfaces by tests/cpp/adapter_chain_test.cpp, checked against the oracle."""
import subprocess
from pathlib import Path

import pytest

from gpu_util import require_gpu

BIN = Path(__file__).resolve().parent / "cpp" / "adapter_chain_test"


@pytest.mark.gpu
def test_adapter_chain():
    require_gpu()
    assert BIN.exists(), "build first (__graft_entry__.build)"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=180)
    print(p.stdout[-3000:])
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-2000:]
    assert "0 failures" in p.stdout
