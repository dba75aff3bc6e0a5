"""The C++ operator interfaces over Arrow C++ (include/gazelle_amd_arrow.h),
driven by the compiled test binary tests/cpp/test_arrow_layer: Splitter
Make / Split / Stop with the data file read back by Arrow's IPC reader
(uncompressed, LZ4_FRAME, a range split), HashAggregateKernel's
ResultIterator, and HashRelationKernel + ConditionedProbeKernel through
SetDependencies / Process, each against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_arrow_layer")


def test_cpp_interfaces_match_the_oracle():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_arrow_layer is not built (run __graft_entry__.build())")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    for case in ("split hash", "split hash lz4", "split range", "aggregate", "join inner", "join semi",
                 "join anti", "join null-aware anti"):
        assert f"PASS {case}\n" in res.stdout, case
    assert res.stdout.rstrip().endswith("ALL PASS")
