"""The stream probe (cdh_stream_rate) on an MI355X: the three state-pattern
streams of 4M cells (600 MB) report rates between 1 TB/s and the 8 TB/s
peak, and a read-only stream is not slower than a read-and-write one."""
import sys

import pytest

from cardioid_amd import hip

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    hip.lib()
    yield


def test_stream_rates_are_plausible():
    print("stream probe, 4M cells", flush=True)
    sys.stdout.flush()
    r = {p: hip.stream_rate(p, 4 * 1024 * 1024, 5) for p in ("read", "write", "inplace")}
    print(r, flush=True)
    for p, v in r.items():
        assert 1000.0 < v < 8000.0, (p, v)
    assert r["read"] >= 0.9 * r["inplace"]
