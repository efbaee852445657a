"""The reference's known-answer tests on the GPU (FP64), through both the
positional-array op forms and the state forms of libswalbe_hip."""
import pytest

import kat

pytestmark = pytest.mark.gpu

CASES = kat.load_cases()


@pytest.fixture(scope="module", params=["raw", "state"])
def backend(request):
    import hip_backend as HB
    print(f"[gpu-kat] backend={request.param}", flush=True)
    return HB.RawBackend() if request.param == "raw" else HB.StateBackend()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_kat(backend, case):
    print(f"[gpu-kat] case={case['name']}", flush=True)
    kat.run_case(backend, case)
