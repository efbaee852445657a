"""The store modes reach the machine code (CPU test of the built library).

The fused steps choose their stores' cache policy (store mode 0: cached, 1:
every output streaming, 2: distributions and pressure streaming) as a
template argument. Chosen by a run-time branch instead, the compiler merged
the streaming and the plain form of each store into one plain store: every
mode then ran cached stores, and the tests that compare the modes' results
(the same bits by design) could not tell. This test disassembles the
library's gfx950 code objects and counts the non-temporal stores of the
kernels whose modes are timed: the FP64 step per mode, the FP32 single-film
and batch kernels, and the 1D per-step kernel."""
import glob
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "swalbe.jl_amd", "lib", "libswalbe_hip.so")
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"


@pytest.fixture(scope="module")
def stores():
    """{kernel symbol: (global stores, of which non-temporal)}."""
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    if not os.path.exists(OBJDUMP):
        pytest.skip("llvm-objdump not found")
    d = tempfile.mkdtemp()
    try:
        so = os.path.join(d, "lib.so")
        shutil.copy(LIB, so)
        subprocess.run([OBJDUMP, "--offloading", so], cwd=d, check=True, capture_output=True,
                       timeout=120)
        objs = sorted(glob.glob(os.path.join(d, "lib.so.*gfx950")))
        assert objs, "no gfx950 code object in the library"
        found = {}
        for o in objs:
            out = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", o], check=True,
                                 capture_output=True, text=True, timeout=300).stdout
            cur = None
            for line in out.splitlines():
                m = re.match(r"^[0-9a-f]+ <(\S+)>:", line)
                if m:
                    cur = m.group(1)
                    continue
                if cur and "global_store" in line:
                    n, nt = found.get(cur, (0, 0))
                    found[cur] = (n + 1, nt + (1 if re.search(r"\bnt\b", line) else 0))
        return found
    finally:
        shutil.rmtree(d)


def pick(stores, *parts):
    ks = [k for k in stores if all(p in k for p in parts)]
    assert ks, parts
    return [stores[k] for k in ks]


def test_fp64_step_follows_the_mode(stores):
    # k_step<double, 8, READ_FT = false, THERMAL = true, NTS, AL>: per node 9 f, h, ux,
    # uy and p (with AL, p is stored ahead of phase D, once per row of the thread)
    for mode in (0, 1, 2):
        for n, nt in pick(stores, f"k_stepIdLi8ELb0ELb1ELi{mode}E"):
            want = {0: 0, 1: n, 2: n - 3}[mode]  # mode 2: h, ux, uy cached
            assert n >= 13 and nt == want, (mode, n, nt)


def test_fp32_kernels_carry_streaming_and_cached_copies(stores):
    # one copy of phase D per mode in each kernel: some stores stream, some not
    for parts in (("k_step_f32ILi64ELi256ELi16ELb0ELb1ELb0ELb1E",),
                  ("k_step_f32ILi64ELi256ELi16ELb0ELb1ELb0ELb0E",),
                  ("k_batch_f32ILb0E",), ("k_batch_f32ILb1E",)):
        for n, nt in pick(stores, *parts):
            assert 0 < nt < n, (parts, n, nt)


def test_1d_step_streams_f_and_p(stores):
    # per output node: 3 f and p streaming, h and u cached
    for n, nt in pick(stores, "k1_step"):
        assert n % 6 == 0 and 3 * nt == 2 * n, (n, nt)


def test_store_mode_rule():
    """The default store mode of time_loop and bench.py: the mode measured
    fastest for the film and the pass."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "swalbe.jl_amd"))
    import swalbe as S
    assert S.store_mode(4096, 4096, False, False) == 1  # large FP32, every output written
    assert S.store_mode(4096, 4096, False, True) == 0   # large FP32, final outputs only
    assert S.store_mode(1024, 1024, False, False) == 0  # 12.6 MB of inputs: stays in L2
    assert S.store_mode(1024, 1024, False, True) == 1
    assert S.store_mode(4096, 4096, True) == S.store_mode(64, 48, True, True) == 2
    # the boundary: 32 MB of FP32 inputs (h, ux, uy)
    assert S.store_mode(1024, 2731, False) == 1 and S.store_mode(1024, 2730, False) == 0
