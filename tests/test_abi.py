"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and its host-side arithmetic (counter-based normals,
cos(πx)) equals the oracle's to the bit."""
import os
import re

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(swalbe(?:1d)?_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from swalbe import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_gpu_fails_loudly_or_runs():
    import swalbe as S
    if S.device_count() == 0:
        try:
            S.HipState(5, 5)
        except S.SwalbeError:
            return
        raise AssertionError("creating a state without a GPU must raise")


def test_host_normals_match_oracle_bitwise():
    import swalbe as S
    z0, z1 = S.normals(5000, seed=77, step=3, tag=0)
    for i in range(0, 5000, 7):
        a, b = O.normal_pair(77, 3, i, 0)
        assert z0[i] == a and z1[i] == b
    u = S.uniforms(3000, seed=5)
    assert all(u[i] == O.uniform(5, 0, i, 2) for i in range(0, 3000, 11))


def test_host_cospi_matches_oracle_bitwise():
    import swalbe as S
    for x in np.concatenate([np.linspace(-3, 3, 997), [0, 0.5, 1, 1 / 9, 1 / 6, 1 / 3, 2.5]]):
        assert S.cospi(x) == O.cospi(x)


def test_host_random_film_matches_oracle():
    import swalbe as S
    h = S.random_film(13, 7, seed=9, base=1.0, amp=0.001)
    assert np.array_equal(h, O.init_normal(13, 7, 9, 1.0, 0.001))
    h = S.random_film(13, 7, seed=9, base=1.0, amp=0.01, dist="uniform")
    assert np.array_equal(h, O.init_uniform(13, 7, 9, 1.0, 0.01))


def test_singledroplet_reference_kat():
    # test/initialvalues.jl:16-34: max == rad (1 - cospi θ) at the centre
    import kat
    import swalbe as S
    case = [c for c in kat.load_cases() if c["op"] == "singledroplet"][0]

    class B:
        singledroplet = staticmethod(S.singledroplet)
        cospi = staticmethod(S.cospi)
    kat.run_case(B, case)


def test_struct_layouts_match_the_header(tmp_path):
    # sizeof / offsetof of every field of the option structs, from gcc on the
    # header, against the ctypes mirrors (and so the field order the Julia
    # binding copies)
    import ctypes as C
    import subprocess

    from swalbe import _lib
    structs = {"swalbe_params": _lib.Params, "swalbe_step_opts": _lib.StepOpts,
               "swalbe1d_step_opts": _lib.StepOpts1D, "swalbe_film_init": _lib.FilmInit}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "swalbe_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out if ln)}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
