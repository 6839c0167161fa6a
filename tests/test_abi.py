"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and rejects bad input without a GPU (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import qmbcuda_amd as q
from qmbcuda_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(qmb_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _abi.SIGNATURES}
    assert names == bound, names ^ bound


def test_device_count_without_gpu_is_safe():
    assert q.device_count() >= 0


def test_group_validation_errors():
    lib = q.lib()
    h = C.c_void_p()
    ge = np.array([[1, 0], [0, 1]], np.uint32)  # element 0 is not the identity
    ch = np.ones((2, 2), np.complex128).view(np.float64)
    st = lib.qmb_group_create(2, 2, ge.ctypes.data, None, 2, ch.ctypes.data, C.byref(h))
    assert st == 1 and b"identity" in lib.qmb_last_error()
    ge = np.array([[0, 1], [0, 0]], np.uint32)  # not a permutation
    st = lib.qmb_group_create(2, 2, ge.ctypes.data, None, 2, ch.ctypes.data, C.byref(h))
    assert st == 1
    ge = np.array([[0, 1, 2], [1, 2, 0]], np.uint32)  # not closed under inversion
    ch = np.ones((2, 2), np.complex128).view(np.float64)
    st = lib.qmb_group_create(3, 2, ge.ctypes.data, None, 2, ch.ctypes.data, C.byref(h))
    assert st == 1 and b"inversion" in lib.qmb_last_error()


def test_group_detects_torus_path():
    for lat in (q.models.Lattice(12), q.models.Lattice(4, 4), q.models.Lattice(6, 6), q.models.Lattice(3, 2, 2)):
        ge, fl, ch = q.models.translation_group(lat, spin_flip=True)
        g = q.Group(ge, ch, fl)
        inv, torus = g.info()
        assert torus
        G = ge.shape[0]
        for e in range(G):  # inverse composes to the identity
            comp = ge[inv[e]][ge[e]]
            assert np.array_equal(comp, np.arange(lat.n_sites)) and fl[inv[e]] == fl[e]


def test_group_generic_path_for_reflections():
    # translations + reflection of a 6-ring: not all cyclic -> generic permutation path
    n = 6
    els = []
    for t in range(n):
        els.append([(s + t) % n for s in range(n)])
    for t in range(n):
        els.append([(t - s) % n for s in range(n)])
    ge = np.array(els, np.uint32)
    ch = np.ones((1, 2 * n), np.complex128)
    g = q.Group(ge, ch)
    inv, torus = g.info()
    assert not torus


def test_terms_reject_sz_violation_and_bad_ops():
    lib = q.lib()
    h = C.c_void_p()
    op = np.array([1, 1], np.int32)  # Sp Sp: changes Sz by 2
    site = np.array([0, 1], np.int32)
    sc = np.array([1.0, 0.0])
    nops = np.array([2], np.int32)
    st = lib.qmb_terms_create(4, 1, 2, op.ctypes.data, site.ctypes.data, sc.ctypes.data, nops.ctypes.data,
                              C.byref(h))
    assert st == 1 and b"Sz" in lib.qmb_last_error()
    op = np.array([3, 2], np.int32)  # fermion operator type
    st = lib.qmb_terms_create(4, 1, 2, op.ctypes.data, site.ctypes.data, sc.ctypes.data, nops.ctypes.data,
                              C.byref(h))
    assert st == 6
    op = np.array([1, 2], np.int32)
    site = np.array([0, 9], np.int32)  # site out of range
    st = lib.qmb_terms_create(4, 1, 2, op.ctypes.data, site.ctypes.data, sc.ctypes.data, nops.ctypes.data,
                              C.byref(h))
    assert st == 1


def test_ctx_create_without_gpu_reports_no_device():
    if q.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert q.lib().qmb_ctx_create(0, C.byref(h)) == 5


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the header's sizes and offsets
    (a C program compiled against include/qmb.h prints them)."""
    import subprocess

    structs = {
        "qmb_lanczos_opts": _abi.LanczosOpts,
        "qmb_lanczos_report": _abi.LanczosReport,
        "qmb_basis_info": _abi.BasisInfo,
        "qmb_ell_info": _abi.EllInfo,
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "qmb.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
