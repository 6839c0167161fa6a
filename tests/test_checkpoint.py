"""Restartable runs: the on-disk ELL format and the Lanczos checkpoint.

A saved and reloaded matrix is the same matrix (export and SpMV bit for bit);
files for another basis, truncated files and files whose indexes point outside
the arrays are refused before any kernel reads them. A Lanczos run stopped and
resumed from its checkpoint (once or many times, plain or with partial or full
reorthogonalisation) ends with the same eigenvalues, step count and Krylov
coefficients as the uninterrupted run. Independently of libqmb, a saved file
read by the oracle's numpy reader is the oracle's own matrix bit for bit, and
resumed runs reach the oracle's converged and the dense eigenvalues.
"""
import struct
import sys

import numpy as np
import pytest

import oracle
import qmbcuda_amd as q
from oracle.qell import read_qell
from tests.helpers import chain_case, ell_dense_nonnull, oracle_ell, oracle_orbits, square_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def objects(ctx, case, n_up, alpha, layout=q.LAYOUT_COMPACT, **kw):
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(case["n_sites"], case["table"])
    b = q.Basis(ctx, g, n_up, alpha, layout)
    return g, t, b, q.Ell(ctx, b, t, **kw)


@pytest.mark.parametrize("alpha,layout,fp64", [(0, q.LAYOUT_COMPACT, False), (0, q.LAYOUT_REFERENCE, False),
                                               (5, q.LAYOUT_COMPACT, False), (0, q.LAYOUT_COMPACT, True)])
def test_ell_round_trip(ctx, tmp_path, alpha, layout, fp64):
    case = chain_case(14, 1.0, 1.0, 0.3, flip=True)
    say(f"ell save/load irrep={alpha} layout={layout} fp64={fp64}")
    g, t, b, e = objects(ctx, case, 7, alpha, layout, keep_fp64_values=fp64)
    path = tmp_path / "m.qell"
    e.save(path)
    e2 = q.Ell.load(ctx, b, path, t)
    i1, i2 = e.info(), e2.info()
    for f in ("n_rows", "nnz", "stored", "max_row_len", "max_terms", "value_type", "packed", "n_values"):
        assert getattr(i1, f) == getattr(i2, f)
    for u, v in zip(e.export_reference(), e2.export_reference()):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    x = np.random.default_rng(3).normal(size=b.n_rows)
    if e.is_complex:
        x = x + 1j * np.random.default_rng(4).normal(size=b.n_rows)
    assert np.array_equal(e.spmv(x), e2.spmv(x))


def test_ell_load_refuses_bad_files(ctx, tmp_path):
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    g, t, b, e = objects(ctx, case, 6, 0)
    path = tmp_path / "m.qell"
    e.save(path)
    raw = path.read_bytes()
    say("ell load: another basis")
    b_other = q.Basis(ctx, g, 6, 3, q.LAYOUT_COMPACT)
    with pytest.raises(q.QMBError):
        q.Ell.load(ctx, b_other, path)
    say("ell load: truncated")
    (tmp_path / "short.qell").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(q.QMBError):
        q.Ell.load(ctx, b, tmp_path / "short.qell")
    say("ell load: column out of range")
    # header (see checkpoint.hip): 8 + 4 + 4 + 6*4 + 4*8 + 4*4 + 8 + 2*4 bytes
    hdr = 8 + 4 + 4 + 24 + 32 + 16 + 8 + 8
    n_rows, n_chunks, nnz, stored = struct.unpack_from("<qqqq", raw, 8 + 4 + 4 + 24)
    off_col = hdr + (n_chunks + 1) * 8 + n_rows * 2
    bad = bytearray(raw)
    # slot 1 of row 0 (chunk 0 starts at offset 0): an index far beyond n_rows
    struct.pack_into("<I", bad, off_col + 64 * 4, (1 << 20) - 1)
    (tmp_path / "bad.qell").write_bytes(bytes(bad))
    with pytest.raises(q.QMBError):
        q.Ell.load(ctx, b, tmp_path / "bad.qell")
    # the good file still loads
    q.Ell.load(ctx, b, path)


@pytest.mark.parametrize("reorth,n_eigs", [(0, 1), (2, 1), (1, 3)])
def test_lanczos_resume_equals_uninterrupted(ctx, tmp_path, reorth, n_eigs):
    case = chain_case(20, 1.0, 1.0, 0.45, flip=True)
    g, t, b, e = objects(ctx, case, 10, 0)
    say(f"lanczos checkpoint reorth={reorth} n_eigs={n_eigs}")
    kw = dict(n_eigs=n_eigs, max_iter=300, tol=1e-13, reorth=reorth, check_every=5)
    ev0, rep0, kr0, _, ok0 = q.lanczos_checkpointed(e, tmp_path / "unused.qlz", **kw)
    assert ok0
    path = tmp_path / "run.qlz"
    ev, rep, kr, _, ok = q.lanczos_checkpointed(e, path, stop_after=7, **kw)
    assert not ok and rep.iters < rep0.iters and path.exists()
    stops = 1
    while not ok:
        ev, rep, kr, _, ok = q.lanczos_checkpointed(e, path, resume=True, stop_after=9, **kw)
        stops += 1
        assert stops < 100
    assert stops >= 3
    assert np.array_equal(ev, ev0)
    assert rep.iters == rep0.iters and rep.n_reorth == rep0.n_reorth
    assert np.array_equal(kr, kr0)


def test_lanczos_resume_refuses_other_options(ctx, tmp_path):
    case = chain_case(16, 1.0, 1.0, 0.3, flip=True)
    g, t, b, e = objects(ctx, case, 8, 0)
    path = tmp_path / "run.qlz"
    say("lanczos checkpoint: mismatched options")
    q.lanczos_checkpointed(e, path, stop_after=5, max_iter=200, check_every=5)
    with pytest.raises(q.QMBError):
        q.lanczos_checkpointed(e, path, resume=True, max_iter=200, check_every=5, seed=7)
    _, _, b2, e2 = objects(ctx, case, 8, 3)
    with pytest.raises(q.QMBError):
        q.lanczos_checkpointed(e2, path, resume=True, max_iter=200, check_every=5)


@pytest.mark.parametrize("label,mk,n_up,alpha,layout,fp64", [
    ("chain14_flip", lambda: chain_case(14, 1.0, 1.0, 0.3, flip=True), 7, 0, q.LAYOUT_COMPACT, False),
    ("chain14_flip", lambda: chain_case(14, 1.0, 1.0, 0.3, flip=True), 7, 0, q.LAYOUT_REFERENCE, False),
    ("chain14_flip", lambda: chain_case(14, 1.0, 1.0, 0.3, flip=True), 7, 5, q.LAYOUT_COMPACT, False),
    ("chain14_flip", lambda: chain_case(14, 1.0, 1.0, 0.3, flip=True), 7, 5, q.LAYOUT_REFERENCE, True),
    ("square4x4_flip", lambda: square_case(4, 4, 1.0, 0.8, flip=True), 8, 0, q.LAYOUT_COMPACT, False),
])
def test_saved_file_is_the_oracle_matrix(ctx, tmp_path, label, mk, n_up, alpha, layout, fp64):
    """A saved matrix read back by the independent numpy reader (oracle/qell.py,
    no libqmb) equals the oracle's own reference-layout build bit for bit."""
    case = mk()
    say(f"qell vs oracle {label} irrep={alpha} layout={layout} fp64={fp64}")
    orb = oracle_orbits(case, n_up)
    ref = oracle_ell(case, orb, alpha)
    g, t, b, e = objects(ctx, case, n_up, alpha, layout, keep_fp64_values=fp64, max_terms=ref.W)
    path = tmp_path / "m.qell"
    e.save(path)
    hd, ind, val, track = read_qell(path)
    assert hd["n_sites"] == case["n_sites"] and hd["n_up"] == n_up and hd["irrep"] == alpha
    assert hd["packed"] == (0 if fp64 else 1)
    W = ref.W
    if layout == q.LAYOUT_REFERENCE:
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val, ref.values)
    else:
        keep = ~orb.null_mask(alpha)
        cidx = np.cumsum(keep) - 1
        rows = np.flatnonzero(keep)
        assert hd["n_rows"] == len(rows)
        for i, r in enumerate(rows):
            n = int(ref.track[r])
            assert track[i] == n
            assert ind[i * W] == i
            assert np.array_equal(ind[i * W + 1:i * W + n], cidx[ref.ind[r * W + 1:r * W + n]])
            assert np.array_equal(val[i * W:i * W + n], ref.values[r * W:r * W + n])


@pytest.mark.parametrize("fp64", [False, True])
def test_saved_locality_file_is_the_oracle_matrix(ctx, tmp_path, fp64):
    """A matrix built with QMB_LOCALITY_SLOTS, saved and read back by the numpy
    reader, equals the oracle's reference-layout build of the table in locality
    order bit for bit; loading it back gives the same SpMV as the built one."""
    case = square_case(4, 4, 1.0, 0.8, flip=True)
    say(f"qell locality slots vs oracle fp64={fp64}")
    orb = oracle_orbits(case, 8)
    ref = oracle_ell(dict(case, table=oracle.locality_order(case["table"])), orb, 0)
    g, t, b, e = objects(ctx, case, 8, 0, q.LAYOUT_REFERENCE, keep_fp64_values=fp64, max_terms=ref.W,
                         locality_slots=True)
    path = tmp_path / "m.qell"
    e.save(path)
    hd, ind, val, track = read_qell(path)
    assert hd["packed"] == (0 if fp64 else 1)
    assert np.array_equal(track, ref.track)
    assert np.array_equal(ind, ref.ind)
    assert np.array_equal(val, ref.values)
    e2 = q.Ell.load(ctx, b, path, terms=t)
    x = np.random.default_rng(5).normal(size=b.n_rows)
    assert np.array_equal(e2.spmv(x), e.spmv(x))


def test_padding_words_are_never_read(ctx, tmp_path):
    """A corrupt padding word (a slot past its row's length inside the chunk
    width) is accepted by the loader and changes nothing: neither the SpMV nor
    the export reads slots past a row's length."""
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    g, t, b, e = objects(ctx, case, 6, 0)
    path = tmp_path / "m.qell"
    e.save(path)
    raw = path.read_bytes()
    hd, _, _, _ = read_qell(path)
    n, nc = hd["n_rows"], hd["n_chunks"]
    hdr = 104
    cp = np.frombuffer(raw[hdr:hdr + (nc + 1) * 8], "<i8")
    rl = np.frombuffer(raw[hdr + (nc + 1) * 8:hdr + (nc + 1) * 8 + n * 2], "<u2")
    off_col = hdr + (nc + 1) * 8 + n * 2
    # the first row shorter than its chunk's width
    pad = None
    for r in range(n):
        L = (cp[r // 64 + 1] - cp[r // 64]) // 64
        if rl[r] < L:
            pad = cp[r // 64] + 64 * int(rl[r]) + r % 64
            break
    assert pad is not None
    say("ell load: corrupt padding word")
    bad = bytearray(raw)
    struct.pack_into("<I", bad, off_col + 4 * int(pad), 0xFFFFFFFF)
    (tmp_path / "pad.qell").write_bytes(bytes(bad))
    e2 = q.Ell.load(ctx, b, tmp_path / "pad.qell", t)
    for u, v in zip(e.export_reference(), e2.export_reference()):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    x = np.random.default_rng(3).normal(size=b.n_rows)
    assert np.array_equal(e.spmv(x), e2.spmv(x))


@pytest.mark.parametrize("reorth", [0, 2])
def test_resumed_run_matches_oracle_and_dense(ctx, tmp_path, reorth):
    """A run stopped three times and resumed from its checkpoint reaches the
    oracle's converged Lanczos eigenvalue and the dense ground state (1e-10)."""
    case = chain_case(16, 1.0, 0.8, 0.35, flip=True)
    orb = oracle_orbits(case, 8)
    ref = oracle_ell(case, orb, 0)
    dense = np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, 0))[0]
    st, e0_or, _, _, _ = ref.lanczos_converge(400, 2212456, tol=1e-13, check_every=5)
    assert st == 0
    g, t, b, e = objects(ctx, case, 8, 0)
    say(f"resumed checkpointed run vs oracle reorth={reorth}")
    kw = dict(max_iter=300, tol=1e-13, reorth=reorth, check_every=5)
    path = tmp_path / "run.qlz"
    ev, rep, _, _, ok = q.lanczos_checkpointed(e, path, stop_after=6, **kw)
    stops = 0
    while not ok:
        ev, rep, _, _, ok = q.lanczos_checkpointed(e, path, resume=True, stop_after=6, **kw)
        stops += 1
        assert stops < 100
    assert stops >= 2 and rep.converged
    assert abs(ev[0] - e0_or) < 1e-10 * abs(e0_or)
    assert abs(ev[0] - dense) < 1e-10 * abs(dense)
