"""The C++ facade (include/qmb/*.hpp), driven like the reference's own example.

CPU part: the facade's term tables, group arrays and characters equal the
oracle's restatement of the reference's model and symmetry code.
GPU part: HeisenbergXXZAbelianSymms solves every irrep; energies match dense
diagonalisation of the oracle's sector matrices.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import model as OM
from tests.helpers import dense_spectrum, ell_dense_nonnull, oracle_ell, oracle_orbits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_facade")


def run(mode, case):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_facade is not built (run __graft_entry__.build())")
    out = subprocess.run([EXE, mode, case], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def oracle_case(name):
    if name == "chain16":
        geom, info = OM.chain_xxz_info(16, 1.0, 1.0, 0.0)
        flip = False
    elif name == "chain12_mg_flip":
        geom, info = OM.chain_xxz_info(12, 1.0, 1.0, 0.5)
        flip = True
    elif name == "square2x4_example":
        geom = OM.square_geometry(2, 4)
        info = OM.extended_xxx_info(geom, [1.0, 0.15])
        flip = False
    elif name == "triangular3x3":
        geom = OM.triangular_geometry(3, 3)
        info = OM.extended_xxx_info(geom, [1.0])
        flip = False
    elif name == "square4x4_flip":
        geom, info = OM.square_xxz_info(4, 4, 1.0, 1.0)
        flip = True
    elif name == "honeycomb2x3_phase":
        from tests.test_lattices import honeycomb

        geom, info, _, _ = honeycomb(2, 3)
        flip = False
    terms = OM.heisenberg_operator(geom, info)
    table = OM.conjugated_term_table(terms)
    ge, ch = OM.translation_group(geom)
    fl = None
    if flip:
        ge, fl, ch = OM.with_spin_flip(ge, ch)
    return dict(n_sites=geom.n_sites, terms=terms, table=table, group_el=ge, chars=ch, flip=fl)


CASES = ["chain16", "chain12_mg_flip", "square2x4_example", "triangular3x3", "square4x4_flip", "honeycomb2x3_phase"]


@pytest.mark.parametrize("name", CASES)
def test_facade_tables_match_oracle(name):
    got = run("tables", name)
    ref = oracle_case(name)
    op, site, sc, nops, hcol = ref["table"]
    assert got["H_size"] == len(nops) and got["H_col"] == hcol
    assert got["op"] == op.tolist()
    assert got["site"] == site.tolist()
    np.testing.assert_array_equal(np.array(got["scalar"]).reshape(-1, 2), np.c_[sc.real, sc.imag])
    G = ref["group_el"].shape[0]
    assert np.array_equal(np.array(got["group_el"]).reshape(G, -1), ref["group_el"])
    if ref["flip"] is not None:
        assert got["flip"] == ref["flip"].tolist()
    else:
        assert not any(got["flip"])
    chars = np.array(got["chars"]).reshape(ref["chars"].shape[0], G, 2)
    np.testing.assert_array_equal(chars[..., 0] + 1j * chars[..., 1], ref["chars"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,n_up", [("chain16", 8), ("chain12_mg_flip", 6), ("square2x4_example", 4),
                                       ("triangular3x3", 4), ("honeycomb2x3_phase", 6)])
def test_facade_driver_solves_every_irrep(name, n_up):
    print(f"[gpu-case] facade solve {name}", flush=True)
    got = run("solve", name)
    ref = oracle_case(name)
    orb = oracle_orbits(ref, n_up)
    E = got["E"]
    assert len(E) == orb.n_irreps
    lowest = []
    for a in range(orb.n_irreps):
        ell = oracle_ell(ref, orb, a)
        M = ell_dense_nonnull(ell, orb, a)
        if M.shape[0] == 0:
            assert E[a] is None
            continue
        e0 = np.linalg.eigvalsh(M)[0]
        lowest.append(e0)
        assert abs(E[a] - e0) < 1e-10 * max(1, abs(e0))
    dense = dense_spectrum(ref, n_up)
    assert abs(got["E_gs"] - dense[0]) < 1e-10 * max(1, abs(dense[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["chain16", "square2x4_example"])
def test_facade_driver_batched_equals_sequential(name):
    """EDOptions::batched solves every irrep in one batched call: same energies,
    same ground-state irrep, a Krylov vector for it."""
    print(f"[gpu-case] facade solve_batched {name}", flush=True)
    seq = run("solve", name)
    bat = run("solve_batched", name)
    assert bat["GS_irrep"] == seq["GS_irrep"]
    assert abs(bat["E_gs"] - seq["E_gs"]) < 1e-10 * max(1, abs(seq["E_gs"]))
    for a, b in zip(seq["E"], bat["E"]):
        assert (a is None) == (b is None)
        if a is not None:
            assert abs(a - b) < 1e-10 * max(1, abs(a))
    assert bat["krylov_len"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["chain16", "square2x4_example"])
def test_facade_locality_slots(name):
    """EDOptions::locality_slots (QMB_LOCALITY_SLOTS on every ELL the driver
    builds): the same energies per irrep, the same ground-state irrep, and the
    same ground-state expectation values (the regenerated state uses the same
    slot order as the first solve)."""
    print(f"[gpu-case] facade solve_locality {name}", flush=True)
    seq = run("solve", name)
    loc = run("solve_locality", name)
    assert loc["GS_irrep"] == seq["GS_irrep"]
    for a, b in zip(seq["E"], loc["E"]):
        assert (a is None) == (b is None)
        if a is not None:
            assert abs(a - b) < 1e-10 * max(1, abs(a))
    ex, exl = run("expect", name), run("expect_locality", name)
    assert abs(ex["E_gs"] - exl["E_gs"]) < 1e-10 * max(1, abs(ex["E_gs"]))
    for key, tol in (("szsz", 1e-5), ("H", 1e-10)):  # the state carries the Lanczos tolerance
        v, w = complex(*ex[key]), complex(*exl[key])
        assert abs(v - w) < tol * max(1, abs(v)), (key, v, w)


@pytest.mark.gpu
def test_facade_fixed_iteration_mode():
    print("[gpu-case] facade solve_fixed chain16", flush=True)
    got = run("solve_fixed", "chain16")
    assert abs(got["E_gs"] - (-7.142296360617)) < 1e-8
    assert got["krylov_len"] == 60


@pytest.mark.gpu
@pytest.mark.parametrize("name,n_up", [("chain16", 8), ("square2x4_example", 4)])
def test_facade_ground_state_expectation_values(name, n_up):
    """ComputeStaticExpValZeroT on the regenerated ground state against <Psi|A|Psi>
    from dense diagonalisation of the ground-state sector (the facade's vector
    carries the Lanczos tolerance: 1e-5 for operators that do not commute with H)."""
    import math

    from tests.helpers import embedding, sector_ground_state

    print(f"[gpu-case] facade expect {name}", flush=True)
    got = run("expect", name)
    ref = oracle_case(name)
    orb = oracle_orbits(ref, n_up)
    a = got["GS_irrep"]
    e0, x = sector_ground_state(ref, orb, a)
    assert abs(got["E_gs"] - e0) < 1e-10 * max(1, abs(e0))
    psi = embedding(ref, orb, a) @ np.conj(x)
    n = ref["n_sites"]
    n1, n2 = (16, 1) if name == "chain16" else (2, 4)
    ph = [np.exp(2j * math.pi * (s // n2) / n1) for s in range(n)]
    SZ, SP, SM = OM.OP_SZ, OM.OP_SP, OM.OP_SM
    ops = {
        "szsz": [[(0, SZ, 1.0), (3, SZ, 1.0)]],
        "hop": [[(0, SP, 0.4 + 0.9j), (1, SM, 1.0)]],
        "szz": [[(r, SZ, np.conj(ph[r]) / math.sqrt(n)), (rp, SZ, ph[rp] / math.sqrt(n))]
                for r in range(n) for rp in range(n)],
        "H": ref["terms"],
    }
    for key, terms in ops.items():
        want = psi.conj() @ OM.dense_sector_hamiltonian(n, n_up, terms) @ psi
        v = complex(*got[key])
        tol = 1e-10 if key == "H" else 1e-5
        assert abs(v - want) < tol * max(1, abs(want)), (key, v, want)


def product_case(name):
    import qmbcuda_amd.models as LM

    if name == "chain12_reflection":
        lat, info = LM.chain(12, 1.0, 1.0, 0.4)
        P = LM.reflection_group(lat)
        n_up = 6
    else:
        lat, info = LM.square(4, 4, 1.0, 1.0)
        P = LM.rotation_group(lat, 4)
        n_up = 8
    ge, fl, ch, pairs = LM.product_group(LM.translation_group(lat, spin_flip=True), P)
    return dict(n_sites=lat.n_sites, table=LM.term_table(lat, info), group_el=ge, chars=ch, flip=fl), n_up


@pytest.mark.parametrize("name", ["chain12_reflection", "square4x4_c4"])
def test_facade_product_group_tables(name):
    """ProductGroup (translations x spin flip x a point operation, folded into the
    torus) gives the element arrays, flips and character rows of the Python
    builder."""
    got = run("tables", name)
    ref, _ = product_case(name)
    G = len(ref["group_el"])
    assert np.array_equal(np.array(got["group_el"]).reshape(G, -1), ref["group_el"])
    assert np.array_equal(np.array(got["flip"]), ref["flip"])
    chars = np.array(got["chars"]).reshape(len(ref["chars"]), G, 2)
    np.testing.assert_allclose(chars[..., 0] + 1j * chars[..., 1], ref["chars"], atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["chain12_reflection", "square4x4_c4"])
def test_facade_driver_with_point_group(name):
    print(f"[gpu-case] facade solve {name}", flush=True)
    got = run("solve", name)
    ref, n_up = product_case(name)
    orb = oracle_orbits(ref, n_up)
    E = got["E"]
    assert len(E) == orb.n_irreps
    for a in range(orb.n_irreps):
        M = ell_dense_nonnull(oracle_ell(ref, orb, a, max_terms=40), orb, a)
        if M.shape[0] == 0:
            assert E[a] is None
            continue
        e0 = np.linalg.eigvalsh(M)[0]
        assert abs(E[a] - e0) < 1e-10 * max(1, abs(e0))


@pytest.mark.gpu
def test_facade_reference_free_function_on_device_arrays():
    """LanczosEigVals_CUDA with borrowed device pointers equals LanczosEigVals on
    the same host arrays bit for bit, both reach the dense ground state of the
    irrep, and the driver answers to the reference's class name."""
    print("[gpu-case] facade ref_cuda chain16", flush=True)
    got = run("ref_cuda", "chain16")
    assert got["E_dev"] == got["E_host"] and got["res_dev"] == got["res_host"] and got["same_krylov"]
    ref = oracle_case("chain16")
    orb = oracle_orbits(ref, 8)
    e0 = np.linalg.eigvalsh(ell_dense_nonnull(oracle_ell(ref, orb, 0), orb, 0))[0]
    assert abs(got["E_dev"] - e0) < 1e-9 * abs(e0)
    assert abs(got["E_driver_irrep0"] - e0) < 1e-10 * abs(e0)
