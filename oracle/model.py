"""Restatement of the reference's host-side model construction (test infrastructure).

TEST INFRASTRUCTURE ONLY: tests and the benchmark's CPU-baseline leg use this to
produce the term tables, group-element arrays and character tables that the
HIP library's own host helpers must reproduce, and to build dense full-sector
Hamiltonians for the dense oracle. The product library never imports it.

Every function cites the reference code it restates (paths relative to the
reference repository root; Q/ = QMBCuda/). Floating coordinates are kept in
float64 where the reference uses float32; the convention (which site goes
where, in which order) is the reference's.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

OP_SZ, OP_SP, OP_SM = 0, 1, 2  # Q/quantum_operators/SingleParticleOperators.h:4 (enum order)


@dataclass
class Geometry:
    """LatticeGeometryInfo (Q/geometry/base.h:5-11)."""

    rx_alpha: list
    ry_alpha: list
    A_mat: list  # A_mat[row][col]; columns are the primitive vectors
    N1: int
    N2: int

    @property
    def M(self) -> int:
        return len(self.rx_alpha)

    @property
    def n_sites(self) -> int:
        return self.N1 * self.N2 * self.M


def square_geometry(n1: int, n2: int) -> Geometry:
    """create_square_lattice_info (Q/geometry/base.cpp:54-69)."""
    return Geometry([0.0], [0.0], [[1.0, 0.0], [0.0, 1.0]], n1, n2)


def triangular_geometry(n1: int, n2: int) -> Geometry:
    """create_triangular_lattice_info (Q/geometry/base.cpp:71-87)."""
    sq3 = math.sqrt(3.0) / 2.0
    return Geometry([0.0], [0.0], [[1.0, 0.5], [0.0, sq3]], n1, n2)


def site_coordinates(geom: Geometry):
    """SpinHamiltonian::CreateGeometry (Q/lattice_models/ArbitrarySpinLattice.cu:19-36):
    site index (i*N2 + j)*M + alpha, position i*a1 + j*a2 + r_alpha."""
    a1 = (geom.A_mat[0][0], geom.A_mat[1][0])
    a2 = (geom.A_mat[0][1], geom.A_mat[1][1])
    r1, r2 = [], []
    for i in range(geom.N1):
        for j in range(geom.N2):
            for al in range(geom.M):
                r1.append(i * a1[0] + j * a2[0] + geom.rx_alpha[al])
                r2.append(i * a1[1] + j * a2[1] + geom.ry_alpha[al])
    return np.array(r1), np.array(r2)


@dataclass
class HeisenbergInfo:
    """HeisenbergInfo<T> (Q/lattice_models/Heisenberg.h:10-30). Each term row is
    (di, dj, a, b): target cell offset, target sublattice a, source sublattice b."""

    Jxy_terms: list = field(default_factory=list)
    Jxy_couplings: list = field(default_factory=list)  # complex allowed
    Jz_terms: list = field(default_factory=list)
    Jz_couplings: list = field(default_factory=list)  # real
    intra_c_bool: bool = True
    inter_c_bool: bool = True


def heisenberg_operator(geom: Geometry, info: HeisenbergInfo):
    """Heisenberg<T>::CreateHamiltonian (Q/lattice_models/Heisenberg.cu:22-79).

    Returns the ManyBodyOperator as a list of terms; each term is a list of
    (site, op_type, scalar). For every unit cell uc (row-major), the Jxy list
    then the Jz list; a Jxy row adds J * Sp_i * Sm_j, a Jz row adds Jz * Sz_i * Sz_j
    (the scalar rides on the first operator: ArbitraryOperator.cu:153-201).
    The boundary filter is the reference's literally, including its
    inter_c_bool condition (:57), which drops every term when inter_c_bool is false.
    """
    M, N1, N2 = geom.M, geom.N1, geom.N2
    terms = []
    if not info.intra_c_bool and not info.inter_c_bool:
        return terms
    lists = [info.Jxy_terms, info.Jz_terms]
    for uc in range(N1 * N2):
        j1, j2 = uc // N2, uc % N2
        for jii in range(2):
            for ji, row in enumerate(lists[jii]):
                j_site = uc * M + row[3]
                i1 = j1 + row[0]
                i2 = j2 + row[1]
                if not info.intra_c_bool and ((i1 < 0 or i1 >= N1) or (i2 < 0 or i2 >= N2)):
                    continue
                if not info.inter_c_bool and ((i1 >= 0 or i1 < N1) or (i2 >= 0 or i2 < N2)):
                    continue
                i1 = ((i1 % N1) + N1) % N1
                i2 = ((i2 % N2) + N2) % N2
                i_site = row[2] + M * (N2 * i1 + i2)
                if jii == 0:
                    terms.append([(i_site, OP_SP, complex(info.Jxy_couplings[ji])), (j_site, OP_SM, 1.0 + 0j)])
                else:
                    terms.append([(i_site, OP_SZ, complex(info.Jz_couplings[ji])), (j_site, OP_SZ, 1.0 + 0j)])
    return terms


def extended_xxx_info(geom: Geometry, J_vec, intra=True, inter=True) -> HeisenbergInfo:
    """CreateExtendedHeisenbergXXX (Q/lattice_models/Heisenberg.cu:88-176): every
    (ni, nj, a, b) with ni in [-N1, N1), nj in [-N2, N2), stable-sorted by bond
    length; shells of equal length (1e-3) take J_vec[0], J_vec[1], ...; zero
    length is skipped; both Jxy and Jz get the same coupling. Distances are
    computed in float32 like the reference so that shell boundaries agree."""
    A = geom.A_mat
    M = geom.M
    rows, dist = [], []
    f32 = np.float32
    for ni in range(-geom.N1, geom.N1):
        for nj in range(-geom.N2, geom.N2):
            for ai in range(M):
                for bi in range(M):
                    rows.append((ni, nj, ai, bi))
                    dx = f32(ni) * f32(A[0][0]) + f32(nj) * f32(A[0][1]) + f32(geom.rx_alpha[ai]) - f32(geom.rx_alpha[bi])
                    dy = f32(ni) * f32(A[1][0]) + f32(nj) * f32(A[1][1]) + f32(geom.ry_alpha[ai]) - f32(geom.ry_alpha[bi])
                    # pow(float, 2) promotes to double; sqrtf of the float-rounded sum
                    d2 = float(dx) * float(dx) + float(dy) * float(dy)
                    dist.append(float(np.sqrt(f32(d2))))
    order = sorted(range(len(rows)), key=lambda k: dist[k])  # Python sort is stable
    rows = [rows[k] for k in order]
    dist = [dist[k] for k in order]
    info = HeisenbergInfo(intra_c_bool=intra, inter_c_bool=inter)
    J_cur = J_vec[0]
    J_counter = -1
    d_cur = dist[0]
    thold = 0.001
    for k in range(len(rows)):
        if dist[k] < thold:
            continue
        if dist[k] > d_cur + thold:
            J_counter += 1
            if J_counter >= len(J_vec):
                break
            J_cur = J_vec[J_counter]
            d_cur = dist[k]
        info.Jxy_terms.append(list(rows[k]))
        info.Jz_terms.append(list(rows[k]))
        info.Jxy_couplings.append(complex(J_cur))
        info.Jz_couplings.append(float(J_cur))
    return info


def chain_xxz_info(n: int, J1=1.0, delta=1.0, J2=0.0) -> tuple:
    """Textbook periodic J1-J2 XXZ chain H = sum_r sum_{d=1,2} J_d (Sx Sx + Sy Sy + delta Sz Sz)
    written through the reference's Heisenberg constructor: Jxy rows at offsets
    +d and -d with J_d/2 each (so every bond gets Sp Sm + Sm Sp) and one Jz row at
    offset +d with J_d*delta. Geometry: N1 = n cells, N2 = 1, one site per cell."""
    geom = square_geometry(n, 1)
    info = HeisenbergInfo()
    for d, J in ((1, J1), (2, J2)):
        if J == 0.0:
            continue
        info.Jxy_terms += [[d, 0, 0, 0], [-d, 0, 0, 0]]
        info.Jxy_couplings += [J / 2.0, J / 2.0]
        info.Jz_terms.append([d, 0, 0, 0])
        info.Jz_couplings.append(J * delta)
    return geom, info


def square_xxz_info(n1: int, n2: int, J=1.0, delta=1.0) -> tuple:
    """Textbook periodic square-lattice nearest-neighbour XXZ model through the
    reference's Heisenberg constructor (Jxy rows at (+-1,0), (0,+-1) with J/2,
    Jz rows at (1,0), (0,1) with J*delta)."""
    geom = square_geometry(n1, n2)
    info = HeisenbergInfo()
    info.Jxy_terms = [[1, 0, 0, 0], [-1, 0, 0, 0], [0, 1, 0, 0], [0, -1, 0, 0]]
    info.Jxy_couplings = [J / 2.0] * 4
    info.Jz_terms = [[1, 0, 0, 0], [0, 1, 0, 0]]
    info.Jz_couplings = [J * delta] * 2
    return geom, info


def conjugated_term_table(terms):
    """BuildOperatorVecs (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336)
    with GetConjugateOperatorType (Q/quantum_operators/SingleParticleOperators.cu:7-29):
    op types conjugated (Sp <-> Sm), scalar = product of conj(op scalars), sites
    padded with -1, NOTerms per term. Returns numpy arrays
    (op_type[H_size*H_col], site[H_size*H_col], scalar[H_size] complex, n_ops[H_size], H_col)."""
    conj = {OP_SZ: OP_SZ, OP_SP: OP_SM, OP_SM: OP_SP}
    h_size = len(terms)
    h_col = max((len(t) for t in terms), default=0)
    op = np.zeros(h_size * h_col, dtype=np.int32)  # enum default value 0 (Sz) like a value-initialised host_vector
    site = np.full(h_size * h_col, -1, dtype=np.int32)
    scalar = np.ones(h_size, dtype=np.complex128)
    n_ops = np.zeros(h_size, dtype=np.int32)
    for hi, t in enumerate(terms):
        n_ops[hi] = len(t)
        for k, (s, ty, sc) in enumerate(t):
            op[hi * h_col + k] = conj[ty]
            site[hi * h_col + k] = s
            scalar[hi] = scalar[hi] * np.conj(sc)
    return op, site, scalar, n_ops, h_col


def max_terms_rule(n_terms: int, n_up: int, n_sites: int, hop_f: int = -1) -> int:
    """ELL width of the reference (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:71-72)."""
    if hop_f > 0:
        return n_up * hop_f
    return n_terms * n_up // n_sites + 1


def _tnm_general(r1, r2, nx, ny, geom: Geometry):
    """Tnm_general (Q/symmetries/SymmOperators.cu:62-126), in float64."""
    tol = 0.001
    a1 = np.array([geom.A_mat[0][0], geom.A_mat[1][0]])
    a2 = np.array([geom.A_mat[0][1], geom.A_mat[1][1]])
    w12 = float(a1 @ a2)
    winv = np.linalg.inv(np.array([[1.0, w12], [w12, 1.0]]))
    x = r1 + nx * a1[0] + ny * a2[0]
    y = r2 + nx * a1[1] + ny * a2[1]
    out1, out2 = np.empty_like(x), np.empty_like(y)
    for i in range(len(x)):
        v = np.array([a1 @ np.array([x[i], y[i]]), a2 @ np.array([x[i], y[i]])])
        al = winv @ v
        al0 = math.fmod(al[0], geom.N1)
        al1 = math.fmod(al[1], geom.N2)
        if abs(al0 - geom.N1) < tol:
            al0 = 0.0
        if abs(al1 - geom.N2) < tol:
            al1 = 0.0
        out1[i] = al0 * a1[0] + al1 * a2[0]
        out2[i] = al0 * a1[1] + al1 * a2[1]
    return out1, out2


def root_of_unity(num: int, den: int) -> complex:
    """exp(2 pi i num/den) with exact values on quarter turns."""
    num %= den
    if (4 * num) % den == 0:
        return [1 + 0j, 1j, -1 + 0j, -1j][(4 * num) // den]
    ang = 2.0 * math.pi * num / den
    return complex(math.cos(ang), math.sin(ang))


def translation_group(geom: Geometry):
    """Tnm_group (Q/symmetries/Groups.cu:49-113) + SymmetryClass::GetGroupElemArr
    (Q/symmetries/SymmetryClass.cu:26-71) + TranformGroupEls_ArmaToDev
    (Q/utils/misc_funcs.cu:80-90).

    Element e = jx*N2 + jy is the translation by jx*a1 + jy*a2 (Groups.cu:66-73);
    group_el[e][dest(s)] = s where dest(s) is the site that s moves to, i.e. the
    array applies the inverse element. Irrep a = nx*N2 + ny has character
    exp(+2 pi i (nx*jx/N1 + ny*jy/N2)) (Groups.cu:89-112), evaluated exactly.
    Returns (group_el[G, LS] uint32, chars[NIr, G] complex128)."""
    N1, N2 = geom.N1, geom.N2
    r1, r2 = site_coordinates(geom)
    LS = len(r1)
    G = N1 * N2
    group_el = np.zeros((G, LS), dtype=np.uint32)
    for e in range(G):
        jx, jy = e // N2, e % N2
        t1, t2 = _tnm_general(r1, r2, jx, jy, geom)
        for s in range(LS):
            q1 = np.nonzero(np.abs(r1 - t1[s]) < 0.001)[0]
            q2 = np.nonzero(np.abs(r2 - t2[s]) < 0.001)[0]
            q = sorted(set(q1.tolist()) & set(q2.tolist()))
            group_el[e, q[0]] = s
    chars = np.zeros((G, G), dtype=np.complex128)
    for a in range(G):
        nx, ny = a // N2, a % N2
        for e in range(G):
            jx, jy = e // N2, e % N2
            num = ((nx * jx) % N1) * N2 + ((ny * jy) % N2) * N1
            chars[a, e] = root_of_unity(num, N1 * N2)
    return group_el, chars


def with_spin_flip(group_el, chars):
    """Spin-inversion extension (no reference counterpart): element f*G + e is
    translation e followed by flipping every spin if f = 1; irrep p*NIr + a has
    character chi_a(e) * (-1)^(p*f)."""
    G, LS = group_el.shape
    nir = chars.shape[0]
    ge = np.concatenate([group_el, group_el], axis=0)
    flip = np.concatenate([np.zeros(G, np.uint8), np.ones(G, np.uint8)])
    ch = np.zeros((2 * nir, 2 * G), dtype=np.complex128)
    for p in range(2):
        for a in range(nir):
            for f in range(2):
                sgn = -1.0 if (p and f) else 1.0
                ch[p * nir + a, f * G:(f + 1) * G] = chars[a] * sgn
    return ge, flip, ch


def sector_basis(n_sites: int, n_up: int) -> np.ndarray:
    """Ascending popcount-n_up integers (closed form of BuildBasisStates)."""
    from itertools import combinations

    states = [sum(1 << p for p in c) for c in combinations(range(n_sites), n_up)]
    return np.array(sorted(states), dtype=np.uint64)


def dense_sector_hamiltonian(n_sites: int, n_up: int, terms) -> np.ndarray:
    """Dense full-Sz-sector Hamiltonian straight from the operator terms (NOT the
    conjugated table): each term c * O_1 O_2 ... O_n acts on a ket right to left
    (O_n first). Sz = +-1/2, Sp raises, Sm lowers. Independent of the reference's
    conjugate-and-apply-left-to-right device trick, so it pins that trick."""
    states = sector_basis(n_sites, n_up)
    index = {int(s): i for i, s in enumerate(states)}
    D = len(states)
    H = np.zeros((D, D), dtype=np.complex128)
    for t in terms:
        coef0 = 1.0 + 0j
        for (_, _, sc) in t:
            coef0 *= sc
        for col, s in enumerate(states):
            st = int(s)
            c = coef0
            ok = True
            for (site, ty, _) in reversed(t):
                bit = (st >> site) & 1
                if ty == OP_SZ:
                    c *= 0.5 if bit else -0.5
                elif ty == OP_SP:
                    if bit:
                        ok = False
                        break
                    st |= 1 << site
                elif ty == OP_SM:
                    if not bit:
                        ok = False
                        break
                    st &= ~(1 << site)
            if ok:
                H[index[st], col] += c
    return H


def ell_to_dense(n_rows: int, W: int, ind, val, track) -> np.ndarray:
    """Dense matrix of a reference-layout ELL (duplicate columns summed)."""
    E = np.zeros((n_rows, n_rows), dtype=np.complex128 if np.iscomplexobj(val) else np.float64)
    ind = ind.reshape(n_rows, W)
    val = val.reshape(n_rows, W)
    for r in range(n_rows):
        for i in range(int(track[r])):
            E[r, ind[r, i]] += val[r, i]
    return E
