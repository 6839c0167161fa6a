"""Host-side model tables for the Python bindings: lattice Heisenberg XXZ term
tables and translation (+ spin-flip) group tables in the reference's
conventions. The C++ facade in include/qmb/ is the drop-in class API; this
module gives Python callers (tests, benchmark) the same tables.

Conventions (reference paths relative to its root, Q/ = QMBCuda/):
  site (i, j, alpha) -> (i*N2 + j)*M + alpha   (Q/lattice_models/ArbitrarySpinLattice.cu:19-36)
  term order: unit cell row-major, then Jxy rows (J * Sp_i Sm_j), then Jz rows
  (Jz * Sz_i Sz_j)                               (Q/lattice_models/Heisenberg.cu:22-79)
  table = conjugated op types, product of conjugated scalars
                                                 (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336)
  element e = jx*N2 + jy translates by (jx, jy); group_el[e][dest(s)] = s
                                                 (Q/symmetries/Groups.cu:49-113, SymmetryClass.cu:26-71)
  chi_a(e) = exp(+2 pi i (nx*jx/N1 + ny*jy/N2)), a = nx*N2 + ny   (Groups.cu:89-112)
Translations are computed with integer cell arithmetic (exact for the
Bravais lattices the reference's coordinate maps describe).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

OP_SZ, OP_SP, OP_SM = 0, 1, 2


@dataclass
class Lattice:
    N1: int
    N2: int = 1
    M: int = 1

    @property
    def n_sites(self) -> int:
        return self.N1 * self.N2 * self.M

    def site(self, i: int, j: int, a: int = 0) -> int:
        return (i * self.N2 + j) * self.M + a


@dataclass
class XXZInfo:
    """HeisenbergInfo (Q/lattice_models/Heisenberg.h:10-30): rows (di, dj, a, b)."""

    Jxy_terms: list = field(default_factory=list)
    Jxy_couplings: list = field(default_factory=list)
    Jz_terms: list = field(default_factory=list)
    Jz_couplings: list = field(default_factory=list)
    intra_c_bool: bool = True
    inter_c_bool: bool = True


def chain(n: int, J1: float = 1.0, delta: float = 1.0, J2: float = 0.0):
    """Periodic J1-J2 XXZ chain, textbook H = sum_d J_d sum_r (Sx Sx + Sy Sy + delta Sz Sz):
    Jxy rows at +-d with J_d/2, one Jz row at +d with J_d*delta."""
    info = XXZInfo()
    for d, J in ((1, J1), (2, J2)):
        if J == 0.0:
            continue
        info.Jxy_terms += [[d, 0, 0, 0], [-d, 0, 0, 0]]
        info.Jxy_couplings += [J / 2.0, J / 2.0]
        info.Jz_terms.append([d, 0, 0, 0])
        info.Jz_couplings.append(J * delta)
    return Lattice(n, 1, 1), info


def square(n1: int, n2: int, J: float = 1.0, delta: float = 1.0):
    """Periodic nearest-neighbour square-lattice XXZ model (textbook convention)."""
    info = XXZInfo(
        Jxy_terms=[[1, 0, 0, 0], [-1, 0, 0, 0], [0, 1, 0, 0], [0, -1, 0, 0]],
        Jxy_couplings=[J / 2.0] * 4,
        Jz_terms=[[1, 0, 0, 0], [0, 1, 0, 0]],
        Jz_couplings=[J * delta] * 2,
    )
    return Lattice(n1, n2, 1), info


def term_table(lat: Lattice, info: XXZInfo):
    """Conjugated flat term table of the Heisenberg Hamiltonian.
    Returns (op_type[H*2], site[H*2], scalar[H] complex, n_ops[H], max_ops=2)."""
    ops, sites, scal = [], [], []
    if info.intra_c_bool or info.inter_c_bool:
        rows = [info.Jxy_terms, info.Jz_terms]
        for uc in range(lat.N1 * lat.N2):
            j1, j2 = uc // lat.N2, uc % lat.N2
            for kind in range(2):
                for ji, r in enumerate(rows[kind]):
                    src = uc * lat.M + r[3]
                    i1, i2 = j1 + r[0], j2 + r[1]
                    if not info.intra_c_bool and (i1 < 0 or i1 >= lat.N1 or i2 < 0 or i2 >= lat.N2):
                        continue
                    if not info.inter_c_bool:
                        continue  # the reference's condition (Heisenberg.cu:57) is always true here
                    i1 %= lat.N1
                    i2 %= lat.N2
                    tgt = r[2] + lat.M * (lat.N2 * i1 + i2)
                    if kind == 0:
                        ops += [OP_SM, OP_SP]  # conj(Sp_tgt), conj(Sm_src)
                        sites += [tgt, src]
                        scal.append(np.conj(complex(info.Jxy_couplings[ji])))
                    else:
                        ops += [OP_SZ, OP_SZ]
                        sites += [tgt, src]
                        scal.append(complex(info.Jz_couplings[ji]))
    n = len(scal)
    return (np.array(ops, np.int32), np.array(sites, np.int32), np.array(scal, np.complex128),
            np.full(n, 2, np.int32), 2)


def _root_of_unity(num: int, den: int) -> complex:
    num %= den
    if (4 * num) % den == 0:
        return (1 + 0j, 1j, -1 + 0j, -1j)[(4 * num) // den]
    ang = 2.0 * math.pi * num / den
    return complex(math.cos(ang), math.sin(ang))


def translation_group(lat: Lattice, spin_flip: bool = False):
    """Translation group (optionally times spin inversion) in the reference's
    element/irrep order. With spin_flip, element f*G + e and irrep p*G + a
    (chi multiplied by (-1)^(p*f)). Returns (group_el[G', LS] uint32,
    flip[G'] uint8, chars[G', G'] complex128)."""
    N1, N2, M = lat.N1, lat.N2, lat.M
    G, LS = N1 * N2, lat.n_sites
    ge = np.zeros((G, LS), np.uint32)
    for e in range(G):
        jx, jy = e // N2, e % N2
        for i in range(N1):
            for j in range(N2):
                for a in range(M):
                    s = lat.site(i, j, a)
                    dest = lat.site((i + jx) % N1, (j + jy) % N2, a)
                    ge[e, dest] = s
    ch = np.zeros((G, G), np.complex128)
    for a in range(G):
        nx, ny = a // N2, a % N2
        for e in range(G):
            jx, jy = e // N2, e % N2
            ch[a, e] = _root_of_unity(((nx * jx) % N1) * N2 + ((ny * jy) % N2) * N1, N1 * N2)
    if not spin_flip:
        return ge, np.zeros(G, np.uint8), ch
    ge2 = np.concatenate([ge, ge])
    fl = np.concatenate([np.zeros(G, np.uint8), np.ones(G, np.uint8)])
    ch2 = np.zeros((2 * G, 2 * G), np.complex128)
    for p in range(2):
        for f in range(2):
            ch2[p * G:(p + 1) * G, f * G:(f + 1) * G] = ch * (-1.0 if (p and f) else 1.0)
    return ge2, fl, ch2


def reflection_group(lat: Lattice):
    """Site reflection of a chain, i -> -i (mod N): elements [identity, P], irreps
    [even, odd]. Same return form as translation_group."""
    n = lat.n_sites
    assert lat.N2 == 1 and lat.M == 1, "a chain"
    ge = np.zeros((2, n), np.uint32)
    for s in range(n):
        ge[0, s] = s
        ge[1, (-s) % n] = s
    ch = np.array([[1, 1], [1, -1]], np.complex128)
    return ge, np.zeros(2, np.uint8), ch


def rotation_group(lat: Lattice, order: int = 4):
    """Rotations of an L x L square torus about site (0, 0): element k turns by
    2 pi k/4 ((i, j) -> (-j, i) per quarter turn, mod L); order 2 keeps
    {identity, pi}. Irrep a has chi(k) = exp(2 pi i a k / order)."""
    L = lat.N1
    assert lat.N2 == L and lat.M == 1 and order in (2, 4), "an L x L square torus"
    n = lat.n_sites
    step = 4 // order
    ge = np.zeros((order, n), np.uint32)
    for k in range(order):
        for i in range(L):
            for j in range(L):
                x, y = i, j
                for _ in range(k * step):
                    x, y = (-y) % L, x
                ge[k, lat.site(x, y)] = lat.site(i, j)
    ch = np.array([[_root_of_unity(a * k, order) for k in range(order)] for a in range(order)], np.complex128)
    return ge, np.zeros(order, np.uint8), ch


def product_group(A, B):
    """Product set {a o b} (b first) of two groups on one lattice, each given as
    (group_el, flip, chars). Element ia*|B| + ib; characters chi_i(a) chi_j(b)
    for every pair (i, j) that is a one-dimensional representation of the product
    (checked on the composition table; a momentum that the point operation does
    not fix gives none), rows in (i, j) order. Returns (group_el, flip, chars, pairs)."""
    geA, flA, chA = (np.asarray(x) for x in A)
    geB, flB, chB = (np.asarray(x) for x in B)
    GA, GB = len(geA), len(geB)
    n = geA.shape[1]
    G = GA * GB
    ge = np.zeros((G, n), np.uint32)
    fl = np.zeros(G, np.uint8)
    for ia in range(GA):
        for ib in range(GB):
            ge[ia * GB + ib] = geB[ib][geA[ia]]
            fl[ia * GB + ib] = flA[ia] ^ flB[ib]
    index = {}
    for e in range(G):
        index.setdefault((ge[e].tobytes(), int(fl[e])), e)
    comp = np.zeros((G, G), np.int64)
    for x in range(G):
        for y in range(G):
            key = (ge[y][ge[x]].tobytes(), int(fl[x] ^ fl[y]))
            if key not in index:
                raise ValueError("the product is not closed under composition")
            comp[x, y] = index[key]
    rows, pairs = [], []
    for i in range(chA.shape[0]):
        for j in range(chB.shape[0]):
            chi = np.array([chA[i][e // GB] * chB[j][e % GB] for e in range(G)])
            if np.allclose(chi[comp], chi[:, None] * chi[None, :], atol=1e-12):
                rows.append(chi)
                pairs.append((i, j))
    return ge, fl, np.array(rows, np.complex128), pairs


def momentum_irrep(lat: Lattice, kx: int, ky: int = 0, flip_parity: int | None = None) -> int:
    """Irrep index of momentum (2 pi kx/N1, 2 pi ky/N2) and, with spin flip, parity (0 even, 1 odd)."""
    a = (kx % lat.N1) * lat.N2 + (ky % lat.N2)
    if flip_parity is not None:
        a += flip_parity * lat.N1 * lat.N2
    return a


def is_real_irrep(chars, irrep: int) -> bool:
    return bool(np.all(np.asarray(chars)[irrep].imag == 0.0))


def square_windows(lat: Lattice):
    """8-site windows of a periodic square lattice (one site per cell) for the
    learned start: a 2x4 and a 4x2 block anchored at every cell, bit k of a
    window = site k in row-major order inside the block. Returns [2*N, 8] uint8."""
    import numpy as np

    assert lat.M == 1
    wins = []
    for (h, w) in ((2, 4), (4, 2)):
        for i in range(lat.N1):
            for j in range(lat.N2):
                wins.append([lat.site((i + di) % lat.N1, (j + dj) % lat.N2) for di in range(h) for dj in range(w)])
    return np.array(wins, np.uint8)


def sublattice_mask(lat: Lattice) -> int:
    """Sites with (i + j) even: the A sublattice of the bipartite square lattice
    (the Marshall sign prior counts up spins there)."""
    assert lat.M == 1
    m = 0
    for i in range(lat.N1):
        for j in range(lat.N2):
            if (i + j) % 2 == 0:
                m |= 1 << lat.site(i, j)
    return m


# ---- operators for expectation values ----------------------------------------
# An operator is a list of products; a product is a list of (site, op, scalar)
# with op in {OP_SZ, OP_SP, OP_SM}, acting on a ket right to left (the last
# factor first), like ManyBodyOperator (Q/quantum_operators/ArbitraryOperator.h:7-80).


def operator_table(terms):
    """Conjugated flat table of an operator (BuildOperatorVecs,
    Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336): op
    types conjugated (Sp <-> Sm), scalar = product of the conjugated scalars.
    Returns (op_type, site, scalar, n_ops, max_ops) for qmbcuda_amd.Terms."""
    conj = {OP_SZ: OP_SZ, OP_SP: OP_SM, OP_SM: OP_SP}
    h = len(terms)
    w = max((len(t) for t in terms), default=0)
    op = np.zeros(h * w, np.int32)
    site = np.full(h * w, -1, np.int32)
    scalar = np.ones(h, np.complex128)
    n_ops = np.zeros(h, np.int32)
    for i, t in enumerate(terms):
        n_ops[i] = len(t)
        for k, (s, ty, sc) in enumerate(t):
            op[i * w + k] = conj[ty]
            site[i * w + k] = s
            scalar[i] *= np.conj(complex(sc))
    return op, site, scalar, n_ops, w


def szsz(i: int, j: int, scalar=1.0):
    """Sz_i Sz_j (SzSz_correlator, Q/quantum_operators/ArbitraryOperator.cu:234-242)."""
    return [[(i, OP_SZ, complex(scalar)), (j, OP_SZ, 1.0 + 0j)]]


def spin_spin(i: int, j: int, scalar=1.0):
    """S_i . S_j = (Sp_i Sm_j + Sm_i Sp_j)/2 + Sz_i Sz_j."""
    c = complex(scalar)
    return [[(i, OP_SP, c / 2), (j, OP_SM, 1.0 + 0j)], [(i, OP_SM, c / 2), (j, OP_SP, 1.0 + 0j)],
            [(i, OP_SZ, c), (j, OP_SZ, 1.0 + 0j)]]


def cell_positions(lat: Lattice):
    """Integer cell coordinates (i, j) of every site (sublattice offsets are not
    needed for the momenta of a Bravais torus)."""
    pos = np.zeros((lat.n_sites, 2), np.int64)
    for i in range(lat.N1):
        for j in range(lat.N2):
            for a in range(lat.M):
                pos[lat.site(i, j, a)] = (i, j)
    return pos


def structure_factor_zz(lat: Lattice, kx: int, ky: int = 0):
    """S^zz(k) = (1/N) sum_{r, r'} exp(i k.(r - r')) Sz_r Sz_r' at k = 2 pi (kx/N1, ky/N2),
    i.e. A_k^dagger A_k for the Fourier operator A_k = N^-1/2 sum_r exp(i k.r) Sz_r
    (FourierOperator, Q/quantum_operators/ArbitraryOperator.cu:248-285)."""
    pos = cell_positions(lat)
    n = lat.n_sites
    ph = np.exp(2j * math.pi * (pos[:, 0] * kx / lat.N1 + pos[:, 1] * ky / lat.N2))
    terms = []
    for r in range(n):
        for rp in range(n):
            terms.append([(r, OP_SZ, np.conj(ph[r]) * ph[rp] / n), (rp, OP_SZ, 1.0 + 0j)])
    return terms
