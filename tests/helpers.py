"""Shared builders for the tests: model cases in the oracle's restatement and in
the library's own host tables, plus independent brute-force checks."""
from __future__ import annotations

import numpy as np

import oracle
from oracle import model as OM


def chain_case(n, J1=1.0, delta=1.0, J2=0.0, flip=False):
    """Oracle-side tables of a periodic J1-J2 XXZ chain (textbook convention)."""
    geom, info = OM.chain_xxz_info(n, J1, delta, J2)
    terms = OM.heisenberg_operator(geom, info)
    table = OM.conjugated_term_table(terms)
    ge, ch = OM.translation_group(geom)
    fl = None
    if flip:
        ge, fl, ch = OM.with_spin_flip(ge, ch)
    return dict(n_sites=n, terms=terms, table=table, group_el=ge, chars=ch, flip=fl)


def square_case(n1, n2, J=1.0, delta=1.0, flip=False):
    geom, info = OM.square_xxz_info(n1, n2, J, delta)
    terms = OM.heisenberg_operator(geom, info)
    table = OM.conjugated_term_table(terms)
    ge, ch = OM.translation_group(geom)
    fl = None
    if flip:
        ge, fl, ch = OM.with_spin_flip(ge, ch)
    return dict(n_sites=n1 * n2, terms=terms, table=table, group_el=ge, chars=ch, flip=fl)


def is_real_row(chars, a):
    return bool(np.all(np.asarray(chars)[a].imag == 0.0))


def oracle_orbits(case, n_up):
    return oracle.Orbits(case["n_sites"], n_up, case["group_el"], case["chars"], flip=case["flip"])


def oracle_ell(case, orb, alpha, max_terms=0, is_complex=None):
    if is_complex is None:
        is_complex = not (is_real_row(case["chars"], alpha) and np.all(case["table"][2].imag == 0))
    W = max_terms or OM.max_terms_rule(len(case["table"][3]), orb.n_up, orb.n_sites)
    return oracle.Ell(orb, case["table"], alpha, W, is_complex)


def brute_orbits(n_sites, n_up, group_el, flip=None):
    """Independent orbit enumeration by explicit images (small N only).
    Returns {rep: (stabiliser element list, orbit set)}."""
    states = [int(s) for s in OM.sector_basis(n_sites, n_up)]
    mask = (1 << n_sites) - 1
    G = group_el.shape[0]

    def image(s, g):
        out = 0
        for site in range(n_sites):
            if (s >> site) & 1:
                out |= 1 << int(group_el[g][site])
        if flip is not None and flip[g]:
            out ^= mask
        return out

    seen = set()
    orbits = {}
    for s in states:
        if s in seen:
            continue
        imgs = [image(s, g) for g in range(G)]
        orb = set(imgs)
        seen |= orb
        rep = min(orb)
        stab = [g for g in range(G) if image(rep, g) == rep]
        orbits[rep] = (stab, orb)
    return orbits


def dense_spectrum(case, n_up):
    H = OM.dense_sector_hamiltonian(case["n_sites"], n_up, case["terms"])
    assert np.allclose(H, H.conj().T, atol=1e-13)
    return np.linalg.eigvalsh(H)


def ell_dense_nonnull(ell, orb, alpha):
    E = ell.dense()
    keep = ~orb.null_mask(alpha)
    return E[np.ix_(keep, keep)]


def embedding(case, orb, alpha):
    """Full-sector vectors of the sector's rows: column r = sum_g conj(chi(g)) |g.rep_r>,
    normalised (zero for null orbits). An eigenvector x of the sector matrix is the
    many-body state embedding @ conj(x); tests check that on H itself."""
    n, G = case["n_sites"], orb.G
    states = OM.sector_basis(n, orb.n_up)
    index = {int(s): i for i, s in enumerate(states)}
    chi = np.asarray(case["chars"])[alpha]
    null = orb.null_mask(alpha)
    E = np.zeros((len(states), orb.n_orb), np.complex128)
    for r, rep in enumerate(orb.reps):
        if null[r]:
            continue
        for g in range(G):
            fl = 0 if case["flip"] is None else int(case["flip"][g])
            E[index[oracle.apply_elem(int(rep), case["group_el"][g], fl, n)], r] += np.conj(chi[g])
        E[:, r] /= np.linalg.norm(E[:, r])
    return E


def sector_ground_state(case, orb, alpha):
    """Lowest eigenpair of the oracle's sector matrix (dense), as a reference-layout
    vector over all orbit rows (zeros on null rows)."""
    ell = oracle_ell(case, orb, alpha, is_complex=True)
    keep = ~orb.null_mask(alpha)
    M = ell.dense()[np.ix_(keep, keep)]
    w, V = np.linalg.eigh(M)
    x = np.zeros(orb.n_orb, np.complex128)
    x[keep] = V[:, 0]
    return w[0], x


def burnside_count(case, n_up):
    """Number of orbits of the Sz sector under the group by Burnside's lemma,
    (1/G) sum_g |Fix(g)|, with the fixed states of each element counted from its
    cycles: without flip a fixed state is constant on every cycle (all down or
    all up); with flip (image = complement of the permuted state) it alternates
    along every cycle, so every cycle must be even and gives two states with
    half of its sites up."""
    ge, fl = case["group_el"], case["flip"]
    n = case["n_sites"]
    total = 0
    for g in range(ge.shape[0]):
        perm = [int(v) for v in ge[g]]
        seen = [False] * n
        poly = {0: 1}  # number of fixed states by up-spin count
        flip = fl is not None and fl[g]
        for s in range(n):
            if seen[s]:
                continue
            c, t = 0, s
            while not seen[t]:
                seen[t] = True
                t = perm[t]
                c += 1
            if not flip:
                opts = ((0, 1), (c, 1))
            elif c % 2 == 0:
                opts = ((c // 2, 2),)
            else:
                opts = ()
            nxt = {}
            for u, m in poly.items():
                for du, w in opts:
                    nxt[u + du] = nxt.get(u + du, 0) + m * w
            poly = nxt
        total += poly.get(n_up, 0)
    assert total % ge.shape[0] == 0
    return total // ge.shape[0]


class DeviceBuffer:
    """Device memory for tests that hand device pointers to the C ABI, taken from
    the HIP runtime libqmb itself links (libamdhip64.so.7). torch ships a HIP
    runtime of its own, so its tensors cannot be shared with libqmb in one process."""

    _hip = None

    @classmethod
    def hip(cls):
        import ctypes as C

        if cls._hip is None:
            import qmbcuda_amd as q

            q.lib()  # loads the runtime as libqmb's dependency
            try:
                h = C.CDLL("libamdhip64.so.7")
            except OSError:
                h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMalloc.restype = C.c_int
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemcpy.restype = C.c_int
            h.hipFree.argtypes = [C.c_void_p]
            h.hipFree.restype = C.c_int
            h.hipMemGetAddressRange.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_void_p]
            h.hipMemGetAddressRange.restype = C.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        import ctypes as C

        self.nbytes = int(nbytes)
        p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        assert cls.hip().hipMemcpy(buf.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return buf

    def reported_size(self):
        """Size of the allocation the runtime reports for this buffer (what the
        library's range checks see)."""
        import ctypes as C

        base, size = C.c_void_p(), C.c_size_t()
        assert self.hip().hipMemGetAddressRange(C.byref(base), C.byref(size), self.ptr) == 0
        return base.value, size.value

    def to_array(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
