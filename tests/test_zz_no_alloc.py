"""qmb_expectation_noalloc (the entry point that allocates nothing) against
qmb_expectation on the same inputs, through the C ABI: two calls in a row, the
second with a smaller basis and term table, and the scratch limit."""
import ctypes as C
import sys

import numpy as np
import pytest

import qmbcuda_amd as q
from qmbcuda_amd import models as LM
from tests.helpers import DeviceBuffer

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc(ctx, basis, terms, x_dev, cplx, flags=0):
    fn = q.lib().qmb_expectation_noalloc
    out = np.zeros(2, np.float64)
    avg = C.c_int(-1)
    st = fn(ctx.h, basis.h, terms.h, int(cplx), x_dev, flags, out.ctypes.data, C.byref(avg))
    return st, complex(out[0], out[1]), avg.value


def sector(ctx, n, k, delta=1.0, j2=0.0):
    lat, info = LM.chain(n, 1.0, delta, j2)
    ge, fl, ch = LM.translation_group(lat)
    g = q.Group(ge, ch)
    t = q.Terms(n, LM.term_table(lat, info))
    b = q.Basis(ctx, g, n // 2, k, q.LAYOUT_COMPACT)
    return lat, g, t, b, q.Ell(ctx, b, t)


def test_noalloc_expectation_equals_allocating_entry(ctx):
    cases = [("N=20 chain, k=0", 20, 0, 0.9, 0.3), ("N=14 chain, k=3 (complex)", 14, 3, 1.0, 0.0)]
    for label, n, k, d, j2 in cases:  # the second call is the smaller one
        say(f"noalloc expectation: {label}")
        lat, g, t, b, e = sector(ctx, n, k, d, j2)
        ev, rep, _, x = q.lanczos(e, max_iter=500, tol=1e-13, want_eigvec=True)
        cplx = np.iscomplexobj(x)
        ops = [t, q.Terms(n, LM.operator_table(LM.szsz(0, 3))),
               q.Terms(n, LM.operator_table(LM.structure_factor_zz(lat, n // 2)))]
        x_dev = DeviceBuffer.from_array(x.view(np.float64) if cplx else x)
        for op in ops:
            ref, ref_avg = b.expectation(op, x)
            st, got, avg = noalloc(ctx, b, op, x_dev.ptr, cplx)
            assert st == 0, q.lib().qmb_last_error()
            # same kernel, same block partials, summed in the same order: equal bit for bit
            assert got == ref, (got, ref)
            assert avg == int(ref_avg)
        # forced group average
        ref, _ = b.expectation(ops[1], x, group_average=True)
        st, got, avg = noalloc(ctx, b, ops[1], x_dev.ptr, cplx, flags=1)
        assert st == 0 and got == ref and avg == 1
        x_dev.free()


def test_noalloc_expectation_limits(ctx):
    say("noalloc expectation: limits")
    lat, g, t, b, e = sector(ctx, 12, 0)
    x = np.random.default_rng(2).normal(size=b.n_rows)
    x_dev = DeviceBuffer.from_array(x)
    # a term table beyond the scratch limit is refused with a clear error
    big = LM.operator_table([term for _ in range(16000) for term in LM.szsz(0, 1)])
    st, _, _ = noalloc(ctx, b, q.Terms(12, big), x_dev.ptr, False)
    assert st == 1 and b"scratch limit" in q.lib().qmb_last_error()
    # a host vector is refused before any kernel runs
    st, _, _ = noalloc(ctx, b, t, x.ctypes.data, False)
    assert st == 1 and b"not device memory" in q.lib().qmb_last_error()
    # a device vector shorter than the basis is refused
    short = DeviceBuffer(8 * (b.n_rows // 2))
    if short.reported_size() == (short.ptr, short.nbytes):  # the runtime reports the exact extent
        st, _, _ = noalloc(ctx, b, t, short.ptr, False)
        assert st == 1 and b"smaller than" in q.lib().qmb_last_error()
    # and the entry still works after the refusals
    ref, _ = b.expectation(t, x)
    st, got, _ = noalloc(ctx, b, t, x_dev.ptr, False)
    assert st == 0 and got == ref
    x_dev.free()
    short.free()
