"""pp_repr_inner_product_noalloc (the S-metric inner product with every intermediate in a
fixed device array) against pp_repr_inner_product through the C ABI: bitwise equal on two
calls in a row, the second on a smaller wavefunction of another representation (the array
keeps the first call's contents), and a wavefunction above the fixed maximum is rejected
with an error before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests.conftest import crandn

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inner(lib, fn, repr_, a, b, work=None):
    from pyprop_amd import _native as N, device as D
    out = N.PPCplx()
    if work is None:
        rc = fn(repr_.h, D.ptr(a), D.ptr(b), C.byref(out), D.stream_handle())
    else:
        rc = fn(repr_.h, D.ptr(a), D.ptr(b), D.ptr(work), C.byref(out), D.stream_handle())
    return rc, complex(out.re, out.im)


def test_noalloc_inner_product_equals_allocating_one(rng):
    import torch
    assert torch.cuda.is_available()
    from pyprop_amd import _native as N
    from pyprop_amd.propagator import CombinedRepresentation
    from pyprop_amd.systems import helium_system, hydrogen_system
    lib = N.load()
    cases = [("helium 42 x 60 x 60", helium_system()),
             ("hydrogen 3 x 40", hydrogen_system(lmax=2, order=6, n_breakpoints=39, rmax=30.0, n_inner=10,
                                                 gamma=3.0))]
    for name, sysm in cases:
        say("noalloc inner product, %s (N = %d)" % (name, sysm.size))
        repr_ = CombinedRepresentation(sysm.shape, sysm.bases)
        a = torch.from_numpy(crandn(rng, sysm.shape)).to("cuda")
        b = torch.from_numpy(crandn(rng, sysm.shape)).to("cuda")
        work = torch.zeros_like(b)
        rc1, ref = _inner(lib, lib.pp_repr_inner_product, repr_, a, b, work)
        rc2, got = _inner(lib, lib.pp_repr_inner_product_noalloc, repr_, a, b)
        assert rc1 == 0 and rc2 == 0, N.load().pp_last_error()
        say("  allocating %r, no-alloc %r" % (ref, got))
        assert got == ref
        # the inputs are untouched
        assert torch.equal(b, torch.from_numpy(b.cpu().numpy()).to("cuda"))
        # and the value is <a|S b>: against the host with the band overlap
        from tests.test_oracle import oracle_system
        s = oracle_system(sysm)
        host = s.inner_product(a.cpu().numpy(), b.cpu().numpy())
        assert abs(got - host) < 1e-12 * abs(host)


def test_noalloc_rejects_oversize():
    import torch
    from pyprop_amd import _native as N
    from pyprop_amd.propagator import CombinedRepresentation
    lib = N.load()
    n = (1 << 20) + 1
    repr_ = CombinedRepresentation((1, n), [None, None])
    a = torch.zeros(n, dtype=torch.complex128, device="cuda")
    rc, _ = _inner(lib, lib.pp_repr_inner_product_noalloc, repr_, a, a)
    msg = lib.pp_last_error().decode()
    say("  oversize: rc %d, %s" % (rc, msg))
    assert rc == -1 and "exceed" in msg
