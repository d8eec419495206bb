"""pgrav_integrate_device_noalloc (no device allocation: its scratch is a static
device array of the library) against pgrav_integrate_device, through the C ABI
on device tables: two calls in a row, the second smaller, each bit for bit
(the same kernels and summation order); a size beyond its scratch is refused
before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _table(n, seed):
    from pgrav.integrate import BODY
    rng = np.random.default_rng(seed)
    t = np.zeros(n, BODY)
    t["mass"] = rng.random(n) + 0.1
    t["pos"] = rng.normal(size=(n, 3))
    t["vel"] = rng.normal(size=(n, 3))
    t["acc"] = rng.normal(size=(n, 3)) * 10
    t["phi"] = -rng.random(n)
    t["ident"] = np.arange(n)
    return t


def test_noalloc_matches_allocating_entry_point(gpu_lib):
    import torch
    from pgrav import integrate
    from pgrav.integrate import BODY
    for n, seed in ((70001, 1), (3000, 2)):  # second call smaller than the first
        print("integrate noalloc n=%d" % n, flush=True)
        a = _table(n, seed)
        d_ref = torch.from_numpy(a.view(np.uint8).copy()).cuda()
        d_new = torch.from_numpy(a.view(np.uint8).copy()).cuda()
        r_ref = integrate.integrate_device(d_ref.data_ptr(), n, BODY, 0.013, 0.027, 5.0, 4.987)
        r_new = integrate.integrate_device(d_new.data_ptr(), n, BODY, 0.013, 0.027, 5.0, 4.987,
                                           noalloc=True)
        assert r_new == r_ref, (r_new, r_ref)
        np.testing.assert_array_equal(d_new.cpu().numpy(), d_ref.cpu().numpy())


def test_noalloc_refuses_oversize(gpu_lib):
    import torch
    from pgrav import PgravError, integrate
    from pgrav.integrate import BODY
    a = _table(16, 3)
    d = torch.from_numpy(a.view(np.uint8).copy()).cuda()
    with pytest.raises(PgravError, match="134217728"):
        integrate.integrate_device(d.data_ptr(), (1 << 27) + 1, BODY, 0.1, 0.1, 0.0, 0.0,
                                   noalloc=True)
    np.testing.assert_array_equal(d.cpu().numpy(), a.view(np.uint8))  # nothing ran
