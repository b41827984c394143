"""The allocation-free five-point wall fit (nodals_wall_profile5_noalloc)
against the allocating entry point it varies (nodals_wall_profile5), through
the C ABI: bitwise equal results on two calls in a row, the second with
fewer faces, so the staging array keeps the first call's data beyond the
second call's range."""
import numpy as np
import pytest

from tests.sst_cases import profile_inputs

pytestmark = pytest.mark.gpu


def test_noalloc_profile5_matches_allocating_entry_point():
    from nodals_amd import _lib

    U, Us, y, nu = profile_inputs()
    n = U.shape[0]
    rng = np.random.default_rng(7)
    order = rng.permutation(n)
    for label, idx in (("all faces", np.arange(n)), ("smaller, shuffled", order[: n // 3])):
        print(f"\n[gpu] profile5 noalloc vs allocating, {label}, n={idx.size}", flush=True)
        a, av = _lib.wall_profile5(U[idx], Us[idx], y[idx], nu)
        b, bv = _lib.wall_profile5_noalloc(U[idx], Us[idx], y[idx], nu)
        assert np.array_equal(av, bv), label
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), label
