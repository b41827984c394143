"""The reference's portable SST-2003 and wall-law cores, compiled from the
reference tree into oracle/_ref, checked against closed forms of the
published models (so that the GPU tests' checker is itself sane)."""
import numpy as np
import pytest

from oracle import refbind
from tests.sst_cases import spalding_inputs, sst_inputs

pytestmark = pytest.mark.skipif(not refbind.sst_available(), reason="reference SST cores not built")


def test_spalding_root_satisfies_the_law_and_limits():
    R = refbind.RefSst()
    slip, y, nu = spalding_inputs()
    out, ok = R.spalding(slip, y, nu)
    m = ok.astype(bool) & (np.abs(slip) * y / nu > 1e-10)
    ut, yp = out[m, 0], out[m, 1]
    up = np.abs(slip[m]) / ut
    x = 0.4 * up
    law = up + np.exp(-0.4 * 5.5) * (np.expm1(x) - x - x * x / 2 - x ** 3 / 6)
    assert np.allclose(law, yp, rtol=1e-9)
    # viscous sublayer: u+ = y+ = sqrt(Re_y) for Re_y << 1
    tiny = (np.abs(slip[m]) * y[m] / nu) < 1e-3
    assert np.allclose(up[tiny], np.sqrt(np.abs(slip[m][tiny]) * y[m][tiny] / nu), rtol=1e-6)


def test_sst_local_closed_forms():
    R = refbind.RefSst()
    c = sst_inputs(200)
    out, fl = R.sst_local(c["gu"], c["k"], c["omega"], c["grad_k"], c["grad_omega"], c["nu"], c["radius"],
                          pipe_xy=c["xy"])
    nut, F1, F2 = out[:, 0], out[:, 1], out[:, 2]
    assert ((F1 >= 0) & (F1 <= 1) & (F2 >= 0) & (F2 <= 1)).all()
    k = np.maximum(c["k"], 1e-8)
    om = np.maximum(c["omega"], 1e-3)
    gu = c["gu"].reshape(-1, 3, 3)
    S = np.sqrt(np.maximum(0, 2 * (0.25 * (gu + gu.transpose(0, 2, 1)) ** 2).sum(axis=(1, 2))))
    assert np.allclose(nut, 0.31 * k / np.maximum(0.31 * om, S * F2), rtol=1e-12)
    assert np.allclose(out[:, 5], F1 * 0.075 + (1 - F1) * 0.0828, rtol=1e-14)
    assert (fl[20:30, 0] == 1).all() and (fl[30:40, 1] == 1).all()


@pytest.mark.skipif(not refbind.PORT_SST_LIB.exists(), reason="host build of the product SST functions missing")
def test_product_sst_functions_reproduce_reference_bitwise_on_host():
    """The pointwise functions the GPU kernels run, compiled for the host,
    give the reference cores' results bit for bit on the seeded sets."""
    from tests.sst_cases import omega_inputs, profile_inputs

    R, P = refbind.RefSst(), refbind.PortSst()
    c = sst_inputs()
    for dev in (False, True):
        a = R.sst_local(c["gu"], c["k"], c["omega"], c["grad_k"], c["grad_omega"], c["nu"], c["radius"],
                        dist=c["dist"], deviatoric=dev)
        b = P.sst_local(c["gu"], c["k"], c["omega"], c["grad_k"], c["grad_omega"], c["nu"], c["radius"],
                        dist=c["dist"], deviatoric=dev)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    slip, y, nu = spalding_inputs()
    a, b = R.spalding(slip, y, nu), P.spalding(slip, y, nu)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
    U, Us, yy, nu = profile_inputs()
    a, b = R.profile5(U, Us, yy, nu), P.profile5(U, Us, yy, nu)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
    inp, sh = omega_inputs()
    assert np.array_equal(R.of_auto(inp, sh), P.of_auto(inp, sh))
