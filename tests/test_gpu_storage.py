"""The 16-bit storage format of the rest populations.

With FP16/BF16 storage the moving populations are kept as shifted 16-bit
floats, f - w_l rho0, and the rest pair (l = 0) as a 16-bit float with three
exponent and twelve mantissa bits: |h| = 2^(e - 9) (1 + m / 4096) for
e = 1..7, m 2^-20 for e = 0, up to 2^-1 (1 - 2^-13), for the plain model's
exponent offset.  These tests pin that format through the C ABI: the rest
value a uniform start stores reads back within half its spacing (a quarter
of FP16's above 2^-8), and a start whose rest value lies beyond the range is
reported as LBM_ERR_STORE_RANGE by the next step.
"""
import numpy as np
import pytest

from lbcuda_amd.abi import default_config, INIT_RHO, STORE_F16, STORE_BF16
from lbcuda_amd.lib import Box

pytestmark = pytest.mark.gpu

W0 = np.float32(float.fromhex("0x1.555552p-2"))   # the library's rest weight
RHO0 = 0.5                                         # storage shift density


def spacing(h):
    """Spacing of the rest format at |h| (exponent offset 0)."""
    a = np.abs(h).astype(np.float64)
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -30)))
    return np.where(a < 2.0 ** -8, 2.0 ** -20, 2.0 ** (e - 12))


@pytest.mark.parametrize("storage", [STORE_F16, STORE_BF16])
def test_rest_pair_reads_back_within_half_spacing(storage):
    print(f"rest format storage={storage}", flush=True)
    n = 16
    rng = np.random.default_rng(11)
    # colour densities across the whole range the format holds (rho0 + 1.5)
    rr = rng.uniform(0.0, 1.99, (n, n, n)).astype(np.float32)
    rb = rng.uniform(0.0, 0.02, (n, n, n)).astype(np.float32)
    rr.flat[:4] = (0.5, 0.5 + 1e-4, 0.5 - 3e-3, 0.0)   # h = 0, subnormal range, f = 0
    g = Box(default_config(n, storage=storage, store_rho0=RHO0))
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    f0 = g.pops(0)[..., 0]
    g.close()
    want = (W0 * rr).astype(np.float32)          # f^eq_0 at rest
    h = want.astype(np.float64) - float(W0) * RHO0
    err = np.abs(f0.astype(np.float64) - want.astype(np.float64))
    bound = 0.5 * spacing(h) + 1e-7               # + float rounding of w_0 rho
    print(f"  max error {err.max():.2e}, max error / half spacing {(err / (0.5 * spacing(h))).max():.3f}")
    assert np.all(err <= bound)
    # above 2^-8 the spacing is a quarter of FP16's at the same magnitude
    big = np.abs(h) >= 2.0 ** -8
    fp16_half_ulp = 0.5 * 2.0 ** (np.floor(np.log2(np.abs(h[big]))) - 10)
    assert np.all(err[big] <= 0.25 * fp16_half_ulp + 1e-7)
    assert err.flat[0] < 1e-7                     # h = 0 is stored exactly


@pytest.mark.parametrize("rho_r,code", [(1.95, 0), (2.5, 12)])
def test_rest_value_beyond_range_is_reported(rho_r, code):
    """rho_r = 2.5 gives f_0 - w_0 rho0 = 0.75 > 0.5; 1.95 stays inside."""
    print(f"rest range rho_r={rho_r}", flush=True)
    n = 8
    g = Box(default_config(n, storage=STORE_F16, store_rho0=RHO0))
    rr = np.full((n, n, n), rho_r, np.float32)
    rb = np.zeros_like(rr)
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert g.step(1) == code
    g.close()
