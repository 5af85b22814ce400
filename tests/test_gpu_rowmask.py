"""The occlusion kernel's light-row mask primitive (reciprocal fast path
with an exact-division fallback) against the oracle's IEEE-division
restatement, on random and adversarial (D, B) pairs (MI355X only)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(rng, R):
    n = 200_000
    D = (rng.normal(size=n) * rng.choice([1e-3, 1.0, 30.0, 1e4], size=n)).astype(np.float32)
    B = (rng.normal(size=n) * rng.choice([1e-3, 1.0, 30.0, 1e4], size=n)).astype(np.float32)
    # quotients within a few ulps of an integer: D = -k * B (+- ulps)
    k = rng.integers(0, R + 2, size=n // 2).astype(np.float32)
    Bq = (rng.normal(size=n // 2) * 3).astype(np.float32)
    Dq = (-k * Bq).astype(np.float32)
    Dq = np.nextafter(Dq, np.where(rng.random(n // 2) < 0.5, np.inf, -np.inf).astype(np.float32))
    Dq[::3] = (-k[::3] * Bq[::3]).astype(np.float32)
    # special values: zeros, denormals, huge, infinities, NaN
    sp = np.array([0.0, -0.0, 1e-40, -1e-40, 1e-30, 3e38, -3e38, np.inf, -np.inf, np.nan, 1.0, -1.0], np.float32)
    Ds, Bs = np.meshgrid(sp, sp)
    return np.concatenate([D, Dq, Ds.ravel()]), np.concatenate([B, Bq, Bs.ravel()])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("R", [1, 2, 7, 16, 32, 33, 64])
def test_device_rowmask_matches_oracle(fd, oracle, R, mode):
    rng = np.random.default_rng(R)
    D, B = _cases(rng, R)
    got = fd.selftest_rowmask(D, B, R, mode)
    want = np.array([oracle.rowmask(float(d), float(b), R - 1, R)[0] for d, b in zip(D, B)], np.uint64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(float(D[i]), float(B[i]), int(got[i]), int(want[i])) for i in bad[:5]]
