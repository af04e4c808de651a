import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
os.environ.setdefault("HJ_QUIET", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


def haar(n, rng):
    """Haar-distributed orthogonal matrix (QR of N(0,1), column signs from diag R)."""
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def svd_input(m, n, s, seed):
    """G = U diag(s) V^T with Haar U (m x n slice), Haar V."""
    rng = np.random.default_rng(seed)
    U = haar(m, rng)[:, :n]
    V = haar(n, rng)
    return (U * np.asarray(s)) @ V.T


def graded_input(m, n, seed, lo=-4.0, hi=4.0):
    """G_ij ~ N(0,1) times column scales 10^U(lo,hi)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((m, n)) * 10.0 ** rng.uniform(lo, hi, n)


@pytest.fixture(scope="session")
def strat0_cycwor():
    import gpujachx_amd as g
    return g.strat_generate("cycwor", 32).astype(np.uint8)


def assert_per_value(fast, plain, spread, ctx="", factor=10.0):
    """Per-value agreement of a fast form's spectrum `fast` with the plain
    FP64 run's `plain` (sigma, or D = j sigma^2): every sorted value within
    `factor` times the largest relative spread of another FP64 run `spread`
    of the same matrix (1e-12 at least), the values of at least a tenth of
    the largest to 1e-11.  A wrong small value fails here even where a
    normwise comparison cannot see it."""
    from gpujachx_amd.accuracy import rel_diffs
    rel_f, ref = rel_diffs(fast, plain)
    rel_s, _ = rel_diffs(spread, plain)
    tol = max(factor * rel_s.max(), 1e-12)
    assert rel_f.max() <= tol, "%s: per-value max %.3g > %.3g (FP64 spread %.3g)" % (
        ctx, rel_f.max(), tol, rel_s.max())
    big = ref >= 0.1 * ref.max()
    assert np.all(rel_f[big] <= 1e-11), "%s: large values %.3g" % (ctx, rel_f[big].max())
    return rel_f.max(), rel_s.max()
