"""Running variance of the render targets (UpdateIter,
output_handler_variance.hpp:7-31, called with the 0-based iteration index,
output_handler.hpp:138-139,223-225). The recursion
  var_i = var_{i-1} + (diff * (x_i - mean_i) - var_{i-1}) / max(1, i)
averages Welford's M2 increments over i steps, which makes it the unbiased
(n - 1 divisor) sample variance of the n = i + 1 per-iteration samples. The
test recovers those samples from the running means of runs of 1..n
iterations (every run restarts the same RNG sequence) and checks the oracle's
variance against numpy's var(ddof=1): a wrong divisor (n instead of n - 1)
is off by n / (n - 1) = 14 % at n = 8."""
import numpy as np

import pt_util as P
from mufflon_amd import scenes as S


def test_oracle_variance_is_sample_variance(oracle):
    c = S.cornell_materials(6, 12, 12, 12)
    bg = P.background_of(c, "envmap")
    ps = P.oracle_pt_scene_materials(oracle, c, bg)
    n = 8
    targets = [0, 1, 3, 6]
    means = []
    for k in range(1, n + 1):
        m, _ = oracle.pt_render_targets(ps, 1, 0, 6, 1, k, targets)
        means.append({t: m[t].astype(np.float64) for t in targets})
    mean, var, nans = oracle.pt_render_targets(ps, 1, 0, 6, 1, n, targets, variance=[0, 3])
    assert set(var) == {0, 3}
    for t in targets:
        assert np.array_equal(mean[t], means[-1][t].astype(np.float32))
    for t in (0, 3):
        x = [means[0][t]] + [(k + 1) * means[k][t] - k * means[k - 1][t] for k in range(1, n)]
        x = np.stack(x)
        expect = x.var(axis=0, ddof=1)
        scale = np.maximum(np.abs(x).max(axis=0), 1e-6) ** 2
        err = np.abs(var[t] - expect) / scale
        assert err.max() < 2e-5, (t, float(err.max()))
        assert (var[t] >= 0).all()
    # one iteration: variance 0 everywhere
    _, v1, _ = oracle.pt_render_targets(ps, 1, 0, 6, 1, 1, [0], variance=[0])
    assert not v1[0].any()
