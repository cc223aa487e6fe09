"""The allocation-free prior apply (aco_prior_apply_noalloc) against
aco_prior_apply through the C ABI: two calls in a row, the second on a
smaller colony, each heuristic bit-identical to the allocating entry point's
on an identical colony; a colony beyond the scratch limit is refused."""
import numpy as np
import pytest

from acogpu_amd import capi, instances

pytestmark = pytest.mark.gpu


def test_prior_apply_noalloc_equals_apply_on_two_calls():
    rng = np.random.default_rng(41)
    params = (rng.standard_normal(capi.PRIOR_PARAMS) * 0.2).astype(np.float32)
    with capi.Prior(capi.default_prior_config()) as pr:
        pr.set_params(params)
        for n, k, seed in ((1000, 20, 3), (300, 16, 4)):  # second call: smaller input
            print(f"case: prior apply no-alloc vs apply, n={n} k={k}", flush=True)
            xy = instances.uniform(n, 1000, seed=seed)
            p = capi.default_params(n=n, m=32, k=k, beta=5.0, rho=0.5, seed=seed)
            with capi.Colony(p, xy=xy) as a, capi.Colony(p, xy=xy) as b:
                a.initialize()
                b.initialize()
                pr.apply(a)
                pr.apply_noalloc(b)
                ha, hb = a.heuristic_nn(), b.heuristic_nn()
                assert ha.shape == (n, k)
                np.testing.assert_array_equal(ha.view(np.uint32), hb.view(np.uint32))
                assert np.all(hb > 0)


def test_prior_apply_noalloc_refuses_beyond_limit():
    print("case: prior apply no-alloc, n=13200 k=20 (264000 edges > 262144)", flush=True)
    xy = instances.uniform(13200, 100000, seed=5)
    p = capi.default_params(n=13200, m=8, k=20, beta=2.0, rho=0.5, seed=5)
    with capi.Prior() as pr, capi.Colony(p, xy=xy) as c:
        c.initialize()
        with pytest.raises(capi.AcoError, match="allocation-free limit"):
            pr.apply_noalloc(c)
