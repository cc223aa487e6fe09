"""C ABI checks that need no GPU: the in-tree library loads, exports every
entry point include/acogpu.h declares, and rejects invalid parameters with an
error code and message before touching the device."""
import ctypes
import re

import pytest

from acogpu_amd import capi


def declared_functions():
    text = open(capi.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aco_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), f"{name} declared in acogpu.h but not exported"
    assert set(names) == set(capi.SIGNATURES), "ctypes signatures out of sync with the header"


def test_params_struct_layout_matches_header():
    # 24 4-byte fields before seed (8-byte aligned) + stream_id + construct_path + reserved[6]
    assert ctypes.sizeof(capi.Params) == 4 * 20 + 8 + 4 * 8
    assert capi.Params.seed.offset == 80


def test_reference_defaults():
    """aco_params_init reproduces inc/default.h:8-39 and RANDOM_SEED."""
    p = capi.default_params()
    assert (p.m, p.alpha, p.beta, p.rho, p.w) == (10, 1.0, 5.0, pytest.approx(0.1), 5)
    assert p.tau0 == pytest.approx(1e-6)
    assert (p.min_age, p.max_age, p.local_perf, p.global_perf, p.min_pop, p.max_pop) == (30, 50, 10, 10, 40, 100)
    assert p.seed == 12345678
    assert p.variant == capi.VARIANT_AS and p.mode == capi.MODE_CANDIDATE


@pytest.mark.parametrize("kw,frag", [
    (dict(n=1), "n must be"),
    (dict(m=0), "m must be"),
    (dict(k=0), "candidate list length"),
    (dict(k=65), "candidate list length"),
    (dict(n=10, k=10), "candidate list length"),
    (dict(rho=1.5), "rho"),
    (dict(alpha=-1.0), "alpha"),
    (dict(variant=7), "variant"),
    (dict(mode=5), "mode"),
    (dict(mode=1, n=9000), "full-row"),
    (dict(ants_per_block=9), "ants_per_block"),
    (dict(ants_per_block=-1), "ants_per_block"),
    (dict(construct_path=3), "construct_path"),
])
def test_invalid_parameters_are_rejected(kw, frag):
    base = dict(n=50, m=8, k=10)
    base.update(kw)
    p = capi.default_params(**base)
    with pytest.raises(capi.AcoError, match=frag):
        capi.Colony(p, xy=[[0, 0]] * p.n)


def test_exactly_one_input_required():
    L = capi.lib()
    p = capi.default_params(n=10, m=2, k=4)
    ctx = ctypes.c_void_p()
    assert L.aco_create(ctypes.byref(p), None, None, ctypes.byref(ctx)) == -1
    assert b"exactly one" in L.aco_last_error(None)
    assert L.aco_destroy(None) == 0
    assert L.aco_version().startswith(b"acogpu_amd")


def test_prior_config_layout_and_validation():
    # uint64 seed, 6 floats, int32 minibatch, int32 transform, int32 reserved[7] -> 68 bytes, 8-aligned
    assert ctypes.sizeof(capi.PriorConfig) == 72
    c = capi.default_prior_config()
    assert c.minibatch == 65536 and c.gamma == 3.0 and c.transform == 1 and c.pos_weight == 1.0
    L = capi.lib()
    out = ctypes.c_void_p()
    bad = capi.default_prior_config(lr=0.0)
    assert L.aco_prior_create(ctypes.byref(bad), ctypes.byref(out)) == -1  # ACO_E_ARG
    assert b"invalid prior configuration" in L.aco_prior_last_error(None)
    assert not out.value
