"""The device's libm (tm5_libm through pydyn_amd.libm: fdlibm exp, log and pow, the functions
every kernel of the library calls) against the oracle's C twin, bit for bit, over every binade,
the branches of each method and the special cases of C99 Annex F."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D
    from pydyn_amd import libm

    return D, libm


def same_bits(got, want):
    both_nan = np.isnan(got) & np.isnan(want)
    return np.all(both_nan | (got.view(np.int64) == want.view(np.int64)))


def test_exp_log_pow_match_the_host_twin(dev):
    from oracle import libm as host

    D, libm = dev
    rng = np.random.default_rng(21)
    special = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 0.5, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 1.7e308])
    x_exp = np.concatenate([rng.uniform(-750.0, 710.0, 200000), rng.uniform(-0.5, 0.5, 50000), special])
    x_log = np.concatenate([np.exp(rng.uniform(-744.0, 709.0, 200000)), rng.uniform(0.5, 2.0, 100000),
                            1.0 + rng.uniform(-1e-6, 1e-6, 20000), np.ldexp(1.0, np.arange(-1074, 1024)),
                            special])  # fmt: skip
    xs = np.concatenate([np.exp(rng.uniform(-30.0, 30.0, 100000)), rng.uniform(0.1, 3.0, 100000),
                         1.0 + rng.uniform(-1e-7, 1e-7, 2000), -np.exp(rng.uniform(-5.0, 5.0, 2000)),
                         np.repeat(special, special.size), np.full(2098, 2.0),
                         rng.uniform(0.0, 1e6, 2000)])  # fmt: skip
    ys = np.concatenate([rng.uniform(-20.0, 20.0, 100000), rng.uniform(-3.0, 3.0, 100000),
                         rng.uniform(-3e9, 3e9, 2000), np.round(rng.uniform(-20.0, 20.0, 2000)),
                         np.tile(special, special.size), np.arange(-1074.0, 1024.0),
                         np.full(2000, 0.5)])  # fmt: skip (y = 0.5: the correctly rounded sqrt)
    for name, args, want in (("exp", (x_exp,), host.exp(x_exp)), ("log", (x_log,), host.log(x_log)),
                             ("pow", (xs, ys), host.pow(xs, ys))):  # fmt: skip
        got = D.download(getattr(libm, name)(*(D.upload(a) for a in args)))
        assert same_bits(got, want), f"{name}: {np.sum(~same_bits_each(got, want))} differ"


def same_bits_each(got, want):
    return (np.isnan(got) & np.isnan(want)) | (got.view(np.int64) == want.view(np.int64))
