"""Calc_Kzz on the device (pydyn_amd.gpu.diffusion.calc_kzz, tm5_calc_kzz) against the
reference's own calc_kzz on its synthetic inputs (tests/golden/kzz.npz, written with the fdlibm
twin), bit for bit: dkg and the intermediates ustar, gph, pblh, kvh, at two shapes."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_oracle import kzz_case  # noqa: E402


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D
    from pydyn_amd.gpu import diffusion as KD

    return D, KD


@pytest.mark.parametrize("case", ["a", "b"])
def test_calc_kzz_matches_the_reference(dev, case):
    D, KD = dev
    args, grid, c = kzz_case(case)
    parts = {}
    dkg = KD.calc_kzz(*args, **grid, parts=parts)
    got = {"dkg": D.download(dkg), **{k: D.download(v) for k, v in parts.items()}}
    for k, v in got.items():
        w = c[k]
        differ = v.view(np.int64) != w.view(np.int64)
        assert not differ.any(), f"{k}: {differ.sum()} of {differ.size} differ, first {np.argwhere(differ)[0]}"
