"""The CH4 loss-rate construction on the device (pydyn_amd.gpu.kernels.loss_rates, tm5_loss_rates)
against the reference's own loss_rates and decay (tests/golden/chem_loss.npz, written by
pydyn/chem.py with the fdlibm twin), bit for bit: the zonal columns alone, with the
tropospheric Cl and the soil sink, and with the model's floor above the zonal columns."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_oracle import _chem_case, _loss_inputs  # noqa: E402


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D
    from pydyn_amd.gpu import kernels as K

    D.context().flags(clear=True)
    return D, K


@pytest.mark.parametrize("case", ["plain", "trop", "high"])
def test_loss_rates_match_the_reference(dev, case):
    D, K = dev
    c = _chem_case(case)
    kw = _loss_inputs(c)
    loss, decay = K.loss_rates(1800.0, **kw)
    assert np.array_equal(D.download(loss).view(np.int64), c["loss"].view(np.int64))
    assert np.array_equal(D.download(decay).view(np.int64), c["decay"].view(np.int64))


def test_loss_rates_reject_unordered_layers(dev):
    D, K = dev
    kw = _loss_inputs(_chem_case("plain"))
    kw["phlb"] = kw["phlb"].copy()
    kw["phlb"][2, 1, 5] = kw["phlb"][2, 1, 4] + 1.0  # one column's layer upside down
    with pytest.raises(ValueError, match="not strictly descending"):
        K.loss_rates(1800.0, **kw)
    kw = _loss_inputs(_chem_case("plain"))
    kw["strat_press"] = kw["strat_press"][:, ::-1].copy()
    with pytest.raises(ValueError, match="not strictly descending"):
        K.loss_rates(1800.0, **kw)
