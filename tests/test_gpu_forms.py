"""Plan forms: the plain generic kernels, the default (specialised where
available) and the autotuned choice compute the same transform; the
autotuner's choice is recorded as wisdom and reused."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64, 1000, 4096, 6720, 8192, 15625])
def test_forms_agree_with_oracle(fy, orc, cuda, n):
    import torch
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    ref = orc.fft_rows(x.astype(np.complex128))
    tol = 1e-7 * math.log2(n) + 1e-7
    descs = []
    for form in (fy.FORM_AUTO, fy.FORM_PLAIN, fy.FORM_TUNED):
        with fy.Plan(n, 5, fy.F32, form) as p:
            descs.append(p.describe())
            y = p.execute(xd, None, fy.FORWARD)
            z = p.execute(y, None, fy.INVERSE)
            torch.cuda.synchronize()
            assert orc.rel_l2(y.cpu().numpy(), ref) < tol, (form, p.describe())
            assert orc.rel_l2(z.cpu().numpy(), x) < 2 * tol
    assert "generic" in descs[1]
    assert "(tuned)" in descs[2]


def test_tuned_records_and_reuses_wisdom(fy, cuda):
    fy.wisdom_forget()
    with fy.Plan(2048, 64, fy.F32, fy.FORM_TUNED) as p:
        d1 = p.describe()
    w = fy.wisdom_export()
    assert "lds 2048 0 " in w
    fy.wisdom_forget()
    fy.wisdom_import(w)
    with fy.Plan(2048, 64, fy.F32, fy.FORM_TUNED) as p:
        assert p.describe() == d1
    fy.wisdom_forget()
