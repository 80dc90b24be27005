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


# N = 10007: the candidate lengths M = 20160 (four-step) and 20480 (LDS)
# time within a few percent of each other, so either may win
@pytest.mark.parametrize("n,batch,kind", [(100000, 6, "fourstep(tuned)"), (65537, 3, "rader-fourstep(tuned)"),
                                          (10007, 8, "bluestein-"), (100003, 3, "bluestein-fourstep(tuned)")])
def test_tuned_structures_vs_oracle(fy, orc, cuda, n, batch, kind):
    """Tuned four-step splits, Rader convolution splits and Bluestein
    convolution lengths compute the same transform, and their choice is
    recorded for this GPU model."""
    import torch
    fy.wisdom_forget()
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    xd = torch.from_numpy(x).cuda()
    with fy.Plan(n, batch, fy.F32, fy.FORM_TUNED) as p:
        d = p.describe()
        assert d.startswith(kind) and "(tuned)" in d, d
        y = p.execute(xd, None, fy.FORWARD)
        z = p.execute(y, None, fy.INVERSE)
    torch.cuda.synchronize()
    tol = 1e-7 * math.log2(n) + 1e-7
    assert orc.rel_l2(y.cpu().numpy(), orc.fft_rows(x.astype(np.complex128))) < tol, d
    assert orc.rel_l2(z.cpu().numpy(), x) < 2 * tol, d
    w = fy.wisdom_export()
    dev = w.splitlines()[1]
    assert dev.startswith("device gfx9"), w
    entry = {"fourstep(tuned)": f"fs {n} 0 0 ", "rader-fourstep(tuned)": f"fs {n} 0 1 "}.get(kind, f"bluestein {n} 0 ")
    assert entry in w, (entry, w)
    # the same plan again comes from wisdom: same structure
    fy.wisdom_forget()
    fy.wisdom_import(w)
    with fy.Plan(n, batch, fy.F32, fy.FORM_TUNED) as p:
        assert p.describe() == d
    # wisdom of another GPU model is refused and changes nothing
    other = w.replace(dev, "device gfx000 1")
    with pytest.raises(fy.FyfftError):
        fy.wisdom_import(other)
    assert fy.wisdom_export() == w
    fy.wisdom_forget()


def test_grouped_tuned_vs_oracle(fy, orc, cuda):
    import torch
    lens = [16, 100, 1000, 4096, 6720, 12289, 1031, 40000, 65537, 100003, 4096, 1000]
    rng = np.random.default_rng(3)
    tot = sum(lens)
    x = (rng.standard_normal(tot) + 1j * rng.standard_normal(tot)).astype(np.complex64)
    fy.wisdom_forget()
    with fy.GroupedPlan(lens, fy.F32, fy.GROUPED_TUNED) as p:
        y = p.execute(torch.from_numpy(x).cuda(), None, fy.FORWARD).cpu().numpy()
    w = fy.wisdom_export()
    assert "ldsg " in w and "fs 40000 0 0 " in w, w
    fy.wisdom_forget()
    off = 0
    for n in lens:
        err = orc.rel_l2(y[off:off + n], orc.fft_ext(x[off:off + n].astype(np.complex128)))
        assert err < 1e-7 * math.log2(n) + 1e-7, (n, err)
        off += n
