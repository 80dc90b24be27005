"""FP32 accuracy against the compiled reference's own error figures.

For every length whose FP32 forward error the survey measured on the
compiled reference (golden/reference_facts.json, fp32_fwd_rel_l2: one
transform of uniform[-1, 1) inputs from mt19937 seed 0, against an FP64
naive DFT), the GPU transform of the SAME input vector (tests/refinputs.py)
must stay within 1.25x the reference's figure.  The tolerance is written
here: rel-L2(GPU, FP64 DFT) <= 1.25 * reference rel-L2.
"""
import json
import os

import numpy as np
import pytest

from tests.refinputs import reference_inputs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FACTS = json.load(open(os.path.join(HERE, "golden", "reference_facts.json")))["fp32_fwd_rel_l2"]
RATIO = 1.25


@pytest.mark.parametrize("n", sorted(int(k) for k in FACTS))
def test_fp32_error_within_reference(fy, orc, cuda, n):
    import torch
    x = reference_inputs(n)
    ref = orc.dft_naive(x.astype(np.complex128))
    with fy.Plan(n, 1, fy.F32) as p:
        desc = p.describe()
        y = p.execute(torch.from_numpy(x[None, :]).cuda(), None, fy.FORWARD)
    torch.cuda.synchronize()
    err = orc.rel_l2(y.cpu().numpy()[0], ref)
    bar = FACTS[str(n)]
    print(f"N={n} gpu {err:.3e} reference {bar:.3e} ratio {err / bar:.3f} [{desc}]")
    assert err <= RATIO * bar, (n, err, bar, desc)
