"""rbmgpu_apply_S_noalloc (S.x with its block partials in the library's
static device scratch, no allocation) against rbmgpu_apply_S on the same
sample set, through the C ABI.  Both run the same kernels with the same
launch layout and reduction order, so the products must agree bit for bit.
Two calls in a row per case, the second on a smaller sample set (fewer
block partials than the first left in the scratch), then a sample set
whose partials exceed the scratch, which must be refused."""
import numpy as np
import pytest

import pyoracle as po
import rbmgpu
from helpers import ini, random_params

pytestmark = pytest.mark.gpu

CASES = [
    # (spec, precision, keep_o): structured FP32, materialised FP32, FP64, symmetric RBM
    (ini("toric", 4, alpha=2), 32, False),
    (ini("toric", 4, alpha=2), 32, True),
    (ini("kitaev", 3), 64, False),
    (ini("kitaev", 6, rbm="symmetry", alpha=4), 32, False),
]


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(np.float64)).to("cuda:0")


@pytest.mark.parametrize("spec,prec,keep_o", CASES)
def test_apply_S_noalloc_matches_apply_S(spec, prec, keep_o):
    import torch

    print(f"[gpu-test] noalloc S.x {spec.split()[3:6]} prec={prec} keep_o={keep_o}", flush=True)
    rng = np.random.default_rng(7)
    m = rbmgpu.Model(spec + ("[gpu]\nkeep_o = true\n" if keep_o else ""))
    g = m.gpu(precision=prec)
    g.set_params(random_params(rng, m.P, 0.1))
    for S in (6000, 700):  # second call: a smaller sample set
        g.load_samples(po.random_states(rng, S, m.N))
        g.reduce(want_o=False)
        x = rng.normal(size=m.P) + 1j * rng.normal(size=m.P)
        r1, r2, rd = 0.9, 9e-4, 1e-2
        y_ref = g.apply_S(x, r1, r2, rd)
        xd = _dev(x)
        yd = torch.full((2 * m.P,), float("nan"), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        g.apply_S_noalloc(xd.data_ptr(), yd.data_ptr(), r1, r2, rd)
        g.synchronize()
        y = yd.cpu().numpy().view(np.complex128)
        assert np.array_equal(y, y_ref), (S, np.max(np.abs(y - y_ref)))


def test_apply_S_noalloc_refuses_oversized_partials():
    """FP64 with O materialised at P = 2144 and 32768 samples: 1024 blocks of
    2144 complex doubles (35 MB) do not fit the 8 MiB scratch."""
    import torch

    rng = np.random.default_rng(8)
    m = rbmgpu.Model(ini("toric", 4, alpha=2))
    g = m.gpu(precision=64)
    g.set_params(random_params(rng, m.P, 0.1))
    g.load_samples(po.random_states(rng, 32768, m.N))
    g.reduce(want_o=False)
    xd = torch.zeros(2 * m.P, dtype=torch.float64, device="cuda:0")
    yd = torch.zeros(2 * m.P, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="static scratch"):
        g.apply_S_noalloc(xd.data_ptr(), yd.data_ptr(), 0.9, 9e-4, 1e-2)
