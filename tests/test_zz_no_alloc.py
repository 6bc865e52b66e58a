"""cpg_l2norm_noalloc (MeshFn::L2Norm through a fixed device array, no allocation) against
cpg_l2norm on the same equations and states, through the C ABI: two calls in a row, the second
on a smaller equation, so the array is reused with less of it meaningful."""
import numpy as np
import pytest

from cpolyg_amd import hip
from cpolyg_amd.spec import EquationSpec, bs_of
from problems import vortex_mesh, vortex_state

pytestmark = pytest.mark.gpu


def test_l2norm_noalloc_matches_l2norm(cuda):
    import torch

    cases = [(vortex_mesh(32), 3), (vortex_mesh(6), 2)]   # 1,024 cells (p = 3), then 36 cells (p = 2)
    eqs = []
    for m, deg in cases:
        E = hip.Equation(hip.Mesh(m), EquationSpec.euler(periodic_trace=1, jacobian=1), deg)
        u = torch.from_numpy(vortex_state(m, deg, perturb=0.05, seed=deg)).to(cuda)
        eqs.append((E, u))
    for E, u in eqs:
        want = E.l2norm(u)
        got = E.l2norm_noalloc(u)
        assert np.all(np.isfinite(got))
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    # a second round of calls: the device's free memory does not change across them
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for E, u in eqs:
        E.l2norm_noalloc(u)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
