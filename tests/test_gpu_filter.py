"""GPU cone filter: index structure bit-exact against the oracle's
brute-force radius query, values to rounding, all application modes."""
import numpy as np
import pytest
import torch

from oracle import filters as OFl
from oracle import mesh as OM
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,h,r", [((60, 20, 10), (1, 1, 1), 1.5), ((9, 7, 5), (1.0, 0.6, 1.4), 1.5),
                                      ((11, 9, 8), (1, 1, 1), 2.2), ((1, 1, 1), (1, 1, 1), 1.5)])
def test_filter_structure_and_values(gpu, dims, h, r):
    from pyto_amd import hex_mesher
    from pyto_amd.topopt_filters import createSmoothingFilter
    print("filter", dims, h, r, flush=True)
    mesh = hex_mesher.HexMesher().grid_mesh(dims, h)
    H, Hs = createSmoothingFilter(mesh, r)
    om = OM.GridMesh(dims, h)
    Href, Hs_ref = OFl.cone_filter(om, r)
    Hg = H.to_scipy()
    # index structure: identical sparsity pattern
    a = Hg.tocsr(); b = Href.tocsr()
    a.sort_indices(); b.sort_indices()
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.abs(a.data - b.data).max() <= 1e-15
    assert rel_err(Hs.cpu().numpy(), Hs_ref) <= 1e-14
    rng = np.random.default_rng(0)
    v = rng.uniform(0.1, 1.0, mesh.num_elems)
    x = rng.uniform(0.1, 1.0, mesh.num_elems)
    assert rel_err(H * v, Href @ v) <= 1e-14
    assert rel_err(H.density_filter(v), OFl.density_filter(Href, Hs_ref, v)) <= 1e-14
    assert rel_err(H.density_chain(v), OFl.density_filter_chain(Href, Hs_ref, v)) <= 1e-14
    assert rel_err(H.sensitivity_filter(x, v), OFl.sensitivity_filter(Href, Hs_ref, x, v)) <= 1e-14
    # device tensors stay on the device
    out = H * torch.from_numpy(v).to(gpu)
    assert out.is_cuda
