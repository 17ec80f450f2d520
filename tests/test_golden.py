"""The oracle reproduces the committed golden fixtures (tests/golden/)."""
import os

import numpy as np
import scipy.sparse.linalg as spla

from oracle import fea as OF
from oracle import filters as OFl
from oracle import ke as OK
from oracle import mesh as OM
from oracle import sensitivities as OS

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_tiny_fixture():
    d = np.load(os.path.join(HERE, "tiny.npz"))
    g = OM.GridMesh((4, 3, 2), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    assert np.allclose(KE, d["KE"], rtol=0, atol=1e-15)
    K = OF.assemble_K(g, KE, OF.simp_scaling(d["x"])).toarray()
    assert np.allclose(K, d["K"], rtol=0, atol=1e-14)


def test_c1_fixture():
    d = np.load(os.path.join(HERE, "c1.npz"))
    g = OM.GridMesh((60, 20, 10), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = OK.hex8_stiffness(1.0, 0.3, (1, 1, 1))
    x = np.full(g.num_elems, 0.5)
    C, grad = OS.compliance_and_gradient(d["u"], x, g.edofMatStructural, KE)
    assert np.isclose(C, d["C"], rtol=1e-12)
    assert np.allclose(grad, d["grad"], rtol=1e-10, atol=0)
    H, Hs = OFl.cone_filter(g, 1.5)
    H = H.tocsr()
    H.sort_indices()
    assert np.array_equal(H.indptr, d["H_indptr"]) and np.array_equal(H.indices, d["H_indices"])
    assert np.allclose(Hs, d["Hs"], rtol=1e-14)
    post = OF.postprocess(g, d["u"], x, 1.0, 0.3)
    assert np.isclose(post["pNormStress"], d["pnorm"], rtol=1e-12)
