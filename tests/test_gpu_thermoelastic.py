"""GPU parity of the thermo-elastic coupling through the reference-shaped API
(HexThermalFEA.get_thermoelastic_force, HexStructuralFEA with
thermo_elastic_force, compute_thermoelastic_compliance_and_gradient) against
the oracle's restatement with sparse direct solves."""
import numpy as np
import pytest

from tests.helpers import rel_err
from tests.test_oracle_thermoelastic import Coupled

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout,exact", [("reference", False), ("physical", True)])
def test_thermoelastic_chain_matches_oracle(gpu, layout, exact):
    from pyto_amd import bound_cond, hex_mesher, linear_solvers, mat_lib, topopt_common
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd import topopt_obj_cons_sensitivities as S
    from pyto_amd.hex_structural_fea import HexStructuralFEA
    from pyto_amd.hex_thermal_fea import HexThermalFEA
    print("thermoelastic chain", layout, exact, flush=True)
    c = Coupled((12, 6, 4), layout)
    tmm.initialize_SIMP_THERMAL_PENALTY(c.pt)
    tmm.set_KVOID_RELATIVE(c.kv)
    tmm.initialize_SIMP_STRUCTURAL_PENALTY(c.ps)
    tmm.set_EVOID_RELATIVE(c.es)
    try:
        mesh = hex_mesher.HexMesher()
        mesh.grid_mesh((12, 6, 4), (1.0, 1.0, 1.0))
        mat = mat_lib.Material("TE", c.E, c.nu, 1.0, c.k0, 1.0, c.alpha)
        bcT = bound_cond.BC(force=c.q, fixed_dofs=c.fixT, dirichlet_values=c.gT)
        bcS = bound_cond.BC(force=c.fS, fixed_dofs=c.fixS, dirichlet_values=np.zeros(c.fixS.size))
        fe_t = HexThermalFEA(mesh, mat, bcT, linear_solvers.Solvers.DPCG, rtol=1e-13,
                             thermoElasticReferenceTemperature=c.t_ref)
        fe_t.h_layout = layout
        T = fe_t.solve(c.x, tmm.MaterialModel.SIMP)
        f_te = fe_t.get_thermoelastic_force(c.x, tmm.MaterialModel.SIMP)
        fe_s = HexStructuralFEA(mesh, mat, bcS, linear_solvers.Solvers.PCG, rtol=1e-13,
                                maxIters=20000)
        fe_s.thermo_elastic_force = f_te
        d = fe_s.solve(c.x, tmm.MaterialModel.SIMP)
        params = topopt_common.TOParams()
        J, g = S.compute_thermoelastic_compliance_and_gradient(c.x, T, d, params, fe_t, fe_s,
                                                               exact=exact)
        g_ref, T_ref, d_ref, _ = c.gradient(c.x, exact=exact)
        from oracle import sensitivities as OS
        f_ref = OS.thermoelastic_force(c.m, T_ref, c.x, c.H, c.E * c.alpha, c.t_ref, c.ps, c.es)
        assert rel_err(T, T_ref) <= 1e-10
        assert rel_err(f_te, f_ref) <= 1e-9
        assert rel_err(d, d_ref) <= 1e-8
        assert abs(J - float(d_ref @ (OS.fea.assemble_K(
            c.m, c.KE, OS.fea.simp_scaling(c.x, c.ps, c.es)) @ d_ref))) <= 1e-8 * abs(J)
        assert rel_err(g, g_ref) <= 1e-7
    finally:
        tmm.initialize_SIMP_THERMAL_PENALTY(1.0)
        tmm.set_KVOID_RELATIVE(1e-8)
        tmm.initialize_SIMP_STRUCTURAL_PENALTY(3.0)
        tmm.set_EVOID_RELATIVE(1e-8)
