"""CPU oracle for the voxel-hex structural FEA path (TEST INFRASTRUCTURE ONLY).

This package is a plain numpy/scipy FP64 restatement of PyTO's algorithm for the
hex structural solve, its deflated PCG, the cone filter, the compliance and
p-norm stress sensitivities and the OC update.  Each function cites the
reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU
baseline).  The product package ``pyto_amd`` never imports it; the product
path runs on the HIP library and fails loudly when that library is missing.

Pinning (what checks each function independently of its own formula):

* ``ke.hex8_stiffness``: exact rational integration (``ke.hex8_stiffness_exact``),
  symmetry, rank 18 with the six rigid-body null vectors, and the patch test
  u^T KE u = V eps^T D eps for linear fields (tests/test_oracle_ke.py).
* ``mesh``: brute-force restatement of the reference's nested loops on tiny
  grids (tests/test_oracle_mesh.py).
* ``fea.assemble_K`` / ``impose_dirichlet``: dense brute-force assembly,
  symmetry, K*rigid = 0 on a free-floating grid.
* ``deflation.deflated_pcg``: scipy ``splu`` direct solve of the same system
  and the true residual ||f-Ku||/||f||; W columns are exact rigid-body modes.
* ``fea.postprocess``: constant-strain patch fields and closed-form von Mises
  states; the uniaxial bar with sliding supports (sigma = F/A exactly).
* ``sensitivities``: C = f^T u = u^T K u and central finite differences for the
  compliance and the p-norm gradients.
* ``filters.cone_filter``: brute-force all-pairs distance query, closed-form
  weights {1, 1/3, (1.5-sqrt 2)/1.5}, symmetry.
* ``ocm.oc_update``: the bisection meets the volume constraint; move limits.

* ``deflation.create_groups`` (``src/deflation.py:86-229``): box counts per
  axis decided in exact rational arithmetic, the node-to-box assignment against
  an integer brute force, centres against closed-form lattice means, and the
  small-box merge on hand-built grids written out by hand
  (tests/test_oracle_groups.py).

Beyond these pins, identity with the reference's own DPCG iterate sequence
(``src/deflation.py:504-584``, SuperLU refactorised each iteration) is not
checked element by element: the reference cannot be imported here (numba,
pyvista, pyamg, numpy-stl are absent).  Converged displacements and compliance
do not depend on the iterate sequence, and the oracle's DPCG is pinned by the
direct solves above.
"""
