"""Types and helpers of the optimisation layer (src/topopt_common.py)."""
import enum
from dataclasses import dataclass, field
from typing import Tuple

import numpy as np

from .topopt_material_model import MaterialModel

DIRECT_SOLVER_DOF_CUTOFF = 100000  # src/topopt_common.py:18


class TO_METHODS(enum.Enum):
    """src/topopt_common.py:20-24 (the two density methods are on the GPU
    path: DENSITYOCM in topopt_ocm, DENSITYMMA in topopt_mma)."""
    DENSITYMMA = enum.auto()
    DENSITYOCM = enum.auto()
    PARETO = enum.auto()
    LEVELSET = enum.auto()


class FEA_MODE(enum.Enum):
    """src/topopt_common.py:27-30."""
    STRUCTURAL = enum.auto()
    THERMAL = enum.auto()
    THERMO_STRUCTURAL = enum.auto()


class TO_QOI(enum.Enum):
    """src/topopt_common.py:32-47."""
    COMPLIANCE = enum.auto()
    VOLUME_FRACTION = enum.auto()
    MASS = enum.auto()
    PNORM_STRESS = enum.auto()
    MAX_VONMISES_STRESS = enum.auto()
    STRESS_FAILURE_FACTOR = enum.auto()
    GVECTOR = enum.auto()


@dataclass(slots=True)
class TOParams:
    """src/topopt_common.py:50-77 (fields of the structural path)."""
    Comment: str = ""
    Objective: Tuple = (TO_QOI.COMPLIANCE, None)
    Constraints: list = field(default_factory=lambda: [(TO_QOI.VOLUME_FRACTION, None, 0.5)])
    nDOFDesired: int = 25000
    APPLY_FILTER_TO_DENSITY: bool = True
    RelativeFilterRadius: float = 1.5
    ElemsToKeep: list = None
    MaxIterations: int = 150
    Eliminate_Hanging_Elements: bool = False
    materialModel: MaterialModel = MaterialModel.SIMP


def find_elements_with_forces(mesh, force, nDOFPerNode):
    """src/topopt_common.py:79-100."""
    force = np.asarray(force)
    forced_nodes = np.unique(np.nonzero(force != 0)[0] // nDOFPerNode)
    is_forced = np.isin(mesh.elemArray, forced_nodes)
    return np.nonzero(is_forced.any(axis=1))[0]


def createFilters(fe_solver, to_params):
    """src/topopt_common.py:127-159 (smoothing filter only; the symmetry and
    extrusion filters are not part of this path)."""
    from .topopt_filters import createSmoothingFilter
    return createSmoothingFilter(fe_solver.mesh, rel_filter_radius=to_params.RelativeFilterRadius,
                                 operator=fe_solver.stiff_mtrx)


def write_summary_row(result_csv_file, row):
    """One result row into a per-method summary CSV, replacing the row of the
    same 'name' or appending it (src/topopt_run_benchmarks.py:218-250)."""
    import os
    import pandas as pd
    if os.path.exists(result_csv_file):
        existing_df = pd.read_csv(result_csv_file)
        idx = existing_df[existing_df["name"] == row["name"]].index
        if len(idx) > 0:
            existing_df.loc[idx[0]] = row
        else:
            existing_df = pd.concat([existing_df, pd.DataFrame([row])], ignore_index=True)
        existing_df.to_csv(result_csv_file, index=False)
    else:
        d = os.path.dirname(os.path.abspath(result_csv_file))
        os.makedirs(d, exist_ok=True)
        pd.DataFrame([row]).to_csv(result_csv_file, index=False)
