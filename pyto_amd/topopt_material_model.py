"""SIMP material interpolation (src/topopt_material_model.py).

The reference keeps the penalty, void modulus, stress relaxation and p-norm
exponent as module globals (:6-26) with setters for the penalty only
(:35-74); here every one of them has a setter, because the workloads use
eps = 1e-9 and p = 8 where the reference hard-codes 1e-8 and 6.
"""
import enum

import numpy as np

_SIMP_STRUCTURAL_PENALTY_MIN = 1
_SIMP_STRUCTURAL_PENALTY_MAX = 6.0
_SIMP_STRUCTURAL_PENALTY = 3.0
_SIMP_STRESS_RELAXATION = 0.5
_EVOID_RELATIVE = 1e-8
_PNORM_EXPONENT = 6
_SIMP_THERMAL_PENALTY_MIN = 1
_SIMP_THERMAL_PENALTY_MAX = 3
_SIMP_THERMAL_PENALTY = 1.0
_KVOID_RELATIVE = 1e-8


class MaterialModel(enum.Enum):
    """src/topopt_material_model.py:28-31 (only SIMP runs on the GPU path)."""
    SIMP = enum.auto()
    RAMP = enum.auto()
    SIMPPLUS = enum.auto()


def set_SIMP_STRUCTURAL_PENALTY(value):
    global _SIMP_STRUCTURAL_PENALTY
    _SIMP_STRUCTURAL_PENALTY = float(value)


def initialize_SIMP_STRUCTURAL_PENALTY(value=None):
    global _SIMP_STRUCTURAL_PENALTY
    _SIMP_STRUCTURAL_PENALTY = 3.0 if value is None else float(value)


def increment_SIMP_STRUCTURAL_PENALTY(value):
    global _SIMP_STRUCTURAL_PENALTY
    _SIMP_STRUCTURAL_PENALTY = max(min(_SIMP_STRUCTURAL_PENALTY + value, _SIMP_STRUCTURAL_PENALTY_MAX),
                                   _SIMP_STRUCTURAL_PENALTY_MIN)
    return _SIMP_STRUCTURAL_PENALTY


def set_EVOID_RELATIVE(value):
    global _EVOID_RELATIVE
    _EVOID_RELATIVE = float(value)


def set_stress_relaxation(value):
    global _SIMP_STRESS_RELAXATION
    _SIMP_STRESS_RELAXATION = float(value)


def set_pNorm_exponent(value):
    global _PNORM_EXPONENT
    _PNORM_EXPONENT = float(value)


def get_SIMP_STRUCTURAL_PENALTY():
    return _SIMP_STRUCTURAL_PENALTY


def get_EVOID_RELATIVE():
    return _EVOID_RELATIVE


def _require_simp(material_model):
    if material_model not in (MaterialModel.SIMP,):
        raise ValueError("Invalid material model specified (only SIMP is supported on the GPU path).")


def get_structural_material_model_scaling(x, material_model):
    """src/topopt_material_model.py:76-95 (SIMP)."""
    if material_model is None:
        return x
    _require_simp(material_model)
    return _EVOID_RELATIVE + (x ** _SIMP_STRUCTURAL_PENALTY) * (1 - _EVOID_RELATIVE)


def get_structural_material_model_sensitivity(x, material_model):
    """src/topopt_material_model.py:119-139 (SIMP)."""
    if material_model is None:
        return np.ones_like(x)
    _require_simp(material_model)
    return _SIMP_STRUCTURAL_PENALTY * (x ** (_SIMP_STRUCTURAL_PENALTY - 1)) * (1 - _EVOID_RELATIVE)


def get_stress_relaxation_factor():
    return _SIMP_STRESS_RELAXATION


def get_stress_relaxation_correction(x):
    """src/topopt_material_model.py:199-201."""
    return (x ** _SIMP_STRESS_RELAXATION).reshape((-1, 1))


def get_stress_relaxation_factor_sensitivity(x):
    return (_SIMP_STRESS_RELAXATION * (x ** (_SIMP_STRESS_RELAXATION - 1))).reshape((-1, 1))


def get_pNorm_exponent():
    return _PNORM_EXPONENT


def initialize_SIMP_THERMAL_PENALTY(value=None):
    """src/topopt_material_model.py:68-74."""
    global _SIMP_THERMAL_PENALTY
    _SIMP_THERMAL_PENALTY = 1.0 if value is None else float(value)


def increment_SIMP_THERMAL_PENALTY(value):
    """src/topopt_material_model.py:53-58."""
    global _SIMP_THERMAL_PENALTY
    _SIMP_THERMAL_PENALTY = max(min(_SIMP_THERMAL_PENALTY + value, _SIMP_THERMAL_PENALTY_MAX),
                                _SIMP_THERMAL_PENALTY_MIN)
    return _SIMP_THERMAL_PENALTY


def get_SIMP_THERMAL_PENALTY():
    return _SIMP_THERMAL_PENALTY


def set_KVOID_RELATIVE(value):
    global _KVOID_RELATIVE
    _KVOID_RELATIVE = float(value)


def get_KVOID_RELATIVE():
    return _KVOID_RELATIVE


def get_thermal_material_model_scaling(x, material_model):
    """src/topopt_material_model.py:98-109 (SIMP)."""
    if material_model is None:
        return x
    _require_simp(material_model)
    return _KVOID_RELATIVE + (x ** _SIMP_THERMAL_PENALTY) * (1 - _KVOID_RELATIVE)


def get_thermal_material_model_sensitivity(x, material_model):
    """src/topopt_material_model.py:142-157 (SIMP)."""
    _require_simp(material_model)
    return _SIMP_THERMAL_PENALTY * (x ** (_SIMP_THERMAL_PENALTY - 1)) * (1 - _KVOID_RELATIVE)
