"""Material records (src/mat_lib.py:8-159); only E and nu enter the structural path."""
import dataclasses
import difflib


@dataclasses.dataclass
class Material:
    """src/mat_lib.py:8-31."""
    name: str
    youngs_modulus: float
    poissons_ratio: float
    mass_density: float = 0.0
    thermal_conductivity: float = 0.0
    specific_heat: float = 0.0
    thermal_expansion_coefficient: float = 0.0
    cost: float = 0.0
    yield_strength: float = 0.0


materials = [
    Material("Steel", 2.1e11, 0.3, 7850, 50, 490, 1.2e-5, 1.0, 2.5e8),
    Material("Aluminum", 7.0e10, 0.33, 2700, 205, 900, 2.3e-5, 2.0, 2.7e8),
    Material("Titanium", 1.1e11, 0.34, 4500, 22, 520, 8.6e-6, 20.0, 8.8e8),
    Material("Unit", 1.0, 0.3, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0),
]


def get_material(name):
    """Case-insensitive lookup with a close-match hint (src/mat_lib.py:136-159)."""
    for m in materials:
        if m.name.lower() == str(name).lower():
            return m
    close = difflib.get_close_matches(str(name), [m.name for m in materials])
    raise ValueError("unknown material %r%s" % (name, (", did you mean %s?" % close[0]) if close else ""))
