"""Element matrices of the 8-node brick (src/hex_element_stiffness.py).

KE comes from the library's Hadamard-basis element operator applied to unit
vectors (``pyto_ke_hex8``), which equals the 2x2x2 Gauss KE of
src/hex_element_stiffness.py:231-286 to rounding; the thermal KE_T likewise
(``pyto_ke_hex8_thermal``, src/hex_element_stiffness.py:339-385).
"""
import ctypes

import numpy as np

from . import _lib


def isotropic_constitutive_matrix(E, nu):
    """src/hex_element_stiffness.py:220-228."""
    c = E / ((1. + nu) * (1. - 2. * nu))
    g = (1. - 2. * nu) / 2.
    return c * np.array([[1. - nu, nu, nu, 0., 0., 0.],
                         [nu, 1. - nu, nu, 0., 0., 0.],
                         [nu, nu, 1. - nu, 0., 0., 0.],
                         [0., 0., 0., g, 0., 0.],
                         [0., 0., 0., 0., g, 0.],
                         [0., 0., 0., 0., 0., g]])


def hex8_stiffness_matrix_structural(E, nu, elem_size, gauss_order=2):
    """24x24 KE, row/col = 3*local_node + component (src/hex_element_stiffness.py:231-286)."""
    if gauss_order not in (2, 6):
        raise ValueError("gauss_order must be 2 or 6")
    ke = np.zeros((24, 24), dtype=np.float64)
    hx, hy, hz = (float(v) for v in elem_size)
    _lib.check(_lib.lib().pyto_ke_hex8(float(E), float(nu), hx, hy, hz,
                                      ke.ctypes.data_as(ctypes.c_void_p)), "pyto_ke_hex8")
    return ke


def hex8_stiffness_matrix_thermal(thermal_conductivity, elem_size, gauss_order=2):
    """8x8 KE_T, row/col = local node (src/hex_element_stiffness.py:339-385)."""
    if gauss_order not in (2, 6):
        raise ValueError("gauss_order must be 2 or 6")
    ke = np.zeros((8, 8), dtype=np.float64)
    hx, hy, hz = (float(v) for v in elem_size)
    _lib.check(_lib.lib().pyto_ke_hex8_thermal(float(thermal_conductivity), hx, hy, hz,
                                              ke.ctypes.data_as(ctypes.c_void_p)),
               "pyto_ke_hex8_thermal")
    return ke
