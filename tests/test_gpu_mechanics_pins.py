"""The closed forms of tests/test_oracle_mechanics_pins.py on the device
(FP64, hc_stage_mechanics): the regular-octahedron term sizes (link, area,
volume, bending, inner links, membrane viscosity with its cap) and the
platelet's dihedral bending against the brute-force evaluation."""
import pytest

import hemocell_amd as hc
import test_oracle_mechanics_pins as pins

pytestmark = pytest.mark.gpu


def device_box():
    s = hc.Simulation(0)
    s.lattice_init(40, 40, 40, 1.0, periodic=(1, 1, 1), precision=hc.HC_FP64)
    return s


def test_device_octahedron_term_closed_forms():
    pins.check_octahedron_terms(device_box)


def test_device_octahedron_inner_link_closed_forms():
    pins.check_octahedron_inner_links(device_box)


def test_device_octahedron_membrane_viscosity_closed_forms():
    pins.check_octahedron_viscosity(device_box)


def test_device_platelet_dihedral_brute_force():
    pins.check_platelet_dihedral(device_box)
