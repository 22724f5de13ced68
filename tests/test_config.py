"""YAML configuration (reference config::loadConfig) on the host."""
from pathlib import Path

import pytest

NS_ONLY = """\
# Navier-Stokes only, same keys as the reference's navier_stokes_only example
mesh:
  type: cylinder
  length: 30.0
  radius: 2.0
  divisions: 35
  max_element_size: 0.03
  flow_direction: [1.0, 0.0, 0.0]
  inlet_outlet_coverage: 1.0

navier_stokes:
  enabled: true
  density: 1060.0
  viscosity: 0.004
  is_newtonian: true
  is_laminar: true
  inlet_velocity: [0.1, 0.0, 0.0]  # trailing comment
  outlet_pressure: 0.0
  use_dirichlet_pressure: true
  dt: 0.001
  relax_p: 0.5
  relax_u: 0.5

solid_mechanics:
  enabled: false

output:
  prefix: fluid_output
  total_steps: 1000
  output_frequency: 20
  vtk_scale: 1.0
  output_fluid: true
  output_solid: false
"""


def test_reference_style_file(P, tmp_path):
    f = tmp_path / "ns.yaml"
    f.write_text(NS_ONLY)
    c = P.loadConfig(f)
    assert c["mesh_type"] == 1 and c["cylinder_length"] == 30.0 and c["cylinder_divisions"] == 35
    assert c["max_element_size"] == 0.03 and c["flow_direction"] == [1.0, 0.0, 0.0]
    assert c["ns_enabled"] == 1 and c["density"] == 1060.0 and c["viscosity"] == 0.004
    assert c["inlet_velocity"] == [0.1, 0.0, 0.0] and c["dt"] == 0.001
    assert c["relax_p"] == 0.5 and c["relax_u"] == 0.5
    assert c["output_prefix"] == "fluid_output" and c["total_steps"] == 1000
    assert c["output_frequency"] == 20 and c["output_fluid"] == 1 and c["output_solid"] == 0
    assert c["output_dpm"] == 1  # default (YamlConfig.hpp:124)
    assert c["solid_enabled"] == 0 and c["dpm_enabled"] == 0
    # extension keys default to the reference's behaviour
    assert c["n_corrections"] == 1 and c["pressure_solver"] == 0 and c["precision"] == 0
    assert c["momentum_tol"] == 1e-6 and c["pressure_tol"] == 1e-8


def test_defaults_and_disabled_section(P):
    c = P.parseConfig("mesh:\n  type: cylinder\nnavier_stokes:\n  enabled: false\n  density: 5\n")
    # a disabled section keeps every default (YamlConfig.cpp:75-78)
    assert c["ns_enabled"] == 0 and c["density"] == 1060.0 and c["dt"] == 1e-3
    assert c["cylinder_length"] == 30.0 and c["cylinder_radius"] == 2.0
    assert c["inlet_velocity"] == [0.1, 0.0, 0.0]
    assert c["total_steps"] == 1000 and c["output_prefix"] == "output"


def test_extension_keys(P):
    c = P.parseConfig("mesh:\n  type: box\n  size: [2, 1, 1]\n  cells: [8, 4, 4]\n"
                      "navier_stokes:\n  enabled: yes\n  n_corrections: 2\n"
                      "  pressure_solver: cg\n  precision: mixed\n  momentum_tol: 1e-8\n"
                      "  inlet_profile: parabolic\n")
    assert c["mesh_type"] == 3 and c["box_size"] == [2.0, 1.0, 1.0] and c["box_cells"] == [8, 4, 4]
    assert c["n_corrections"] == 2 and c["pressure_solver"] == 1 and c["precision"] == 1
    assert c["momentum_tol"] == 1e-8 and c["inlet_profile"] == 1
    c = P.parseConfig("mesh:\n  type: cylinder\nnavier_stokes:\n  enabled: yes\n"
                      "  pressure_solver: amg\n")
    assert c["pressure_solver"] == 2


@pytest.mark.parametrize("text,msg", [
    ("navier_stokes:\n  enabled: true\n", "Mesh configuration is required"),
    ("mesh:\n  length: 3\n", "Mesh type not specified"),
    ("mesh:\n  type: torus\n", "Unknown mesh type: torus"),
    ("mesh:\n  type: obj\n", "obj_file required"),
    ("mesh:\n  type: cylinder\n  flow_direction: [1, 0]\n", "Invalid Vector3d format"),
    ("mesh:\n  type: cylinder\n  divisions: many\n", "bad conversion"),
    ("mesh:\n  type: cylinder\nnavier_stokes:\n  enabled: true\n  dt: fast\n", "bad conversion"),
])
def test_errors(P, text, msg):
    with pytest.raises(P.CfsaError) as e:
        P.parseConfig(text)
    assert msg in str(e.value)


def test_missing_file(P, tmp_path):
    with pytest.raises(P.CfsaError) as e:
        P.loadConfig(tmp_path / "nope.yaml")
    assert e.value.code == 7


REF_CONFIGS = Path("/root/reference/config")


@pytest.mark.skipif(not REF_CONFIGS.exists(), reason="reference checkout not present")
def test_parses_every_reference_config(P):
    for f in sorted(REF_CONFIGS.glob("*.yaml")):
        c = P.loadConfig(f)
        assert c["mesh_type"] in (0, 1, 2), f.name
        assert c["total_steps"] > 0 and c["output_frequency"] > 0
