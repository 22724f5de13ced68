// Simulation configuration: the mesh / navier_stokes / output sections of the
// reference's YAML files (include/IO/YamlConfig.hpp:8-55,117-155,
// src/IO/YamlConfig.cpp:11-96,180-232), read by a small YAML-subset parser
// (block maps by indentation, plain or quoted scalars, [a, b, c] flow
// sequences, '-' block sequences, '#' comments) because yaml-cpp is not part of
// this build.  Defaults and the "section present but disabled" behaviour
// follow the reference; solid / dpm / fsi sections only report `enabled`.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace cfsa {

struct YamlNode {
  enum Kind { Null, Scalar, Map, Seq } kind = Null;
  std::string scalar;
  std::vector<std::pair<std::string, std::shared_ptr<YamlNode>>> map;
  std::vector<std::shared_ptr<YamlNode>> seq;
  const YamlNode *get(const std::string &key) const;
};

std::shared_ptr<YamlNode> parse_yaml(const std::string &text);

struct MeshConfig {
  enum Type { FromOBJ = 0, Cylinder = 1, IcoSphere = 2, Box = 3 };
  int type = Cylinder;
  std::string obj_file_path;
  double obj_scale = 1.0;
  double cylinder_length = 30.0, cylinder_radius = 2.0;
  int cylinder_divisions = 35;
  double icosphere_radius = 1.0;
  int icosphere_recursion = 2;
  double max_element_size = 0.03;
  double flow_direction[3] = {1.0, 0.0, 0.0};
  double inlet_outlet_coverage = 1.0;
  // extension (type "box"): structuredRectangularPrism(length, width, height, nx, ny, nz)
  double box_size[3] = {1.0, 1.0, 1.0};
  int box_cells[3] = {1, 1, 1};
  // extension (type "msh"): Gmsh MSH 4.1 file
  std::string msh_file_path;
};

struct NavierStokesConfig {
  bool enabled = false;
  double density = 1060.0, viscosity = 0.004;
  bool is_newtonian = true, is_laminar = true;
  double inlet_velocity[3] = {0.1, 0.0, 0.0};
  double outlet_pressure = 0.0;
  bool use_dirichlet_pressure = true;
  double dt = 1e-3, relax_p = 0.5, relax_u = 0.5;
  // extensions, all optional so reference files parse unchanged
  int n_corrections = 1;
  int pressure_solver = 0;  // "bicgstab" | "cg"
  int precision = 0;        // "fp64" | "mixed"
  double momentum_tol = 1e-6, pressure_tol = 1e-8;
  int inlet_profile = 0;    // "uniform" | "parabolic"
};

struct OutputConfig {
  std::string output_prefix = "output";
  int total_steps = 1000, output_frequency = 20;
  double vtk_scale = 1.0;
  bool output_fluid = true, output_solid = true, output_dpm = true;
};

struct SimulationConfig {
  MeshConfig mesh;
  NavierStokesConfig navier_stokes;
  OutputConfig output;
  bool solid_enabled = false, dpm_enabled = false, fsi_enabled = false;
};

// config::loadConfig (YamlConfig.cpp:200-232); throws std::runtime_error with
// the reference's "Configuration error: ..." / "YAML parsing error: ..." texts.
SimulationConfig load_config(const std::string &path);
SimulationConfig parse_config(const std::string &text);

}  // namespace cfsa
