#include "yaml_config.hpp"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace cfsa {

const YamlNode *YamlNode::get(const std::string &key) const {
  if (kind != Map) return nullptr;
  for (const auto &kv : map)
    if (kv.first == key) return kv.second.get();
  return nullptr;
}

namespace {

struct Line {
  int indent;
  std::string text;
  int lineno;
};

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

std::string strip_comment(const std::string &s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

[[noreturn]] void fail(int lineno, const std::string &msg) {
  throw std::runtime_error("YAML parsing error: line " + std::to_string(lineno) + ": " + msg);
}

std::string unquote(const std::string &v) {
  if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') ||
                        (v.front() == '\'' && v.back() == '\'')))
    return v.substr(1, v.size() - 2);
  return v;
}

std::shared_ptr<YamlNode> inline_value(const std::string &raw, int lineno) {
  auto n = std::make_shared<YamlNode>();
  const std::string v = trim(raw);
  if (v.empty() || v == "~" || v == "null") return n;
  if (v.front() == '[') {
    if (v.back() != ']') fail(lineno, "unterminated flow sequence");
    n->kind = YamlNode::Seq;
    const std::string body = trim(v.substr(1, v.size() - 2));
    if (body.empty()) return n;
    std::stringstream ss(body);
    std::string item;
    while (std::getline(ss, item, ',')) {
      auto c = std::make_shared<YamlNode>();
      c->kind = YamlNode::Scalar;
      c->scalar = unquote(trim(item));
      n->seq.push_back(c);
    }
    return n;
  }
  n->kind = YamlNode::Scalar;
  n->scalar = unquote(v);
  return n;
}

std::shared_ptr<YamlNode> parse_block(const std::vector<Line> &L, size_t &i, int indent);

std::shared_ptr<YamlNode> parse_seq(const std::vector<Line> &L, size_t &i, int indent) {
  auto n = std::make_shared<YamlNode>();
  n->kind = YamlNode::Seq;
  while (i < L.size() && L[i].indent == indent && L[i].text[0] == '-') {
    const std::string rest = trim(L[i].text.substr(1));
    const int lineno = L[i].lineno;
    ++i;
    if (rest.empty()) {
      if (i < L.size() && L[i].indent > indent)
        n->seq.push_back(parse_block(L, i, L[i].indent));
      else
        n->seq.push_back(std::make_shared<YamlNode>());
    } else {
      n->seq.push_back(inline_value(rest, lineno));
    }
  }
  return n;
}

std::shared_ptr<YamlNode> parse_block(const std::vector<Line> &L, size_t &i, int indent) {
  if (i < L.size() && L[i].text[0] == '-') return parse_seq(L, i, indent);
  auto n = std::make_shared<YamlNode>();
  n->kind = YamlNode::Map;
  while (i < L.size() && L[i].indent == indent) {
    const std::string &t = L[i].text;
    const int lineno = L[i].lineno;
    size_t colon = std::string::npos;
    char quote = 0;
    for (size_t k = 0; k < t.size(); ++k) {
      if (quote) {
        if (t[k] == quote) quote = 0;
      } else if (t[k] == '"' || t[k] == '\'') {
        quote = t[k];
      } else if (t[k] == ':' && (k + 1 == t.size() || t[k + 1] == ' ' || t[k + 1] == '\t')) {
        colon = k;
        break;
      }
    }
    if (colon == std::string::npos) fail(lineno, "expected 'key: value'");
    const std::string key = unquote(trim(t.substr(0, colon)));
    const std::string rest = trim(t.substr(colon + 1));
    ++i;
    std::shared_ptr<YamlNode> val;
    if (rest.empty()) {
      if (i < L.size() && (L[i].indent > indent || (L[i].indent == indent && L[i].text[0] == '-')))
        val = parse_block(L, i, L[i].indent);
      else
        val = std::make_shared<YamlNode>();
    } else {
      val = inline_value(rest, lineno);
    }
    n->map.push_back({key, val});
  }
  if (i < L.size() && L[i].indent > indent) fail(L[i].lineno, "unexpected indentation");
  return n;
}

// --- scalar conversions (yaml-cpp as<T>() accepts these spellings) ---
double as_double(const YamlNode *n, const std::string &key) {
  if (!n || n->kind != YamlNode::Scalar) throw std::runtime_error("bad conversion of '" + key + "'");
  const char *s = n->scalar.c_str();
  char *end = nullptr;
  errno = 0;
  double v = std::strtod(s, &end);
  if (end == s || *end != '\0' || errno == ERANGE)
    throw std::runtime_error("bad conversion of '" + key + "': " + n->scalar);
  return v;
}
int as_int(const YamlNode *n, const std::string &key) {
  if (!n || n->kind != YamlNode::Scalar) throw std::runtime_error("bad conversion of '" + key + "'");
  const char *s = n->scalar.c_str();
  char *end = nullptr;
  errno = 0;
  long v = std::strtol(s, &end, 10);
  if (end == s || *end != '\0' || errno == ERANGE)
    throw std::runtime_error("bad conversion of '" + key + "': " + n->scalar);
  return (int)v;
}
bool as_bool(const YamlNode *n, const std::string &key) {
  if (!n || n->kind != YamlNode::Scalar) throw std::runtime_error("bad conversion of '" + key + "'");
  static const char *yes[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"};
  static const char *no[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"};
  for (const char *y : yes)
    if (n->scalar == y) return true;
  for (const char *f : no)
    if (n->scalar == f) return false;
  throw std::runtime_error("bad conversion of '" + key + "': " + n->scalar);
}
std::string as_string(const YamlNode *n, const std::string &key) {
  if (!n || n->kind != YamlNode::Scalar) throw std::runtime_error("bad conversion of '" + key + "'");
  return n->scalar;
}

template <class T>
T get_or(const YamlNode &node, const std::string &key, T def);
template <>
double get_or(const YamlNode &n, const std::string &k, double d) {
  const YamlNode *c = n.get(k);
  return c ? as_double(c, k) : d;
}
template <>
int get_or(const YamlNode &n, const std::string &k, int d) {
  const YamlNode *c = n.get(k);
  return c ? as_int(c, k) : d;
}
template <>
bool get_or(const YamlNode &n, const std::string &k, bool d) {
  const YamlNode *c = n.get(k);
  return c ? as_bool(c, k) : d;
}
template <>
std::string get_or(const YamlNode &n, const std::string &k, std::string d) {
  const YamlNode *c = n.get(k);
  return c ? as_string(c, k) : d;
}

void vec3(const YamlNode *n, double out[3]) {
  if (!n || n->kind != YamlNode::Seq || n->seq.size() != 3)
    throw std::runtime_error("Invalid Vector3d format - expected [x, y, z]");
  for (int c = 0; c < 3; ++c) out[c] = as_double(n->seq[(size_t)c].get(), "vector");
}
void ivec3(const YamlNode *n, int out[3]) {
  if (!n || n->kind != YamlNode::Seq || n->seq.size() != 3)
    throw std::runtime_error("Invalid vector format - expected [nx, ny, nz]");
  for (int c = 0; c < 3; ++c) out[c] = as_int(n->seq[(size_t)c].get(), "vector");
}

// YamlConfig.cpp:30-66
MeshConfig parse_mesh(const YamlNode &node) {
  MeshConfig c;
  const YamlNode *t = node.get("type");
  if (!t) throw std::runtime_error("Mesh type not specified");
  const std::string type = as_string(t, "type");
  if (type == "obj") {
    c.type = MeshConfig::FromOBJ;
    if (!node.get("obj_file")) throw std::runtime_error("obj_file required for OBJ mesh type");
    c.obj_file_path = as_string(node.get("obj_file"), "obj_file");
    c.obj_scale = get_or(node, "obj_scale", 1.0);
  } else if (type == "cylinder") {
    c.type = MeshConfig::Cylinder;
    c.cylinder_length = get_or(node, "length", 30.0);
    c.cylinder_radius = get_or(node, "radius", 2.0);
    c.cylinder_divisions = get_or(node, "divisions", 35);
  } else if (type == "icosphere") {
    c.type = MeshConfig::IcoSphere;
    c.icosphere_radius = get_or(node, "radius", 1.0);
    c.icosphere_recursion = get_or(node, "recursion", 2);
  } else if (type == "box") {
    c.type = MeshConfig::Box;
    if (node.get("size")) vec3(node.get("size"), c.box_size);
    if (node.get("cells")) ivec3(node.get("cells"), c.box_cells);
  } else if (type == "msh") {
    c.type = 4;
    if (!node.get("msh_file")) throw std::runtime_error("msh_file required for MSH mesh type");
    c.msh_file_path = as_string(node.get("msh_file"), "msh_file");
  } else {
    throw std::runtime_error("Unknown mesh type: " + type);
  }
  c.max_element_size = get_or(node, "max_element_size", 0.03);
  if (node.get("flow_direction")) vec3(node.get("flow_direction"), c.flow_direction);
  c.inlet_outlet_coverage = get_or(node, "inlet_outlet_coverage", 1.0);
  return c;
}

// YamlConfig.cpp:68-96 (+ optional extension keys)
NavierStokesConfig parse_ns(const YamlNode *node) {
  NavierStokesConfig c;
  if (!node || node->kind != YamlNode::Map) return c;
  c.enabled = get_or(*node, "enabled", false);
  if (!c.enabled) return c;
  c.density = get_or(*node, "density", 1060.0);
  c.viscosity = get_or(*node, "viscosity", 0.004);
  c.is_newtonian = get_or(*node, "is_newtonian", true);
  c.is_laminar = get_or(*node, "is_laminar", true);
  if (node->get("inlet_velocity")) vec3(node->get("inlet_velocity"), c.inlet_velocity);
  c.outlet_pressure = get_or(*node, "outlet_pressure", 0.0);
  c.use_dirichlet_pressure = get_or(*node, "use_dirichlet_pressure", true);
  c.dt = get_or(*node, "dt", 1e-3);
  c.relax_p = get_or(*node, "relax_p", 0.5);
  c.relax_u = get_or(*node, "relax_u", 0.5);
  c.n_corrections = get_or(*node, "n_corrections", 1);
  const std::string ps = get_or(*node, "pressure_solver", std::string("bicgstab"));
  if (ps == "bicgstab") c.pressure_solver = 0;
  else if (ps == "cg") c.pressure_solver = 1;
  else if (ps == "amg") c.pressure_solver = 2;
  else throw std::runtime_error("Unknown pressure_solver: " + ps);
  const std::string pr = get_or(*node, "precision", std::string("fp64"));
  if (pr == "fp64") c.precision = 0;
  else if (pr == "mixed") c.precision = 1;
  else throw std::runtime_error("Unknown precision: " + pr);
  c.momentum_tol = get_or(*node, "momentum_tol", 1e-6);
  c.pressure_tol = get_or(*node, "pressure_tol", 1e-8);
  const std::string ip = get_or(*node, "inlet_profile", std::string("uniform"));
  if (ip == "uniform") c.inlet_profile = 0;
  else if (ip == "parabolic") c.inlet_profile = 1;
  else throw std::runtime_error("Unknown inlet_profile: " + ip);
  return c;
}

// YamlConfig.cpp:180-196
OutputConfig parse_output(const YamlNode *node) {
  OutputConfig c;
  if (!node || node->kind != YamlNode::Map) return c;
  c.output_prefix = get_or(*node, "prefix", std::string("output"));
  c.total_steps = get_or(*node, "total_steps", 1000);
  c.output_frequency = get_or(*node, "output_frequency", 20);
  c.vtk_scale = get_or(*node, "vtk_scale", 1.0);
  c.output_fluid = get_or(*node, "output_fluid", true);
  c.output_solid = get_or(*node, "output_solid", true);
  c.output_dpm = get_or(*node, "output_dpm", true);
  return c;
}

bool enabled(const YamlNode *node) {
  if (!node || node->kind != YamlNode::Map) return false;
  return get_or(*node, "enabled", false);
}

}  // namespace

std::shared_ptr<YamlNode> parse_yaml(const std::string &text) {
  std::vector<Line> lines;
  std::stringstream ss(text);
  std::string raw;
  int lineno = 0;
  while (std::getline(ss, raw)) {
    ++lineno;
    if (raw.find('\t') != std::string::npos && raw.find_first_not_of(" \t") != std::string::npos &&
        raw.find_first_not_of(' ') < raw.size() && raw[raw.find_first_not_of(' ')] == '\t')
      fail(lineno, "tab indentation");
    const std::string s = strip_comment(raw);
    const std::string t = trim(s);
    if (t.empty() || t == "---") continue;
    lines.push_back({(int)s.find_first_not_of(' '), t, lineno});
  }
  size_t i = 0;
  if (lines.empty()) return std::make_shared<YamlNode>();
  auto root = parse_block(lines, i, lines[0].indent);
  if (i != lines.size()) fail(lines[i].lineno, "unexpected content");
  return root;
}

SimulationConfig parse_config(const std::string &text) {
  std::shared_ptr<YamlNode> root;
  root = parse_yaml(text);
  try {
    SimulationConfig cfg;
    const YamlNode *mesh = root->get("mesh");
    if (!mesh) throw std::runtime_error("Mesh configuration is required");
    cfg.mesh = parse_mesh(*mesh);
    cfg.navier_stokes = parse_ns(root->get("navier_stokes"));
    cfg.solid_enabled = enabled(root->get("solid_mechanics"));
    cfg.dpm_enabled = enabled(root->get("dpm"));
    cfg.fsi_enabled = enabled(root->get("fsi_coupling"));
    cfg.output = parse_output(root->get("output"));
    return cfg;
  } catch (const std::exception &e) {
    throw std::runtime_error(std::string("Configuration error: ") + e.what());
  }
}

SimulationConfig load_config(const std::string &path) {
  std::ifstream in(path);
  if (!in.is_open())
    throw std::runtime_error("YAML parsing error: bad file: " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return parse_config(ss.str());
}

}  // namespace cfsa
