// Host-side tetrahedral mesh layer: geometry, P1 gradients, faces, P2 edge-node
// numbering and fluid boundary labels.
//
// Mirrors the reference Mesh class (include/Mesh/Mesh.hpp:15-220,
// src/Mesh/Mesh.cpp:16-685) in behaviour and numbering, but stores everything
// as flat arrays and builds faces / edges with sorts instead of std::map so a
// 5.3M-tet mesh is processed in seconds.  Every floating-point expression that
// feeds an integer decision (face centres for boundary labels) is evaluated in
// the same operation order as the reference so labels agree bit for bit.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace cfsa {

// Values match enum class FluidBCType (include/FEM/NavierStokes.hpp:13-19);
// the integer is also what the VTK writer prints.
enum FluidBC : uint8_t {
  BC_WALL = 0,
  BC_INLET = 1,
  BC_OUTLET = 2,
  BC_INTERNAL = 3,
  BC_UNDEFINED = 4,
};

struct Face {
  int32_t tet_a = -1;
  int32_t tet_b = -1;
  int32_t face_id_a = -1;
  int32_t face_id_b = -1;
  bool is_ecm = false;
  double area = 0.0;
  std::array<int32_t, 3> vertids{};
  std::array<double, 3> normal{};
  std::array<double, 3> center{};
};

class Mesh {
 public:
  // --- construction (reference: Mesh.hpp:70-93) ---
  static Mesh structuredRectangularPrism(double length, double width,
                                         double height, int nx, int ny, int nz);
  static Mesh fromArrays(const double *coords, int64_t nv, const int32_t *tets,
                         int64_t nt);
  static Mesh fromMshFile(const std::string &path);
  // Own cylinder mesher (the reference meshes cylinders with TetGen, which is
  // not part of this build): a ring-structured disc triangulation extruded
  // along +x, every triangular prism split into 3 tets by the
  // smallest-global-index rule so neighbouring prisms agree on quad diagonals.
  static Mesh cylinder(double length, double radius, int divisions,
                       double max_volume);

  // reference: Mesh::setupBoundaryConditions (Mesh.cpp:325-417)
  static void setupBoundaryConditions(const double inlet_to_outlet_dir[3],
                                      double percent_coverage, Mesh &mesh);
  // reference: Mesh::setP2BoundariesFromP1Boundaries (Mesh.hpp:198-215)
  void setP2BoundariesFromP1Boundaries();
  // reference: Mesh::updateGeometry (Mesh.hpp:98-101), which calls
  // computeShapeFunctionGradients once more after computeGeometry (whose last
  // step it already is); the repeat recomputes the same values and is dropped.
  void updateGeometry() { computeGeometry(); }

  int64_t nVertices() const { return (int64_t)vertices.size() / 3; }
  int64_t nTets() const { return (int64_t)tets.size() / 4; }
  int64_t nFaces() const { return (int64_t)faces.size(); }
  int64_t nEdgeNodes() const { return (int64_t)edge_nodes.size() / 3; }
  bool hasDegenerateTet() const;

  // --- data (flat, row-major) ---
  std::vector<double> vertices;          // nv x 3
  std::vector<int32_t> tets;             // nt x 4
  std::vector<double> volumes;           // nt
  std::vector<double> centroids;         // nt x 3
  std::vector<double> gradients;         // nt x 4 x 3  (P1 grad phi_i)
  std::vector<int32_t> tet_faceids;      // nt x 4
  std::vector<int32_t> tet_neighbors;    // nt x 4
  std::vector<Face> faces;               // sorted-vertex-triple order
  std::vector<double> edge_nodes;        // ne x 3 (midpoints)
  std::vector<int32_t> edge_endpoints;   // ne x 2 (lo, hi vertex id)
  std::vector<int32_t> tet_edge_nodes;   // nt x 6 (local edges 01,02,03,12,13,23)
  std::vector<uint8_t> p1_bc;            // nv
  std::vector<uint8_t> p2_bc;            // ne
  double min_edge_length = 0.0;

  // Tet volumes (|6V|/6) and P1 gradients from raw arrays, same arithmetic as
  // computeGeometry / computeShapeFunctionGradients.
  static void tetGeometry(const double *coords, const int32_t *tets, int64_t nt,
                          double *vol, double *grad);

  void finalize();  // connectivity, geometry, normals, gradients, P2 numbering
  void buildConnectivity();
  void computeGeometry();
  void ensureConsistentFaceNormals();
  void computeShapeFunctionGradients();
  void buildP2EdgeNodes();
  void computeMinEdgeLength();
};

}  // namespace cfsa
