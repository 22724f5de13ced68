#include "vtk.hpp"

#include <fstream>

namespace cfsa {

bool write_mesh_vtk(const std::string &path, const double *coords, int64_t nv,
                    const int32_t *tets, int64_t nt, const uint8_t *p1_bc, double scale) {
  std::ofstream out(path);
  if (!out.is_open()) return false;
  out << "# vtk DataFile Version 3.0\n";
  out << "Shape export\n";
  out << "ASCII\n";
  out << "DATASET " << "UNSTRUCTURED_GRID" << "\n";
  out << "POINTS " << nv << " double\n";
  for (int64_t i = 0; i < nv; ++i)
    out << coords[i * 3] * scale << " " << coords[i * 3 + 1] * scale << " "
        << coords[i * 3 + 2] * scale << "\n";
  out << "CELLS " << nt << " " << nt * 5 << "\n";
  for (int64_t t = 0; t < nt; ++t)
    out << "4 " << tets[t * 4] << " " << tets[t * 4 + 1] << " " << tets[t * 4 + 2] << " "
        << tets[t * 4 + 3] << "\n";
  out << "CELL_TYPES " << nt << "\n";
  for (int64_t t = 0; t < nt; ++t) out << "10\n";
  out << "\nPOINT_DATA " << nv << "\n";
  out << "SCALARS FluidBoundaryType int 1\n";
  out << "LOOKUP_TABLE default\n";
  for (int64_t i = 0; i < nv; ++i) out << static_cast<int>(p1_bc[i]) << "\n";
  return static_cast<bool>(out);
}

bool append_ns_fields_vtk(const std::string &path, const double *pressure,
                          const double *velocity, int64_t nv, double scale) {
  std::ofstream out(path, std::ios::app);
  if (!out.is_open()) return false;
  out << "\nSCALARS pressure double 1\n";
  out << "LOOKUP_TABLE default\n";
  for (int64_t i = 0; i < nv; ++i) out << pressure[i] << "\n";
  out << "\nVECTORS velocity double\n";
  for (int64_t i = 0; i < nv; ++i)
    out << velocity[i * 3] * scale << " " << velocity[i * 3 + 1] * scale << " "
        << velocity[i * 3 + 2] * scale << "\n";
  return static_cast<bool>(out);
}

}  // namespace cfsa
