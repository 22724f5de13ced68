// Legacy ASCII VTK writer, reference io::exportToVtk overloads for Mesh and
// NavierStokesSolver (src/IO/VtkExport.cpp:11-16, 49-84, 238-268).  Numbers
// go through std::ostream with its default precision (6 significant digits)
// like the reference, so files compare byte for byte.
#pragma once

#include <cstdint>
#include <string>

namespace cfsa {

bool write_mesh_vtk(const std::string &path, const double *coords, int64_t nv,
                    const int32_t *tets, int64_t nt, const uint8_t *p1_bc, double scale);
// Appends "pressure" (P1) and "velocity" (P1 vertices, times scale).
bool append_ns_fields_vtk(const std::string &path, const double *pressure,
                          const double *velocity, int64_t nv, double scale);

}  // namespace cfsa
