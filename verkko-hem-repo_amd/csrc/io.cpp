// VTU/PVTU output of the order parameter and the Newton update
// (FemGL::output_results, femgl/src/io.cc:105-168): point data du_ij, dv_ij
// (the update) and u_ij, v_ij (the solution) with the reference's component
// names, cell data "subdomain", written as solution_NN.0.vtu plus the
// solution_NN.pvtu record of deal.II's write_vtu_with_pvtu_record.  Q1 cells
// are one VTK hexahedron each; Q2 cells are written as their 8 sub-hexahedra
// through the 27 nodes.
#include <cstdio>
#include <string>
#include <vector>

#include "glhip.h"
#include "handles.hpp"

namespace {

const char *NAMES[18] = {"11", "12", "13", "21", "22", "23", "31", "32", "33",
                         "11", "12", "13", "21", "22", "23", "31", "32", "33"};

void write_field(FILE *f, const char *prefix_u, const char *prefix_v, const double *V, int n_nodes)
{
  for (int c = 0; c < 18; ++c) {
    std::fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s%s\" format=\"ascii\">\n", c < 9 ? prefix_u : prefix_v,
                 NAMES[c]);
    for (int v = 0; v < n_nodes; ++v) std::fprintf(f, "%.17g\n", V[(size_t)v * 18 + c]);
    std::fprintf(f, "        </DataArray>\n");
  }
}

}  // namespace

extern "C" int glh_write_vtu(const glh_mesh *mm, const double *U, const double *dU, const char *dir, int counter)
{
  if (!mm || !U || !dir || counter < 0) return GLH_ERR_ARG;
  const glh::Mesh &m = mm->m;
  const int n1 = m.n1;
  // VTK hexahedron corner order from the lexicographic local order
  const int vtk[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  std::vector<int> hex;
  for (int e = 0; e < m.n_cells; ++e) {
    const int *cn = &m.cell_nodes[(size_t)e * m.ns];
    const int sub = n1 - 1;  // 1 for Q1, 2 for Q2 (8 sub-cells)
    for (int kz = 0; kz < sub; ++kz)
      for (int ky = 0; ky < sub; ++ky)
        for (int kx = 0; kx < sub; ++kx)
          for (int q = 0; q < 8; ++q) {
            const int a = vtk[q];
            const int ax = kx + (a & 1), ay = ky + ((a >> 1) & 1), az = kz + (a >> 2);
            hex.push_back(cn[ax + n1 * (ay + n1 * az)]);
          }
  }
  const int n_hex = (int)(hex.size() / 8);
  char num[16];
  std::snprintf(num, sizeof(num), "%02d", counter);
  const std::string base = std::string("solution_") + num;
  const std::string vtu = std::string(dir) + "/" + base + ".0.vtu", pvtu = std::string(dir) + "/" + base + ".pvtu";
  FILE *f = std::fopen(vtu.c_str(), "w");
  if (!f) return GLH_ERR_ARG;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n  <UnstructuredGrid>\n");
  std::fprintf(f, "    <Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n", m.n_nodes, n_hex);
  std::fprintf(f, "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int v = 0; v < m.n_nodes; ++v)
    std::fprintf(f, "%.17g %.17g %.17g\n", m.node_xyz[3 * v], m.node_xyz[3 * v + 1], m.node_xyz[3 * v + 2]);
  std::fprintf(f, "        </DataArray>\n      </Points>\n      <Cells>\n");
  std::fprintf(f, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (size_t i = 0; i < hex.size(); i += 8)
    std::fprintf(f, "%d %d %d %d %d %d %d %d\n", hex[i], hex[i + 1], hex[i + 2], hex[i + 3], hex[i + 4], hex[i + 5],
                 hex[i + 6], hex[i + 7]);
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (int i = 1; i <= n_hex; ++i) std::fprintf(f, "%d\n", 8 * i);
  std::fprintf(f, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (int i = 0; i < n_hex; ++i) std::fprintf(f, "12\n");
  std::fprintf(f, "        </DataArray>\n      </Cells>\n      <PointData>\n");
  if (dU) write_field(f, "du_", "dv_", dU, m.n_nodes);
  write_field(f, "u_", "v_", U, m.n_nodes);
  std::fprintf(f, "      </PointData>\n      <CellData>\n");
  std::fprintf(f, "        <DataArray type=\"Float32\" Name=\"subdomain\" format=\"ascii\">\n");
  for (int i = 0; i < n_hex; ++i) std::fprintf(f, "0\n");
  std::fprintf(f, "        </DataArray>\n      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  const bool ok_vtu = std::fclose(f) == 0;
  FILE *g = std::fopen(pvtu.c_str(), "w");
  if (!g) return GLH_ERR_ARG;
  std::fprintf(g, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n  <PUnstructuredGrid GhostLevel=\"0\">\n");
  std::fprintf(g, "    <PPointData>\n");
  for (int pass = dU ? 0 : 1; pass < 2; ++pass)
    for (int c = 0; c < 18; ++c)
      std::fprintf(g, "      <PDataArray type=\"Float64\" Name=\"%s%s\"/>\n",
                   pass == 0 ? (c < 9 ? "du_" : "dv_") : (c < 9 ? "u_" : "v_"), NAMES[c]);
  std::fprintf(g, "    </PPointData>\n    <PCellData>\n      <PDataArray type=\"Float32\" Name=\"subdomain\"/>\n"
                  "    </PCellData>\n    <PPoints>\n      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n"
                  "    </PPoints>\n    <Piece Source=\"%s.0.vtu\"/>\n  </PUnstructuredGrid>\n</VTKFile>\n",
               base.c_str());
  const bool ok_p = std::fclose(g) == 0;
  return ok_vtu && ok_p ? GLH_OK : GLH_ERR_ARG;
}
