// Opaque handle types of the C ABI (include/glhip.h) over the host mesh
// layer, shared by the translation units that implement the ABI.
#pragma once
#include "amr.hpp"
#include "mesh.hpp"

struct glh_mesh {
  glh::Mesh m;
};

struct glh_forest {
  glh::Forest f;
  int deg = 1;  // element degree of the meshes built from it (glh_forest_set_degree)
};
