// Constants shared by the host mesh layer and the device code.
#pragma once
#include <cstdint>

namespace glh {

constexpr int NC = 18;  // real components of the order parameter (Re/Im of A_{mu i})
// node_mask flag of a hanging node (its value follows its constraint line)
constexpr uint32_t HANG_BIT = 1u << 31;

}  // namespace glh
