// Internal accessors shared by the engine's translation units.
#pragma once
#include <string>

#include "kernels.h"

namespace ice {
const DevSpace& space_dev(const ice_space* s);
const ice_space* disc_space(const ice_disc* d);
int disc_neq(const ice_disc* d);
void set_error(const std::string& m);
}  // namespace ice
