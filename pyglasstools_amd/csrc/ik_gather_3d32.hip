// ik_gather_3d32.hip -- field-gather instantiations for D = 3, output tensors
// 3 x 3, bond entries in float (one translation unit per form, compiled in
// parallel).
#include "ik_gather_launch.hpp"

namespace pgt {

void launch_gather_3d32(const LaunchCtx &c, unsigned fields, int cg) { launch_gather_c<3, 3, float>(c, fields, cg); }

}  // namespace pgt
