// ik_gather_3d2.hip -- field-gather instantiations for D = 3, output tensors
// 2 x 2, bond entries in double (one translation unit per form, compiled in
// parallel).
#include "ik_gather_launch.hpp"

namespace pgt {

void launch_gather_3d2(const LaunchCtx &c, unsigned fields, int cg) { launch_gather_c<3, 2, double>(c, fields, cg); }

}  // namespace pgt
