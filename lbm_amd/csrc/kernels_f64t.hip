// Kernels and launch wrappers for FP64 storage with tangents (CodecTan, both lattices): the
// forward kernels run in forward-mode FP64 (the tangent run the adjoint is checked against).
#include "launch.h"

namespace lbm {
const Ops* ops_f64t(int q) { return q == 19 ? Launch<19, CodecTan>::ops() : Launch<27, CodecTan>::ops(); }
}  // namespace lbm
