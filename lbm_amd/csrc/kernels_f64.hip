// Kernels and launch wrappers for the F64 storage codec (both lattices).
#include "launch.h"

namespace lbm {
const Ops* ops_f64(int q) { return q == 19 ? Launch<19, CodecF64>::ops() : Launch<27, CodecF64>::ops(); }
}  // namespace lbm
