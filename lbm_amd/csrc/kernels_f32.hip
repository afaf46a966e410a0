// Kernels and launch wrappers for the F32 storage codec (both lattices).
#include "launch.h"

namespace lbm {
const Ops* ops_f32(int q) { return q == 19 ? Launch<19, CodecF32>::ops() : Launch<27, CodecF32>::ops(); }
}  // namespace lbm
