// Kernels and launch wrappers for the F16S storage codec (both lattices).
#include "launch.h"

namespace lbm {
const Ops* ops_f16s(int q) { return q == 19 ? Launch<19, CodecF16S>::ops() : Launch<27, CodecF16S>::ops(); }
}  // namespace lbm
