// Kernels and launch wrappers for the BF16S storage codec (both lattices).
#include "launch.h"

namespace lbm {
const Ops* ops_bf16s(int q) { return q == 19 ? Launch<19, CodecBF16S>::ops() : Launch<27, CodecBF16S>::ops(); }
}  // namespace lbm
