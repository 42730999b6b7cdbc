// zsc_f32.hip -- FP32-compute instantiations of libzsc's kernels: FP32
// storage, FP32 storage with g as deviations, and FP32 compute over bf16 /
// fp16 deviation storage.
#include "zsc_adjoint.h"

namespace zsc {
template struct Ops<float>;
template struct Ops<bf16s>;
template struct Ops<f16s>;
template struct Ops<f32d>;
template struct DropIn<float>;
template struct Adj<float>;
}  // namespace zsc
