// zsc_f64.hip -- FP64 instantiation of libzsc's kernels (parity mode).
// Built with -ffp-contract=off: every multiply and add is rounded on its own,
// as in an IEEE evaluation of the reference's expressions.
#include "zsc_adjoint.h"

namespace zsc {
template struct Ops<double>;
template struct DropIn<double>;
template struct Adj<double>;
// FP32 adjoint lattices over FP64 states (arithmetic in FP64)
template struct Adj<double, float>;
}  // namespace zsc
