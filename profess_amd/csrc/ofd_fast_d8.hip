// Instantiation of the fast evaluation path for (double, kWGC8).
#include "ofd_fast_impl.h"

namespace ofd {
template int fast_eval_k<double, kWGC8>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
}  // namespace ofd
