// Instantiation of the fast evaluation path for (float, kWGC8).
#include "ofd_fast_impl.h"

namespace ofd {
template int fast_eval_k<float, kWGC8>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
}  // namespace ofd
