// Instantiation of the fast evaluation path for (float, kWT4S).
#include "ofd_fast_impl.h"

namespace ofd {
template int fast_eval_k<float, kWT4S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
}  // namespace ofd
