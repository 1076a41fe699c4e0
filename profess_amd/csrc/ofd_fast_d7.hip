// Instantiation of the fast evaluation path for (double, kWT4S).
#include "ofd_fast_impl.h"

namespace ofd {
template int fast_eval_k<double, kWT4S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
}  // namespace ofd
