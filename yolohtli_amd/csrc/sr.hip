// sr.hip -- symmetry-reduction path (placeholder until the kernels land).
#include "plan.h"

struct yh_sr {
  int unused;
};

namespace yh {
void sr_free(yh_sr* sr) { delete sr; }
}  // namespace yh
