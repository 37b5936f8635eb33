// Learned preconditioner (message-passing network on the mesh graph): placeholder until
// the network kernels are in.
#include "rmx_internal.h"

namespace rmx {

bool mpnn_ready(Ctx& c) { (void)c; return false; }

void mpnn_apply(Ctx& c, const double* r, double* z)
{
  (void)c; (void)r; (void)z;
  fail(RMX_ERR_ARG, "learned preconditioner not loaded");
}

}  // namespace rmx
