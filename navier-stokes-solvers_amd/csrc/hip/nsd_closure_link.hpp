// nsd_closure_link.hpp -- what the solver's step uses of the closure network
// beyond its C API: the network's bf16 input buffer, which the inverse row
// transform fills directly, and a forward pass from it.
#pragma once

#include "ns_closure.h"

namespace nsd {

// the network's input buffer (bf16, [batch][n][n]) and the scale the
// vorticity is multiplied by when written into it
unsigned short* closure_input(NsdClosure* c, float* in_scale);
// y = out_scale net(input buffer) for `batch` fields, on `stream`
int closure_eval_staged(NsdClosure* c, float* y, int batch, void* stream);

}  // namespace nsd
