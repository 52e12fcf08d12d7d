// The persistent batched kernel (placeholder until the lockstep kernel
// lands): not available yet.
#pragma once
#include "kx_plain.h"
#include "kx_util.h"

namespace kx {

template <class Net, bool ET>
inline int launch_persistent(CellsArgs a, int mixed, int chunk, int wpc,
                             hipStream_t s) {
  (void)a; (void)mixed; (void)chunk; (void)wpc; (void)s;
  return FLAG_UNSUPPORTED;
}

}  // namespace kx
