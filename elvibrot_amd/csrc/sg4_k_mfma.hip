// Split-precision MFMA path (SG4_PREC_F32X2): the term kernel.
#include "sg4_mfma.h"

namespace sg4 {

__global__ __launch_bounds__(NT_MFMA) void k_hpsi_mfma(MArgs a) { mfma_items(a); }

}  // namespace sg4
