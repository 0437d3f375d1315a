// Split type-1 path, FP32 contractions: the term kernel.  It fits 168 VGPRs
// and may run 768 threads (three waves per SIMD at one workgroup per CU).
#include "sg4_split.h"

namespace sg4 {

__global__ __launch_bounds__(NT32) void k_hpsi_t1s_f32(KArgs a, KMats<float> km) { split_items(a, km); }

}  // namespace sg4
