// Type-1 grid path with FP32 contractions.
#include "sg4_grid.h"

namespace sg4 {

__global__ __launch_bounds__(NT) void k_hpsi_t1_f32(KArgs a, KMats<float> km) {
  extern __shared__ double lds_d[];
  float* lds = (float*)lds_d;
  for (int w = blockIdx.x; w < a.nitems; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();
    const Item it = a.items[w];
    if (it.term < 0) continue;
    if (it.big) term_t1<float, true>(a, km, it, lds);
    else if (a.terms[it.term].ngrp > 0) term_t1_grouped<float>(a, km, it, lds);
    else term_t1<float, false>(a, km, it, lds);
  }
}

}  // namespace sg4
