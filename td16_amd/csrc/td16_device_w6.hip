// The persistent search built for six waves per SIMD (80 VGPRs; the body is
// in td16_search.h).  This file is compiled with LLVM's machine copy
// propagation switched off (-mllvm -disable-copyprop, csrc/Makefile): at a
// tight register budget the pass can move a spill reload into an odd-aligned
// register tuple, which gfx950 rejects ("Subtarget requires even aligned
// vector registers"), depending on unrelated changes elsewhere in the body.
#include "td16_search.h"

namespace td16 {

// Six waves per SIMD (24 per CU): the per-request paths spill a few
// registers to scratch, which costs request-bound batches (the mix of
// 100,000 requests runs best at 20 waves per CU) while node-bound ones gain
// from the sixth wave.  Chosen per launch by the wavefronts-per-CU knob.
__global__ void __launch_bounds__(kBlockThreads, 6)
td16_search_kernel_w6(Params P) {
  __shared__ BlockLds lds;
  search_blocks<false>(P, lds);
}

// Its counted build (Control::candidates), for measurement launches.
__global__ void __launch_bounds__(kBlockThreads, 6)
td16_search_kernel_w6_counted(Params P) {
  __shared__ BlockLds lds;
  search_blocks<true>(P, lds);
}

}  // namespace td16
