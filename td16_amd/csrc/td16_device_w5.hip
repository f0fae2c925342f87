// The persistent search built for five waves per SIMD (96 VGPRs; the body
// is in td16_search.h).  This file is compiled with LLVM's machine copy
// propagation switched off (-mllvm -disable-copyprop, csrc/Makefile): at this
// register budget the pass can move a 96-bit spill reload into an odd-aligned
// register triple, which gfx950 rejects ("Subtarget requires even aligned
// vector registers"), depending on unrelated changes elsewhere in the body.
#include "td16_search.h"

namespace td16 {

// Five waves per SIMD: the per-request paths spill some registers to
// scratch, which costs the request-bound mix ~8 %, while searches of tens of
// nodes per request gain from the fifth wave (the emitter-shaped stream: -6 %
// at 20 waves per CU).  Chosen per launch by the wavefronts-per-CU knob.
__global__ void __launch_bounds__(kBlockThreads, 5)
td16_search_kernel_w5(Params P) {
  __shared__ BlockLds lds;
  search_blocks<false>(P, lds);
}

// Its counted build (Control::candidates), for measurement launches.
__global__ void __launch_bounds__(kBlockThreads, 5)
td16_search_kernel_w5_counted(Params P) {
  __shared__ BlockLds lds;
  search_blocks<true>(P, lds);
}

}  // namespace td16
