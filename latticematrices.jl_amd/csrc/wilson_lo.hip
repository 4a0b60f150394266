// wilson_lo.hip -- one low-precision stencil instantiation per object file
// (compiled once per LMX_LO_COMBO in {0, 1, 2} and LMX_LO_NR in {1, 2, 4},
// so the variants build in parallel): the (spinor, links) pair
// LMX_LO_COMBO_<n> at LMX_LO_NR right-hand sides per work item.
#include "wilson_kernels.h"

#if !defined(LMX_LO_NR) || !defined(LMX_LO_COMBO)
#error "build with -DLMX_LO_COMBO=0|1|2 -DLMX_LO_NR=1|2|4"
#endif

namespace lmx {

#define LMX_INST(In, Out, Lk, Cl, NR) template int launch_lo_nr<In, Out, Lk, Cl, NR>(lmx_context*, const LoArgs&, Lk, Cl, bool);
#if LMX_LO_COMBO == 0
LMX_LO_COMBO_0(LMX_INST, LMX_LO_NR)
#elif LMX_LO_COMBO == 1
LMX_LO_COMBO_1(LMX_INST, LMX_LO_NR)
#else
LMX_LO_COMBO_2(LMX_INST, LMX_LO_NR)
#endif
#undef LMX_INST
// the LDS-staged and RHS-loop (FP16 links) shapes, built with the
// one-RHS-per-work-item objects
#if LMX_LO_NR == 1
#define LMX_INST_LDS(In, Out, Lk, Cl, NR) template int launch_lo_lds<In, Out, Lk, Cl>(lmx_context*, const LoArgs&, Lk, Cl, bool);
#if LMX_LO_COMBO == 0
LMX_LO_COMBO_0(LMX_INST_LDS, 1)
#elif LMX_LO_COMBO == 1
LMX_LO_COMBO_1(LMX_INST_LDS, 1)
#else
LMX_LO_COMBO_2(LMX_INST_LDS, 1)
#endif
#undef LMX_INST_LDS
#endif
#if LMX_LO_NR == 1 && LMX_LO_COMBO == 1
template int launch_lo_loop<SpF32In, SpF32Out, F16Clover>(lmx_context*, const LoArgs&, LinkF16, F16Clover, bool);
#elif LMX_LO_NR == 1 && LMX_LO_COMBO == 2
template int launch_lo_loop<SpH16In, SpH16Out, F16Clover>(lmx_context*, const LoArgs&, LinkF16, F16Clover, bool);
#endif

}  // namespace lmx
