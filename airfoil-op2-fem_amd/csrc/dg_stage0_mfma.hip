// The stage-0 instantiations of the stage kernel (dg_stage_kernel.h),
// compiled with LLVM's iterative ILP scheduler (-amdgpu-sched-strategy=
// iterative-ilp, the Makefile's rule for this file): on the 724x724 vortex
// it takes the stage-0 kernel 349-353 -> 340-345 us (three alternating runs)
// at the same 167 VGPRs, while it made the stage-1 kernel no faster and the
// stage-2 kernel 3% slower, so those keep the default scheduler (launchers
// in dg_stage_mfma.hip).
#include "dg_stage_kernel.h"

namespace dg {

#define DG_DEFINE(N, CG) template DG_STAGE0_KERNEL(N, CG)
DG_STAGE0_INSTANCES(DG_DEFINE)
#undef DG_DEFINE

}  // namespace dg
