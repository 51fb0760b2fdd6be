// Launchers of the stage kernel (dg_stage_kernel.h).  The stage-0
// instantiations are compiled in dg_stage0_mfma.hip with another
// instruction scheduler and only declared here.
#include "dg_stage_kernel.h"

namespace dg {

#define DG_DECLARE(N, CG) extern template DG_STAGE0_KERNEL(N, CG)
DG_STAGE0_INSTANCES(DG_DECLARE)
#undef DG_DECLARE

// Waves per workgroup actually used for order N: 1, 2 or 4, within 64 KB of
// LDS per workgroup.
template <int N, int EPI, bool CG>
static int mfma_wpb(int want) {
  int w = want >= 4 ? 4 : (want >= 2 ? 2 : 1);
  while (w > 1 && (size_t)w * mfma_lds_doubles<N, EPI, CG>() * sizeof(double) > 64 * 1024) w /= 2;
  return w;
}

template <int N, bool GATHER, int EPI, bool CG>
static hipError_t launch_t(const Launch &L, const StageArgs &a) {
  static int per_cu[5] = {-1, -1, -1, -1, -1};  // resident workgroups per CU, by waves per workgroup
  const int wpb = mfma_wpb<N, EPI, CG>(L.wpb);
  const size_t lds = (size_t)wpb * mfma_lds_doubles<N, EPI, CG>() * sizeof(double);
  if (per_cu[wpb] < 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rhs_mfma_kernel<N, GATHER, EPI, CG>, 64 * wpb, lds) !=
            hipSuccess ||
        nb < 1)
      nb = 8 / wpb > 0 ? 8 / wpb : 1;
    per_cu[wpb] = nb;
  }
  const int ntiles = a.ncases > 0 ? a.ntiles_hint : (a.ncells + 15) / 16;
  const int nblk = ntiles > 0 ? (ntiles + wpb - 1) / wpb : -1;
  int grid = L.cus > 0 ? L.cus * per_cu[wpb] : 2048 / wpb;
  if (L.max_blocks > 0 && grid > L.max_blocks) grid = L.max_blocks;
  if (nblk >= 0 && grid > nblk) grid = nblk;
  // Fewer workgroups with equal visits each (4,096 tiles on 2,048 waves
  // rather than 3,072) measured slower: more resident waves hide more
  // latency, so the grid stays at the occupancy limit.
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) return hipSuccess;
  hipLaunchKernelGGL((rhs_mfma_kernel<N, GATHER, EPI, CG>), dim3(grid), dim3(64 * wpb), lds, L.stream, a,
                     L.ops, (unsigned long long)L.fmask4, L.ph, (unsigned long long)L.facenodes, L.walk);
  return hipGetLastError();
}

template <int N, bool CG>
static hipError_t launch_g(const Launch &L, const StageArgs &a, int epi) {
  switch (epi) {
    case EPI_RHS: return launch_t<N, true, EPI_RHS, CG>(L, a);
    case EPI_STAGE0: return launch_t<N, true, EPI_STAGE0, CG>(L, a);
    case EPI_STAGE1: return launch_t<N, true, EPI_STAGE1, CG>(L, a);
    case EPI_STAGE2: return launch_t<N, true, EPI_STAGE2, CG>(L, a);
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t launch_n(const Launch &L, const StageArgs &a, bool gather, int epi) {
  if (!gather) return launch_t<N, false, EPI_RHS, false>(L, a);
  return a.cmet ? launch_g<N, true>(L, a, epi) : launch_g<N, false>(L, a, epi);
}

hipError_t launch_rhs_mfma(const Launch &L, const StageArgs &a, bool gather, int epi) {
  if (!gather && epi != EPI_RHS) return hipErrorInvalidValue;
  if (a.cmet && !gather) return hipErrorInvalidValue;
  switch (L.N) {
    case 1: return launch_n<1>(L, a, gather, epi);
    case 2: return launch_n<2>(L, a, gather, epi);
    case 3: return launch_n<3>(L, a, gather, epi);
    case 4: return launch_n<4>(L, a, gather, epi);
  }
  return hipErrorInvalidValue;
}

}  // namespace dg
