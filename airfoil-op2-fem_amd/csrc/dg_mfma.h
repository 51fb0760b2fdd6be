// FP64 matrix-core and direct-to-LDS helpers shared by the stage kernels and
// the reverse (adjoint) RHS kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace dg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// D = A B + C on one 16x16x4 FP64 block: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[l / 16 + 4 r][l % 16], r = 0..3.
__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Copies nch (>= 1) 16-byte chunks from global src to LDS dst with
// direct-to-LDS loads (global_load_lds_dwordx4: 16 bytes per lane, one
// contiguous KiB per instruction of the wave, no registers holding the
// data).  dst is the same for all lanes of the wave (it goes to M0) and
// holds 64 ceil(MAXCH / 64) chunks: every instruction is issued with all
// lanes, those past nch re-reading the last chunk into their own slots, so
// the number of loads in flight is the same on every path and the
// compiler's counter waits for older loads stay exact.  The data is in LDS
// once the wave's vector-memory counter has drained (s_waitcnt vmcnt(0)).
template <int MAXCH>
__device__ __forceinline__ void copy_to_lds(const double *src, double *dst, int nch, int lane) {
#pragma unroll
  for (int r = 0; r < (MAXCH + 63) / 64; r++) {
    const int ch = min(64 * r + lane, nch - 1);
    __builtin_amdgcn_global_load_lds((glb_void_t *)(src + 2 * ch), (lds_void_t *)(dst + 128 * r), 16, 0, 0);
  }
}

// The same copy issued from inline assembly, for a kernel whose own LDS
// accesses follow the copy before its data is needed: the compiler does not
// see these loads, so it neither waits for them before those LDS accesses
// (it cannot tell that they touch other LDS memory, and drained every load
// in flight there) nor counts them in its waits for later loads (the copy
// is older than those, and loads complete in order, so the waits it
// computes for them stay sufficient).  The caller waits for the copy
// itself (s_waitcnt vmcnt(0)) before reading dst, and makes sure no LDS
// access to dst is pending when it issues the copy.  M0 takes dst's LDS
// address (M0 is reserved: the compiler writes it right before each of its
// own uses, so it cannot be clobbered under it); the s_nop covers the wait
// state between writing M0 and the direct-to-LDS load that reads it.
template <int MAXCH>
__device__ __forceinline__ void copy_to_lds_unseen(const double *src, double *dst, int nch, int lane) {
#pragma unroll
  for (int r = 0; r < (MAXCH + 63) / 64; r++) {
    const int ch = min(64 * r + lane, nch - 1);
    const unsigned lds = (unsigned)(size_t)(lds_void_t *)(dst + 128 * r);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(lds), "v"(src + 2 * ch)
                 : "memory");
  }
}

// Waits for the wave's own LDS accesses in flight; with the memory clobber,
// also a compiler barrier for LDS accesses.  Between phases of a kernel
// whose waves each own their LDS, it replaces __syncthreads(), whose
// workgroup release fence would also wait for the wave's global stores.
__device__ __forceinline__ void lds_wait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Doubles of LDS that copy_to_lds<MAXCH> writes (whole KiB per instruction).
template <int MAXCH>
constexpr int copy_to_lds_doubles() {
  return 128 * ((MAXCH + 63) / 64);
}

// The lane index passed through an empty asm statement inside the tile loop.
// A lane's (cell, node) items are the same on every tile, so the compiler
// hoists their decomposition and addresses out of the loop; held for the
// kernel's lifetime they overflow the register budget and are spilled to
// scratch, and every reload then waits for all loads in flight.  Derived
// from this value they are recomputed per tile (a few integer operations).
// Measured at N=3: the face-map code prefetch of every stage and the Roe
// and Euler-flux items of the stage-2 kernel (the one with the CFL
// epilogue) gain from it; the other stages' items keep their registers.
// At N=4 every stage spills without it (the items of all stages go through it).
__device__ __forceinline__ int lane_in_loop(int l) {
  asm volatile("" : "+v"(l));
  return l;
}

}  // namespace dg
