// kern_table.hip -- converts a caller-supplied generator table in the reference's layout
// (src/precompute.rs:12-69) into this engine's odd-multiple table, with the checks of table.hpp.
// One workgroup of 128 lanes: lanes 0..15 check entry j, lane 0 shares entry 0 through LDS, then
// lane j writes odd multiple 2j + 1 (for Ed25519 also (2j + 1) [2^136] P0 in the second half).
// flags[0] = 0 for a table that passes (else TAB_BAD_* bits).
#include "kernels_impl.hpp"
#include "table.hpp"

namespace sigops {

template <class C>
__global__ __launch_bounds__(128) void k_table_secp(const u32* __restrict__ limbs, u32 lbits,
                                                    uint4* __restrict__ gtab_out, u32* __restrict__ flags) {
  typedef typename C::Fp F;
  __shared__ u32 s_bad;
  __shared__ u32 s_p0[2 * NL];
  const u32 j = threadIdx.x;
  const u32 nl = 256 / lbits + 1;
  if (j == 0) s_bad = 0;
  __syncthreads();
  fe x = fe_zero(), y = fe_zero();
  u32 bad = 0;
  if (j < 16) {
    tab_secp_check_entry<C>(j, limbs, nl, lbits, x, y, bad);
    if (j == 0) {
      const fe xc = fcanon<F>(x), yc = fcanon<F>(y);
#pragma unroll
      for (int k = 0; k < NL; k++) {
        s_p0[k] = xc.v[k];
        s_p0[NL + k] = yc.v[k];
      }
    }
  }
  __syncthreads();
  u32 p0[2 * NL], w[20];
#pragma unroll
  for (int k = 0; k < 2 * NL; k++) p0[k] = s_p0[k];
  tab_secp_build_entry<C>(j, p0, x, y, bad, w);
#pragma unroll
  for (int c = 0; c < 5; c++) gtab_out[j * 5 + c] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
  if (bad) atomicOr(&s_bad, bad);
  __syncthreads();
  if (j == 0) flags[0] = s_bad;
}

__global__ __launch_bounds__(128) void k_table_ed(const u32* __restrict__ limbs, u32 lbits,
                                                  uint4* __restrict__ btab_out, u32* __restrict__ flags) {
  typedef ModPed F;
  __shared__ u32 s_bad;
  __shared__ u32 s_p0[2 * NL];
  const u32 j = threadIdx.x;
  const u32 nl = 256 / lbits + 1;
  if (j == 0) s_bad = 0;
  __syncthreads();
  fe x = fe_zero(), y = fe_zero();
  u32 bad = 0;
  if (j < 16) {
    tab_ed_check_entry(j, limbs, nl, lbits, x, y, bad);
    if (j == 0) {
      const fe xc = fcanon<F>(x), yc = fcanon<F>(y);
#pragma unroll
      for (int k = 0; k < NL; k++) {
        s_p0[k] = xc.v[k];
        s_p0[NL + k] = yc.v[k];
      }
    }
  }
  __syncthreads();
  u32 p0[2 * NL], w[28];
#pragma unroll
  for (int k = 0; k < 2 * NL; k++) p0[k] = s_p0[k];
  tab_ed_build_entry(j, p0, x, y, bad, w);
#pragma unroll
  for (int c = 0; c < 7; c++) btab_out[j * 7 + c] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
  // the high-half table of the lattice-reduced check: (2j+1) [2^136] P0
  tab_ed_build_hi_entry(j, p0, w);
#pragma unroll
  for (int c = 0; c < 7; c++)
    btab_out[(BTAB_ENTRIES + j) * 7 + c] = make_uint4(w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]);
  if (bad) atomicOr(&s_bad, bad);
  __syncthreads();
  if (j == 0) flags[0] = s_bad;
}

// limbs: device copy of the 16 x coords x (256 / lbits + 1) words; lbits in 11..15 (the caller checks)
hipError_t launch_table_secp(int curve, const u32* limbs, u32 lbits, void* gtab_out, u32* flags, hipStream_t s) {
  if (curve == 0)
    k_table_secp<CurveK1><<<1, 128, 0, s>>>(limbs, lbits, (uint4*)gtab_out, flags);
  else
    k_table_secp<CurveR1><<<1, 128, 0, s>>>(limbs, lbits, (uint4*)gtab_out, flags);
  return hipGetLastError();
}

hipError_t launch_table_ed(const u32* limbs, u32 lbits, void* btab_out, u32* flags, hipStream_t s) {
  k_table_ed<<<1, 128, 0, s>>>(limbs, lbits, (uint4*)btab_out, flags);
  return hipGetLastError();
}

}  // namespace sigops
