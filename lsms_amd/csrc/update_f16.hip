// update_f16.hip -- split-FP16 trailing update of the mixed-precision LU.
//
// The mixed path factors an FP32 copy of the KKR matrix (tau_lu.hip, T =
// float) and refines against FP64.  Its dominant kernel is the paired-step
// trailing update C -= L_a U_a + L_b U_b.  Here the products run on the FP16
// matrix cores (v_mfma_f32_16x16x32_f16, FP32 accumulation): every FP32
// panel entry was split by the panel kernel into x 2^-e = hi + lo (FP16 each,
// one power-of-2 scale e per 64x64 tile), and each real product is
//   x y ~ (hx hy + hx ly + lx hy) 2^(ex + ey)
// (the lo*lo term, 2^-22 relative, is dropped), 22 significant bits against
// FP32's 24.  The refinement against the FP64 matrix recovers FP64 accuracy
// exactly as for the FP32 update; its contraction is set by the conditioning
// of M, not by this update.
//
// Two waves per 64x64 C tile (64 rows x 32 columns each), 2 tiles per
// workgroup.  Complex products in 4M
// form (Cr += (-Ar) Br + Ai Bi, Ci += (-Ar) Bi + (-Ai) Br) with two FP32
// accumulator sets in AGPRs.  The MFMA computes D = C^T = U^T L^T: the
// a-operand is a U plane (lane l: column l%16, k = 8(l/16)..+7, contiguous),
// the b-operand an L plane (lane l: row l%16, same k), and accumulator
// register r of lane l is C(row l%16, column 4(l/16) + r) (layout measured
// on gfx950), so C moves in 128 B runs.  C is scaled by the tile exponents
// on entry, between the two steps and on exit (exact powers of two).
//
// Built without -amdgpu-mfma-vgpr-form: the 64 accumulators per lane live
// in AGPRs.
#include "lu_common.hpp"

namespace lsmsx {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef GCx<float> GF2;
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const h8 GH8;
#else
typedef const h8 GH8;
#endif

__device__ __forceinline__ f4 mfma_h(h8 a, h8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ h8 neg8(h8 v) {
  return -v;
}

// acc (4 x 2 blocks of 16x16: bi row block, bj column block of the wave's
// 64x32 part) += product of the split L tile planes Lp (all 64 rows) and the
// U tile planes Up (the wave's 32 columns) over one 64-wide step.
__device__ __forceinline__ void split_step(GH8 *__restrict__ Lp, GH8 *__restrict__ Up, f4 (&cr)[4][2],
                                           f4 (&ci)[4][2]) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  constexpr int P = kTile * kTile / 8;  // plane stride in h8 units
#pragma unroll
  for (int kc = 0; kc < 2; kc++) {  // two k chunks of 32
    h8 ur_h[2], ur_l[2], ui_h[2], ui_l[2];
#pragma unroll
    for (int bj = 0; bj < 2; bj++) {
      const int o = ((16 * bj + l16) * kTile + 32 * kc + 8 * g) / 8;
      ur_h[bj] = Up[o];
      ur_l[bj] = Up[o + P];
      ui_h[bj] = Up[o + 2 * P];
      ui_l[bj] = Up[o + 3 * P];
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
      const int o = ((16 * bi + l16) * kTile + 32 * kc + 8 * g) / 8;
      const h8 nar_h = Lp[o], nar_l = Lp[o + P], ai_h = Lp[o + 2 * P], ai_l = Lp[o + 3 * P];
      const h8 nai_h = neg8(ai_h), nai_l = neg8(ai_l);
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        f4 r = cr[bi][bj], m = ci[bi][bj];
        r = mfma_h(ur_h[bj], nar_h, r);
        r = mfma_h(ur_l[bj], nar_h, r);
        r = mfma_h(ur_h[bj], nar_l, r);
        r = mfma_h(ui_h[bj], ai_h, r);
        r = mfma_h(ui_l[bj], ai_h, r);
        r = mfma_h(ui_h[bj], ai_l, r);
        m = mfma_h(ui_h[bj], nar_h, m);
        m = mfma_h(ui_l[bj], nar_h, m);
        m = mfma_h(ui_h[bj], nar_l, m);
        m = mfma_h(ur_h[bj], nai_h, m);
        m = mfma_h(ur_l[bj], nai_h, m);
        m = mfma_h(ur_h[bj], nai_l, m);
        cr[bi][bj] = r;
        ci[bi][bj] = m;
      }
    }
  }
}

// Paired-step trailing update (see k_lu_update2 in tau_lu.hip for the tile
// ranges) on the split-FP16 planes of steps a and b.
__global__ __launch_bounds__(256) void k_lu_update2_h(const SolveDesc *__restrict__ descs, int step,
                                                      int xcd_swizzle) {
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcd_swizzle) xcd_remap(bx, by);
  const SolveDesc &d = descs[by];
  const int a = step - d.step0, b = a + 1;
  if (b < 0 || b > d.nt - 2) return;
  if (d.conv && *d.conv) return;
  const int nrest = d.nt - 1 - b;
  const int w = threadIdx.x >> 6;
  const int x = 2 * bx + (w >> 1);
  if (x >= nrest * nrest) return;  // per wave: no barriers below
  const int ti = b + 1 + x % nrest, tj = b + 1 + x / nrest;
  const int wc = 32 * (w & 1);      // the wave's 32 columns of the tile
  const int nt = d.nt;
  const int lane = threadIdx.x & 63, l16 = lane & 15, g = lane >> 4;
  const long ld = d.ld;
  GF2 *C = (GF2 *)d.A + (long)(kTile * ti) + ld * (kTile * tj + wc);
  const _Float16 *hp = (const _Float16 *)d.hp;
  const int ns = a >= 0 ? 2 : 1;
  const int s_first = a >= 0 ? a : b;
  auto tile_exp = [&](int s) {  // exponent of L(ti, s) U(s, tj)
    const int par = s & 1;
    return d.he[(par * 2 + 0) * nt + ti] + d.he[(par * 2 + 1) * nt + tj];
  };
  f4 cr[4][2], ci[4][2];
  int e = tile_exp(s_first);
  {
    const float sc = ldexpf(1.f, -e);
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const float2 c = C[(16 * bi + l16) + ld * (16 * bj + 4 * g + r)];
          cr[bi][bj][r] = c.x * sc;
          ci[bi][bj][r] = c.y * sc;
        }
  }
  for (int u = 0; u < ns; u++) {
    const int s = s_first + u;
    if (u > 0) {
      const int e2 = tile_exp(s);
      const float sc = ldexpf(1.f, e - e2);
#pragma unroll
      for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          cr[bi][bj] *= sc;
          ci[bi][bj] *= sc;
        }
      e = e2;
    }
    const int par = s & 1;
    GH8 *Lp = (GH8 *)(hp + (long)((par * 2 + 0) * nt + ti) * kPlaneHalves);
    GH8 *Up = (GH8 *)(hp + (long)((par * 2 + 1) * nt + tj) * kPlaneHalves + wc * kTile);
    split_step(Lp, Up, cr, ci);
  }
  const float sc = ldexpf(1.f, e);
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        C[(16 * bi + l16) + ld * (16 * bj + 4 * g + r)] =
            make_float2(cr[bi][bj][r] * sc, ci[bi][bj][r] * sc);
}

}  // namespace lsmsx
