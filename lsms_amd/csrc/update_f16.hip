// update_f16.hip -- split-FP16 trailing update of the mixed-precision LU.
//
// The mixed path factors an FP32 copy of the KKR matrix (tau_lu.hip, T =
// float) and refines against FP64.  Its dominant kernel is the trailing
// update of a group of elimination steps (eight by default), C -= sum_k
// L_k U_k (K = 512 for eight: C is read and written once per group).  Here
// the products run on the FP16 matrix cores (v_mfma_f32_16x16x32_f16, FP32 accumulation): every FP32
// panel entry was split by the panel kernel into x 2^-e = hi + lo (FP16 each,
// one power-of-2 scale e per 64x64 tile), and each real product is
//   x y ~ (hx hy + hx ly + lx hy) 2^(ex + ey)
// (the lo*lo term, 2^-22 relative, is dropped), 22 significant bits against
// FP32's 24.  The refinement against the FP64 matrix recovers FP64 accuracy
// exactly as for the FP32 update; its contraction is set by the conditioning
// of M, not by this update.
//
// A workgroup of the trailing update owns a 128 x 128 region (C tiles ti,
// ti+1 x tj, tj+1); each of its eight waves computes a 64 x 32 part of one
// tile.  Per 32-wide k chunk the split planes of the two L row tiles and of
// the two U column tiles are staged in LDS (64 KiB, two stages), so every
// plane element is fetched from L2 once per workgroup and read by two or four
// waves; the loads run two chunks ahead of the MFMAs, so the L2 latency hides
// behind them.  (The cross update of a diagonal tile uses one 64 x 64 region
// of two waves and one stage; the other tiles of a step's column and row
// take their cross update inside the fused panel, k_lu_panel_h in
// tau_lu.hip.)  The four 16-byte k groups of a 64-byte LDS row are XOR-swizzled by the row, so the
// lanes of a 128-bit read hit distinct banks.  Complex products in 4M form
// (Cr += (-Ar) Br + Ai Bi, Ci += (-Ar) Bi + (-Ai) Br) with two FP32
// accumulator sets.  The MFMA computes D = C^T = U^T L^T: the a-operand is
// a U plane (lane l: column l%16, k = 8(l/16)..+7), the b-operand an L plane
// (lane l: row l%16, same k), and accumulator register r of lane l is
// C(row l%16, column 4(l/16) + r) (layout measured on gfx950), so C moves
// in 128 B runs.  The accumulators start from zero and are scaled by the
// tile exponents between the steps (exact powers of two); C is loaded during
// the last k chunk, into the registers the plane prefetch no longer needs,
// and added on exit, so no MFMA waits for C's HBM latency.
//
// Built without -amdgpu-mfma-vgpr-form, so that the compiler may keep the 64
// accumulators per lane in AGPRs where the VGPRs run short.
#include "lu_common.hpp"

namespace lsmsx {

typedef GCx<float> GF2;
// C moves with non-temporal loads and stores: it is touched once per group,
// and keeping it out of L2 leaves room for the split planes, which are re-read.
typedef float fv2 __attribute__((ext_vector_type(2)));

// C tiles (r0 + wt, c0 + ct), wt < nrow <= RT, ct < ncol <= NC (a 64 RT x
// 64 NC region of 128 RT NC threads; tiles past nrow / ncol are skipped),
// -= sum over the steps s_first .. s_first+ns-1 of L(., s) U(s, .), on the
// split-FP16 planes of those steps.  Waves: wt = row tile, ct = column tile,
// wc = the 32-column half.  Per 32-wide k chunk the planes of the RT L row
// tiles and the NC U column tiles are staged in LDS, (RT + NC) x 16 KiB, so a
// chunk's staging serves RT x NC output tiles.  DB = 1: one LDS stage, the next
// chunk's loads in flight during this chunk's MFMAs, two barriers per chunk.
// DB = 2: two LDS stages and two register stages, loads issued two chunks
// ahead, one barrier per chunk.
template <int RT, int NC, int DB>
__device__ __forceinline__ void region_h(const SolveDesc &d, int r0, int nrow, int c0, int ncol,
                                         int s_first, int ns, h8 (&Ls)[DB][RT][4][kTile][4],
                                         h8 (&Us)[DB][NC][4][kTile][4]) {
  constexpr int NT = 128 * RT * NC;            // threads
  constexpr int NQ = (RT + NC) * 1024 / NT;    // 8-half loads per thread per chunk
  constexpr int NQL = RT * 1024 / NT;          // of which L planes
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, l16 = lane & 15, g = lane >> 4;
  const int wt = w / (2 * NC), ct = (w >> 1) % NC, wc = 32 * (w & 1);
  const int ti = r0 + wt, tj = c0 + ct;
  const bool valid = wt < nrow && ct < ncol;
  const int nt = d.nt;
  const long ld = tile_ld(d);
  const _Float16 *hp = (const _Float16 *)d.hp;
  const int gmask = d.group - 1;  // plane slot of step s: s & gmask
  // The tile exponents of every step of the range (ns <= 16), read once into
  // the wave's LDS row: a global load per step inside the chunk loop would
  // wait on every load in flight, the plane prefetches included.
  __shared__ int exs[NT / 64][16];
  if (lane < ns) {
    const int par = (s_first + lane) & gmask;
    exs[w][lane] = valid ? d.he[(par * 2 + 0) * nt + ti] + d.he[(par * 2 + 1) * nt + tj] : 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  auto tile_exp = [&](int s) { return exs[w][s - s_first]; };
  GF2 *C = tile_ptr<float>(d, valid ? ti : r0, valid ? tj : c0) + ld * wc;
  // Chunk ch = (step u = ch / 2, k half ch % 2): NQ loads of 8 halves per
  // thread (q < NQL: the two L row tiles, then the U column tiles), all
  // issued together.  Tiles past the end are clamped to a real tile (their
  // waves skip the MFMAs).
  auto chunk_src = [&](int ch, int q) -> GH8 * {
    const int s = s_first + (ch >> 1), kh = ch & 1, par = s & gmask;
    const int e8 = tid + NT * q;
    if (q < NQL) {
      const int tt = e8 >> 10, plane = (e8 >> 8) & 3, rr = (e8 >> 2) & 63, grp = e8 & 3;
      const int t = min(r0 + tt, nt - 1);
      return (GH8 *)(hp + (long)((par * 2 + 0) * nt + t) * kPlaneHalves) +
             ((plane * kTile + rr) * kTile + 32 * kh + 8 * grp) / 8;
    }
    const int f = e8 - RT * 1024, cu = f >> 10, plane = (f >> 8) & 3, rr = (f >> 2) & 63, grp = f & 3;
    const int t = min(c0 + cu, nt - 1);
    return (GH8 *)(hp + (long)((par * 2 + 1) * nt + t) * kPlaneHalves) +
           ((plane * kTile + rr) * kTile + 32 * kh + 8 * grp) / 8;
  };
  auto chunk_dst = [&](int buf, int q) -> h8 * {
    const int e8 = tid + NT * q;
    if (q < NQL) {
      const int tt = e8 >> 10, plane = (e8 >> 8) & 3, rr = (e8 >> 2) & 63, grp = e8 & 3;
      return &Ls[buf][tt][plane][rr][swz_group(rr, grp)];
    }
    const int f = e8 - RT * 1024, cu = f >> 10, plane = (f >> 8) & 3, rr = (f >> 2) & 63, grp = f & 3;
    return &Us[buf][cu][plane][rr][swz_group(rr, grp)];
  };
  f4 cr[4][2], ci[4][2];
  int e = valid ? tile_exp(s_first) : 0;
  // rescale C to the tile exponents of the step chunk ch starts
  auto rescale = [&](int ch) __attribute__((always_inline)) {
    if (ch > 0 && !(ch & 1) && valid) {
      const int e2 = tile_exp(s_first + (ch >> 1));
      if (e2 == e) return;  // a factor of exactly 1
      const float sc1 = ldexpf(1.f, e - e2);
#pragma unroll
      for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          cr[bi][bj] *= sc1;
          ci[bi][bj] *= sc1;
        }
      e = e2;
    }
  };
  auto compute = [&](int buf) __attribute__((always_inline)) {
    if (!valid) return;
    // the wave that reaches its MFMAs first keeps the issue slots while the
    // other wave of its SIMD is still staging, so the two drift apart; the
    // scheduler interleaves the LDS operand reads with the MFMAs
    __builtin_amdgcn_s_setprio(2);
    __builtin_amdgcn_iglp_opt(0);
    // U operands of the wave's two 16-column blocks, with -Re(U) formed once
    // here for the -Im(L) Re(U) term (cheaper than negating Im(L) per row
    // block; Im(L) (-Re U) = -(Im(L) Re(U)) exactly)
    h8 ur_h[2], ur_l[2], ui_h[2], ui_l[2], nur_h[2], nur_l[2];
#pragma unroll
    for (int bj = 0; bj < 2; bj++) {
      const int col = wc + 16 * bj + l16, gg = swz_group(col, g);
      ur_h[bj] = Us[buf][ct][0][col][gg];
      ur_l[bj] = Us[buf][ct][1][col][gg];
      ui_h[bj] = Us[buf][ct][2][col][gg];
      ui_l[bj] = Us[buf][ct][3][col][gg];
      nur_h[bj] = neg8(ur_h[bj]);
      nur_l[bj] = neg8(ur_l[bj]);
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
      const int row = 16 * bi + l16, gg = swz_group(row, g);
      const h8 nar_h = Ls[buf][wt][0][row][gg], nar_l = Ls[buf][wt][1][row][gg];
      const h8 ai_h = Ls[buf][wt][2][row][gg], ai_l = Ls[buf][wt][3][row][gg];
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
        m = mfma_h(nur_h[bj], ai_h, m);
        m = mfma_h(nur_l[bj], ai_h, m);
        m = mfma_h(nur_h[bj], ai_l, m);
        cr[bi][bj] = r;
        ci[bi][bj] = m;
      }
    }
    __builtin_amdgcn_s_setprio(0);
  };
  // C is added after the products: the accumulators start from zero and C's
  // loads are issued during the last chunk, so no MFMA waits for them
  fv2 cv[4][2][4];
  auto issue_c = [&](int half) __attribute__((always_inline)) {
    if (!valid) return;
#pragma unroll
    for (int bi = 2 * half; bi < 2 * half + 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          cv[bi][bj][r] = __builtin_nontemporal_load((const fv2 *)(C + (16 * bi + l16) + ld * (16 * bj + 4 * g + r)));
  };
#pragma unroll
  for (int bi = 0; bi < 4; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) {
      cr[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
      ci[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
    }
  const int nch = 2 * ns;
  if constexpr (DB == 1) {
    h8 pv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) pv[q] = *chunk_src(0, q);
#pragma unroll 1
    for (int ch = 0; ch < nch - 1; ch++) {
      rescale(ch);
      __syncthreads();  // the previous chunk's readers are done
#pragma unroll
      for (int q = 0; q < NQ; q++) *chunk_dst(0, q) = pv[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; q++) pv[q] = *chunk_src(ch + 1, q);
      compute(0);
    }
    // the last chunk, with C's loads in the prefetch registers' place
    rescale(nch - 1);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) *chunk_dst(0, q) = pv[q];
    __syncthreads();
    issue_c(0);
    issue_c(1);
    compute(0);
  } else {
    // chunk c is staged in LDS buffer c & 1 and loaded into pa (even c) or pb
    // (odd c); nch is even
    h8 pa[NQ], pb[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) pa[q] = *chunk_src(0, q);
#pragma unroll
    for (int q = 0; q < NQ; q++) pb[q] = *chunk_src(1, q);
#pragma unroll
    for (int q = 0; q < NQ; q++) *chunk_dst(0, q) = pa[q];
    __syncthreads();
    if (nch > 2) {
#pragma unroll
      for (int q = 0; q < NQ; q++) pa[q] = *chunk_src(2, q);
    }
#pragma unroll 1
    for (int ch = 0; ch < nch - 2; ch += 2) {
      rescale(ch);
      compute(0);
#pragma unroll
      for (int q = 0; q < NQ; q++) *chunk_dst(1, q) = pb[q];
      if (ch + 3 < nch) {
#pragma unroll
        for (int q = 0; q < NQ; q++) pb[q] = *chunk_src(ch + 3, q);
      }
      __syncthreads();  // buffer 1 staged; every wave is done with buffer 0
      compute(1);
#pragma unroll
      for (int q = 0; q < NQ; q++) *chunk_dst(0, q) = pa[q];
      if (ch + 4 < nch) {
#pragma unroll
        for (int q = 0; q < NQ; q++) pa[q] = *chunk_src(ch + 4, q);
      }
      __syncthreads();  // buffer 0 staged; every wave is done with buffer 1
    }
    // the last two chunks, with C's loads in the prefetch registers' place
    rescale(nch - 2);
    issue_c(0);
    compute(0);
#pragma unroll
    for (int q = 0; q < NQ; q++) *chunk_dst(1, q) = pb[q];
    issue_c(1);
    __syncthreads();
    compute(1);
  }
  if (valid) {
    const float sc2 = ldexpf(1.f, e);
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          __builtin_nontemporal_store(fv2{cr[bi][bj][r] * sc2 + cv[bi][bj][r].x,
                                          ci[bi][bj][r] * sc2 + cv[bi][bj][r].y},
                                      (fv2 *)(C + (16 * bi + l16) + ld * (16 * bj + 4 * g + r)));
  }
}

// Trailing update of the step group starting at global step `step` (see
// k_lu_trail in tau_lu.hip for the tile ranges) on the split-FP16 planes of
// the group's steps: 128 x 128 regions (two row tiles x two column tiles) of
// 512 threads.
__global__ __launch_bounds__(512) void k_lu_trail_h(const SolveDesc *__restrict__ descs, int step,
                                                    int xcd_swizzle) {
  __shared__ h8 Ls[2][2][4][kTile][4];
  __shared__ h8 Us[2][2][4][kTile][4];
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcd_swizzle) xcd_remap(bx, by);
  const SolveDesc &d = descs[by];
  const int g0 = step - d.step0;
  if (g0 < 0) return;
  const int ns = min(d.group, d.nt - 1 - g0);
  if (ns <= 0) return;
  if (d.conv && *d.conv) return;
  const int b = g0 + ns - 1;  // last step of the group
  const int nrest = d.nt - 1 - b, np = (nrest + 1) / 2;
  if (bx >= np * np) return;
  // Regions in super-blocks of kBlk x kBlk tile pairs, block by block
  // (columns of blocks within a block row): the workgroups resident on an
  // XCD at one time share the planes of about 16 row and 16 column tiles,
  // which stay in its L2.
  constexpr int kBlk = 8;
  const int br = bx / (kBlk * np), hb = min(kBlk, np - br * kBlk);
  const int off = bx - br * kBlk * np, bc = off / (hb * kBlk), off2 = off - bc * hb * kBlk;
  const int sr = br * kBlk + off2 % hb, sc = bc * kBlk + off2 / hb;
  region_h<2, 2, 2>(d, b + 1 + 2 * sr, min(2, nrest - 2 * sr), b + 1 + 2 * sc,
                 min(2, nrest - 2 * sc), g0, ns, Ls, Us);
}

// Cross update of the diagonal tile of step t = g + u of the group starting
// at global step `step` (g local), before the tile is factored: the steps
// g .. t-1 on their split planes, one 64 x 64 region of 128 threads per
// solve.  The other tiles of column and row t take the same update inside
// the fused panel (k_lu_panel_h).
__global__ __launch_bounds__(128) void k_lu_cross_diag_h(const SolveDesc *__restrict__ descs,
                                                         int step, int u) {
  __shared__ h8 Ls[1][1][4][kTile][4];
  __shared__ h8 Us[1][1][4][kTile][4];
  const SolveDesc &d = descs[blockIdx.x];
  const int g0 = step - d.step0, t = g0 + u;
  if (g0 < 0 || u <= 0 || t > d.nt - 2) return;
  if (d.conv && *d.conv) return;
  region_h<1, 1, 1>(d, t, 1, t, 1, g0, u, Ls, Us);
}

}  // namespace lsmsx
