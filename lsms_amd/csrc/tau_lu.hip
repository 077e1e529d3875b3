// tau_lu.hip -- batched blocked complex LU of KKR matrices on MI355X.
//
// Reference behaviour replaced: solveTau00zgetrf (linearSolvers_CPU.cpp:98-123:
// LU of M, solve [t0;0], keep the top kkrsz_ns rows) and the block-inverse
// route (zblock_lu_cpp.cpp:19-86 + tau_inv_postproc.f:33-57: eliminate the
// non-central blocks with pivoting inside diagonal blocks, then
// (1-delta)^-1 t0).  The matrix is in the solver layout (central site last,
// identity padding first) and the elimination is right-looking over 64-wide
// tiles with partial pivoting inside each 64x64 diagonal tile, the block-LU
// pivoting scheme of the reference's block-inverse solvers.
//
// Per step s (tile column c0 = 64 s), batched over all solves:
//   k_lu_diag    factor A_ss = P^T L U in LDS; store L^-1, U^-1 and P
//   k_lu_panel   L_is = A_is U^-1 (tiles below), U_sj = L^-1 P A_sj (right)
//   k_lu_update  A_ij -= L_is U_sj                     <- MFMA, dominant
// with the steps taken in groups (k_lu_cross, k_lu_trail: the trailing
// matrix is read and written once per group; on the split-FP16 path
// k_lu_panel_h takes the cross updates inside the panel and the trailing
// update is k_lu_trail_h in update_f16.hip); then k_lu_last factors the last
// tile with the pivots of its non-central
// rows and of the central rows kept apart.  In FP64 it also solves
// S tau00 = t0 with the Schur complement S (the plain path); in FP32 it
// stores the last tile's factors for the refinement solves (mixed path).
//
// All kernels are templated on the real type: double (complex FP64, MFMA
// v_mfma_f64_16x16x4f64) or float (complex FP32, v_mfma_f32_16x16x4f32).
#include "lu_common.hpp"

namespace lsmsx {

// In-LDS LU with partial pivoting of a 64x64 tile A[col][row].  Pivot rows
// for column j come from [j, nd) while j < nd and from [j, 64) afterwards,
// so rows >= nd (the central site of the last tile) never pivot into the
// non-central part.  Full-row swaps (LAPACK getrf convention inside the
// tile), piv[] the swap sequence, perm[] the resulting permutation.  256 threads.
template <class T>
__device__ void tile_lu_lds(CxT<T> (*A)[kTile + 1], int nd, int *piv, int *perm, int *sinfo,
                            double *sminpiv, int col_base) {
  typedef CxT<T> V;
  const int tid = threadIdx.x;
  for (int j = 0; j < kTile; j++) {
    const int hi = (j < nd) ? nd : kTile;
    if (tid < 64) {
      // wave 0: pivot search, row swap (lane = column) and column scaling;
      // wave-scope ordering of its LDS traffic is enough between them
      T v = (tid >= j && tid < hi) ? cabs1(A[j][tid]) : T(-1);
      int idx = tid;
      wave_argmax(v, idx);
      if (tid == 0) {
        piv[j] = idx;
        int t = perm[j];
        perm[j] = perm[idx];
        perm[idx] = t;
        double ax = A[j][idx].x, ay = A[j][idx].y;
        double a = sqrt(ax * ax + ay * ay);
        if (a < *sminpiv) *sminpiv = a;
        if (v == T(0) && *sinfo == 0) *sinfo = col_base + j + 1;
      }
      const int p = idx;
      if (p != j) {
        V t = A[tid][j];
        A[tid][j] = A[tid][p];
        A[tid][p] = t;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const V a = A[j][j];
      if ((a.x != T(0) || a.y != T(0)) && tid > j) A[j][tid] = cmul2(A[j][tid], crecip(a));
    }
    __syncthreads();
    const V a = A[j][j];
    const bool nz = (a.x != T(0) || a.y != T(0));
    // rank-1 update of the trailing block: thread (row r = tid % 64, column
    // class tid / 64) takes columns c > j of its class, with l = A(r, j) read
    // once
    if (nz && (tid & 63) > j) {
      const int r = tid & 63;
      const V l = A[j][r];
      int c = tid >> 6;
      if (c <= j) c += ((j - c) / 4 + 1) * 4;
      // four columns at a time, all loads before the stores, so the LDS
      // round trips overlap
      for (; c < 64; c += 16) {
        V u[4], x[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (c + 4 * q < 64) {
            u[q] = A[c + 4 * q][j];
            x[q] = A[c + 4 * q][r];
          }
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (c + 4 * q < 64) {
            x[q].x -= l.x * u[q].x - l.y * u[q].y;
            x[q].y -= l.x * u[q].y + l.y * u[q].x;
            A[c + 4 * q][r] = x[q];
          }
      }
    }
    __syncthreads();
  }
}

// In-place inversion of the factored tile: unit-lower L (waves 0 and 2 own
// rows) and upper U (waves 1 and 3 own rows).  L^-1: for j = 63..0,
// X[r][j] = -(L[r][j] + sum_{j<k<r} X[r][k] L[k][j]) for r > j.  U^-1: for
// j = 0..63, X[j][j] = 1/U[j][j], X[r][j] = -X[j][j] sum_{r<=k<j} X[r][k] U[k][j].
// Each sum is split between the two waves of its part by the parity of k
// (two accumulators per thread), so the dependent LDS loads of a row are a
// quarter as long; the second wave's half goes through LDS.
template <class T>
__device__ void tile_invert_lds(CxT<T> (*A)[kTile + 1]) {
  typedef CxT<T> V;
  __shared__ V part[2][kTile];
  const int tid = threadIdx.x, r = tid & 63, w = tid >> 6;
  const bool lpart = (w & 1) == 0;  // waves 0, 2: L^-1; waves 1, 3: U^-1
  const int half = w >> 1;          // the parity of k this wave sums
  auto madd = [](V &acc, V x, V y) {
    acc.x += x.x * y.x - x.y * y.y;
    acc.y += x.x * y.y + x.y * y.x;
  };
  for (int jj = 0; jj < kTile; jj++) {
    const int jl = kTile - 1 - jj;
    const int ju = jj;
    V s0 = czero<V>(), s1 = czero<V>();
    if (lpart) {
      if (r > jl) {
        int k = jl + 1 + half;
        for (; k + 2 < r; k += 4) {
          madd(s0, A[k][r], A[jl][k]);
          madd(s1, A[k + 2][r], A[jl][k + 2]);
        }
        if (k < r) madd(s0, A[k][r], A[jl][k]);
      }
    } else if (r < ju) {
      int k = r + half;
      for (; k + 2 < ju; k += 4) {
        madd(s0, A[k][r], A[ju][k]);
        madd(s1, A[k + 2][r], A[ju][k + 2]);
      }
      if (k < ju) madd(s0, A[k][r], A[ju][k]);
    }
    s0.x += s1.x;
    s0.y += s1.y;
    if (half == 1) part[lpart ? 0 : 1][r] = s0;
    __syncthreads();
    V nv = czero<V>();
    if (half == 0) {
      const V o = part[lpart ? 0 : 1][r];
      s0.x += o.x;
      s0.y += o.y;
      if (lpart) {
        if (r > jl) {
          const V a = A[jl][r];
          nv.x = -(a.x + s0.x);
          nv.y = -(a.y + s0.y);
        }
      } else {
        const V ujj = A[ju][ju];
        const bool nz = (ujj.x != T(0) || ujj.y != T(0));
        const V xjj = nz ? crecip(ujj) : czero<V>();
        if (r < ju) {
          const V v = cmul2(s0, xjj);
          nv.x = -v.x;
          nv.y = -v.y;
        } else if (r == ju) {
          nv = xjj;
        }
      }
    }
    __syncthreads();  // every read of this step is done
    if (half == 0) {
      if (lpart) {
        if (r > jl) A[jl][r] = nv;
      } else if (r <= ju) {
        A[ju][r] = nv;
      }
    }
    __syncthreads();
  }
}

// Stage a 64 x 64 tile (column stride ld) in LDS, S[col][row] =
// src[rowmap(row) + ld col]: the 16 loads of a thread are all issued before
// its first LDS write, so staging costs one memory round trip, not sixteen.
template <class T, class RM>
__device__ __forceinline__ void stage_tile(CxT<T> (*S)[kTile + 1], const GCx<T> *src, long ld,
                                           RM rowmap) {
  const int tid = threadIdx.x, r = tid & 63, c0 = tid >> 6;
  const int rr = rowmap(r);
  CxT<T> v[kTile / 4];
#pragma unroll
  for (int q = 0; q < kTile / 4; q++) v[q] = ldg<T>(src + rr + ld * (c0 + 4 * q));
#pragma unroll
  for (int q = 0; q < kTile / 4; q++) S[c0 + 4 * q][r] = v[q];
}

// Power-of-2 scale of a 64 x 64 tile for its split-FP16 planes, 256
// threads: ex with max |Re|, |Im| of get(rc, k) times 2^-ex below 2^14
// (inside the FP16 range with room for the lo part; 0 for a zero tile).
template <class F>
__device__ int plane_exponent(F get) {
  __shared__ float wmax[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  __syncthreads();  // the values were written by all four waves
  float m = 0.f;
  for (int e = tid; e < kTile * kTile; e += 256) {
    const float2 v = get(e >> 6, e & 63);
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if (lane == 0) wmax[w] = m;
  __syncthreads();
  m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  int ex = 0;
  if (m > 0.f) {
    frexpf(m, &ex);  // m < 2^ex
    ex -= 14;
  }
  return ex;
}

// Eight complex values x 2^-ex (Re times sgn) split into hi/lo FP16 planes:
// x 2^-ex = hi + lo to 22 bits.
__device__ __forceinline__ void split8(const float2 (&v)[8], float sgn, float sc, h8 &rh, h8 &rl,
                                       h8 &ih, h8 &il) {
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const float re = sgn * v[u].x * sc, im = v[u].y * sc;
    const _Float16 a = (_Float16)re, b = (_Float16)im;
    rh[u] = a;
    rl[u] = (_Float16)(re - (float)a);
    ih[u] = b;
    il[u] = (_Float16)(im - (float)b);
  }
}

// Split the 64 x 64 values get(rc, k) (256 threads) into hi/lo FP16 planes
// with one power-of-2 scale (plane_exponent) for the split-FP16 products.
// Layout, k fastest:
//   L tile (rc = row i, k = tile column):  planes -Re, -Re lo, Im, Im lo ([i][k]);
//   U tile (k = tile row, rc = column j):  planes  Re,  Re lo, Im, Im lo ([j][k]).
// The L planes carry the sign of the update C -= L U.
template <class F>
__device__ void write_planes(F get, bool lower, _Float16 *out, int *eout) {
  const int tid = threadIdx.x;
  const int ex = plane_exponent(get);
  if (tid == 0) *eout = ex;
  const float sc = ldexpf(1.f, -ex);
  const float sgn = lower ? -1.f : 1.f;
  for (int q = tid; q < kTile * 8; q += 256) {
    const int rc = q >> 3, kg = q & 7;  // row i (L) or column j (U); k group
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = get(rc, 8 * kg + u);
    h8 rh, rl, ih, il;
    split8(v, sgn, sc, rh, rl, ih, il);
    h8 *o = (h8 *)(out + rc * kTile + 8 * kg);
    o[0] = rh;
    o[(kTile * kTile) / 8] = rl;
    o[2 * (kTile * kTile) / 8] = ih;
    o[3 * (kTile * kTile) / 8] = il;
  }
}

// The planes of an FP32 panel tile held in LDS as Ts[col][row].
__device__ __forceinline__ void write_split_planes(float2 (*Ts)[kTile + 1], bool lower,
                                                   _Float16 *out, int *eout) {
  if (lower)
    write_planes([&](int i, int k) { return Ts[k][i]; }, true, out, eout);
  else
    write_planes([&](int j, int k) { return Ts[j][k]; }, false, out, eout);
}

// ---------------------------------------------------------------------------
// Diagonal tile of step s: factor, write back L\U, store L^-1 (strict lower)
// and U^-1 (upper incl. diagonal) packed in one tile, and the permutation.
template <class T>
__global__ __launch_bounds__(256) void k_lu_diag(const SolveDesc *__restrict__ descs, int step) {
  typedef CxT<T> V;
  const SolveDesc &d = descs[blockIdx.x];
  step -= d.step0;
  if (step < 0 || step > d.nt - 1 || (step == d.nt - 1 && !d.full_lu)) return;
  if (d.conv && *d.conv) return;
  __shared__ V A[kTile][kTile + 1];
  __shared__ int piv[kTile], perm[kTile];
  __shared__ int sinfo;
  __shared__ double sminpiv;
  const int tid = threadIdx.x;
  const int c0 = step * kTile;
  GCx<T> *Mt = tile_ptr<T>(d, step, step);
  const long ld = tile_ld(d);
  stage_tile<T>(A, Mt, ld, [](int r) { return r; });
  if (tid < kTile) perm[tid] = tid;
  if (tid == 0) {
    sinfo = 0;
    sminpiv = 1e300;
  }
  __syncthreads();
  tile_lu_lds<T>(A, kTile, piv, perm, &sinfo, &sminpiv, c0);
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    Mt[r + ld * c] = A[c][r];
  }
  if (tid < kTile) d.perm[step * kTile + tid] = perm[tid];
  if (tid == 0) {
    if (sinfo != 0 && *d.info == 0) *d.info = sinfo;
    if (sminpiv < *d.minpiv) *d.minpiv = sminpiv;
  }
  __syncthreads();
  tile_invert_lds<T>(A);
  GCx<T> *inv = (GCx<T> *)d.inv + step * d.inv_stride;
  for (int e = tid; e < kTile * kTile; e += 256) {
    int r = e & 63, c = e >> 6;
    inv[r + kTile * c] = A[c][r];
  }
  if constexpr (sizeof(T) == 4) {
    if (d.ihp) {
      // split planes of L^-1 (unit lower, L layout) and U^-1 (U layout) for
      // the fused panel's FP16 products
      _Float16 *ip = (_Float16 *)d.ihp;
      write_planes(
          [&](int i, int k) {
            V v = czero<V>();
            if (i > k) v = A[k][i];
            if (i == k) v.x = 1.f;
            return v;
          },
          true, ip, d.ihe);
      write_planes(
          [&](int j, int k) {
            V v = czero<V>();
            if (k <= j) v = A[j][k];
            return v;
          },
          false, ip + kPlaneHalves, d.ihe + 1);
    }
  }
}

// Cross update of tile (ti, tj) inside the split-FP16 panel, C = A - sum over
// the steps s_first .. s_first+ns-1 of L(ti, s) U(s, tj), written to
// Ts[col][rowpos[row]] (rowpos null: the identity).  The products run on the
// split planes of those steps (slot s mod group) with the FP16 MFMAs of the
// trailing update, in its orientation (a-operand a U plane, lane l: column
// l%16; b-operand an L plane, lane l: row l%16; k = 8(l/16)..+7 of a 32-wide
// chunk; register r of lane l is C(row l%16, column 4(l/16) + r)) and its
// 4M x 3 split products, FP32 accumulators that start from zero and are
// rescaled between the steps by the tile exponents (exact powers of two).
// Wave w takes rows wr.., columns wc.. (32 x 32).  Per chunk the L and U
// planes (16 KiB each) are staged in Ts's memory with the XOR-swizzled k
// groups, the next chunk's loads in flight during this chunk's MFMAs (eight
// 16-byte loads per thread).  A (the tile's FP32 values) is read during the
// last chunk and added at the end.
__device__ __forceinline__ void cross_to_lds(const SolveDesc &d, int ti, int tj, int s_first,
                                             int ns, const GCx<float> *A, long ld,
                                             float2 (*Ts)[kTile + 1], const int *rowpos) {
  __shared__ int exs[16];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, l16 = lane & 15, g = lane >> 4;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const int nt = d.nt, gmask = d.group - 1;
  const _Float16 *hp = (const _Float16 *)d.hp;
  h8(*Ls)[kTile][4] = reinterpret_cast<h8(*)[kTile][4]>(&Ts[0][0]);  // [plane][row][k group]
  h8(*Us)[kTile][4] = Ls + 4;                                          // [plane][column][k group]
  if (tid < ns) {
    const int par = (s_first + tid) & gmask;
    exs[tid] = d.he[(par * 2 + 0) * nt + ti] + d.he[(par * 2 + 1) * nt + tj];
  }
  __syncthreads();
  // chunk ch = (step ch / 2, k half ch % 2); load q of a thread: side (L, U),
  // plane, row (column), k group
  auto src = [&](int ch, int q) __attribute__((always_inline)) -> GH8 * {
    const int s = s_first + (ch >> 1), kh = ch & 1, par = s & gmask;
    const int e8 = tid + 256 * q, side = e8 >> 10, plane = (e8 >> 8) & 3, rr = (e8 >> 2) & 63,
              grp = e8 & 3;
    return (GH8 *)(hp + (long)((par * 2 + side) * nt + (side ? tj : ti)) * kPlaneHalves) +
           ((plane * kTile + rr) * kTile + 32 * kh + 8 * grp) / 8;
  };
  auto dst = [&](int q) __attribute__((always_inline)) -> h8 * {
    const int e8 = tid + 256 * q, side = e8 >> 10, plane = (e8 >> 8) & 3, rr = (e8 >> 2) & 63,
              grp = e8 & 3;
    return &(side ? Us : Ls)[plane][rr][swz_group(rr, grp)];
  };
  f4 cr[2][2], ci[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) {
      cr[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
      ci[bi][bj] = cr[bi][bj];
    }
  int e = exs[0];
  auto rescale = [&](int ch) __attribute__((always_inline)) {
    if (ch > 0 && !(ch & 1)) {
      const int e2 = exs[ch >> 1];
      if (e2 == e) return;
      const float sc1 = ldexpf(1.f, e - e2);
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          cr[bi][bj] *= sc1;
          ci[bi][bj] *= sc1;
        }
      e = e2;
    }
  };
  // Cr += (-Lr) Ur + Li Ui, Ci += (-Lr) Ui + Li (-Ur): the L planes hold -Re
  auto compute = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int bj = 0; bj < 2; bj++) {
      const int col = wc + 16 * bj + l16, gu = swz_group(col, g);
      const h8 ur_h = Us[0][col][gu], ur_l = Us[1][col][gu];
      const h8 ui_h = Us[2][col][gu], ui_l = Us[3][col][gu];
      const h8 nur_h = neg8(ur_h), nur_l = neg8(ur_l);
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        const int row = wr + 16 * bi + l16, gl = swz_group(row, g);
        const h8 nar_h = Ls[0][row][gl], nar_l = Ls[1][row][gl];
        const h8 ai_h = Ls[2][row][gl], ai_l = Ls[3][row][gl];
        f4 r = cr[bi][bj], m = ci[bi][bj];
        r = mfma_h(ur_h, nar_h, r);
        r = mfma_h(ur_l, nar_h, r);
        r = mfma_h(ur_h, nar_l, r);
        r = mfma_h(ui_h, ai_h, r);
        r = mfma_h(ui_l, ai_h, r);
        r = mfma_h(ui_h, ai_l, r);
        m = mfma_h(ui_h, nar_h, m);
        m = mfma_h(ui_l, nar_h, m);
        m = mfma_h(ui_h, nar_l, m);
        m = mfma_h(nur_h, ai_h, m);
        m = mfma_h(nur_l, ai_h, m);
        m = mfma_h(nur_h, ai_l, m);
        cr[bi][bj] = r;
        ci[bi][bj] = m;
      }
    }
  };
  const int nch = 2 * ns;
  h8 pv[8];
#pragma unroll
  for (int q = 0; q < 8; q++) pv[q] = *src(0, q);
#pragma unroll 1
  for (int ch = 0; ch < nch - 1; ch++) {
    rescale(ch);
    __syncthreads();  // the previous chunk's readers are done
#pragma unroll
    for (int q = 0; q < 8; q++) *dst(q) = pv[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) pv[q] = *src(ch + 1, q);
    compute();
  }
  // the last chunk, with A's loads in the prefetch registers' place:
  // A(wr + 16 bi + l16, wc + 16 bj + 4 g + r)
  rescale(nch - 1);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; q++) *dst(q) = pv[q];
  __syncthreads();
  float2 av[2][2][4];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        av[bi][bj][r] = ldg<float>(A + (wr + 16 * bi + l16) + ld * (wc + 16 * bj + 4 * g + r));
  compute();
  __syncthreads();  // every wave is done reading the staged planes
  const float sc2 = ldexpf(1.f, e);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = wr + 16 * bi + l16, col = wc + 16 * bj + 4 * g + r;
        Ts[col][rowpos ? rowpos[row] : row] =
            make_float2(av[bi][bj][r].x + cr[bi][bj][r] * sc2, av[bi][bj][r].y + ci[bi][bj][r] * sc2);
      }
}

// Panel of step s: x < nrest -> L_is = A_is U^-1 ; else U_sj = L^-1 P A_sj.
// The operand tile is staged in LDS (rows gathered by the pivot permutation
// for U_sj) before it is overwritten in place.  With CROSS (the split-FP16
// path) the tile first takes the cross update with the ncross steps before s
// in its group, on the way into LDS (cross_to_lds), so that no separate
// cross kernel reads and writes it.
template <class T, bool CROSS>
__device__ __forceinline__ void panel_tile(const SolveDesc &d, int step, int ncross, int x) {
  typedef CxT<T> V;
  const int nrest = d.nt - 1 - step;
  if (step < 0 || nrest <= 0 || x >= 2 * nrest) return;
  if (d.conv && *d.conv) return;
  __shared__ __align__(16) V Ts[kTile][kTile + 1];  // Ts[col][row]
  const int tid = threadIdx.x, w = tid >> 6;
  const int lane = tid & 63, li = lane & 15, kq = lane >> 4;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const long ld = tile_ld(d);
  const GCx<T> *__restrict__ inv = (const GCx<T> *)d.inv + step * d.inv_stride;
  const bool lower = x < nrest;
  const int ti = lower ? step + 1 + x : step, tj = lower ? step : step + 1 + x - nrest;
  GCx<T> *tile = tile_ptr<T>(d, ti, tj);
  bool staged = false;
  if constexpr (CROSS) {
    if (ncross > 0) {
      // the row of C that lands at LDS row r is perm[r] (U_sj): rowpos = perm^-1
      __shared__ int rowpos[kTile];
      if (!lower && tid < kTile) rowpos[d.perm[step * kTile + tid]] = tid;
      cross_to_lds(d, ti, tj, step - ncross, ncross, tile, ld, Ts, lower ? nullptr : rowpos);
      staged = true;
    }
  }
  if (!staged) {
    if (lower) {
      stage_tile<T>(Ts, tile, ld, [](int r) { return r; });
    } else {
      const int *perm = d.perm + step * kTile;
      stage_tile<T>(Ts, tile, ld, [&](int r) { return perm[r]; });
    }
  }
  __syncthreads();
  V ov[2][2][4];  // output (row wr + 16 bi + li, column wc + 16 bj + acc_row(kq, r))
  if constexpr (CROSS) {
    // split-FP16 product with the diagonal tile's inverse planes (written by
    // k_lu_diag): the tile's side is split here with its own scale, the
    // accumulators take -(A B) as in the cross update (b-operand rows with
    // -Re, a-operand columns with +Re), two 32-wide k chunks
    const int eC = plane_exponent([&](int a, int b) { return Ts[a][b]; });
    const float sc = ldexpf(1.f, -eC);
    const int g = kq, l16 = li;
    GH8 *ip = (GH8 *)((const _Float16 *)d.ihp + (lower ? kPlaneHalves : 0));
    const int eI = d.ihe[lower ? 1 : 0];
    f4 cr[2][2], ci[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        cr[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
        ci[bi][bj] = cr[bi][bj];
      }
#pragma unroll 1
    for (int kh = 0; kh < 2; kh++) {
      // b-operands (rows wr + 16 bi ..): L_is = C U^-1 takes the rows of C
      // from Ts[k][i], U_sj = L^-1 (P C) the rows of L^-1 from its planes
      h8 lo[2][4];
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        if (lower) {
          float2 v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = Ts[32 * kh + 8 * g + u][wr + 16 * bi + l16];
          split8(v, -1.f, sc, lo[bi][0], lo[bi][1], lo[bi][2], lo[bi][3]);
        } else {
#pragma unroll
          for (int p = 0; p < 4; p++)
            lo[bi][p] = ip[((p * kTile + wr + 16 * bi + l16) * kTile + 32 * kh + 8 * g) / 8];
        }
      }
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        // a-operand (columns wc + 16 bj ..): U^-1 from its planes, or P C from Ts[j][k]
        h8 uo[4];
        if (lower) {
#pragma unroll
          for (int p = 0; p < 4; p++)
            uo[p] = ip[((p * kTile + wc + 16 * bj + l16) * kTile + 32 * kh + 8 * g) / 8];
        } else {
          float2 v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = Ts[wc + 16 * bj + l16][32 * kh + 8 * g + u];
          split8(v, 1.f, sc, uo[0], uo[1], uo[2], uo[3]);
        }
        const h8 nur_h = neg8(uo[0]), nur_l = neg8(uo[1]);
#pragma unroll
        for (int bi = 0; bi < 2; bi++) {
          f4 r = cr[bi][bj], m = ci[bi][bj];
          r = mfma_h(uo[0], lo[bi][0], r);
          r = mfma_h(uo[1], lo[bi][0], r);
          r = mfma_h(uo[0], lo[bi][1], r);
          r = mfma_h(uo[2], lo[bi][2], r);
          r = mfma_h(uo[3], lo[bi][2], r);
          r = mfma_h(uo[2], lo[bi][3], r);
          m = mfma_h(uo[2], lo[bi][0], m);
          m = mfma_h(uo[3], lo[bi][0], m);
          m = mfma_h(uo[2], lo[bi][1], m);
          m = mfma_h(nur_h, lo[bi][2], m);
          m = mfma_h(nur_l, lo[bi][2], m);
          m = mfma_h(nur_h, lo[bi][3], m);
          cr[bi][bj] = r;
          ci[bi][bj] = m;
        }
      }
    }
    const float so = -ldexpf(1.f, eC + eI);
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int r = 0; r < 4; r++) ov[bj][bi][r] = make_float2(cr[bi][bj][r] * so, ci[bi][bj][r] * so);
  } else {
    Acc3<T> q;
    zero_acc(q);
    // operand k-steps in flight: three for FP64 (the inverse's L2 latency),
    // one for FP32, where more registers would cost a wave per SIMD
    constexpr int kPanelPF = sizeof(T) == 8 ? 3 : 1;
    if (lower) {
      // A(i,k) = Ts(wr+i, k); B(k,j) = U^-1(k, wc+j), upper incl. diagonal
      wave_gemm3m<T, false, kPanelPF>([&](int i, int k) { return Ts[k][wr + i]; },
                                      [&](int k, int j) {
                                        int jj = wc + j;
                                        V v = czero<V>();
                                        if (k <= jj) v = inv[k + kTile * jj];
                                        return v;
                                      },
                                      q);
    } else {
      // A(i,k) = L^-1(wr+i, k), unit lower; B(k,j) = (P A_sj)(k, wc+j)
      wave_gemm3m<T, false, kPanelPF>([&](int i, int k) {
                                        int ii = wr + i;
                                        V v = czero<V>();
                                        if (ii == k) v.x = T(1);
                                        if (ii > k) v = inv[ii + kTile * k];
                                        return v;
                                      },
                                      [&](int k, int j) { return Ts[wc + j][k]; }, q);
    }
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          ov[bj][bi][r].x = q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
          ov[bj][bi][r].y = q.p[2][bj][bi][r] - q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
        }
  }
  GCx<T> *C = tile + wr + ld * wc;
  const bool planes = sizeof(T) == 4 && d.hp != nullptr;
  if (planes) __syncthreads();  // every wave is done reading Ts
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const V o = ov[bj][bi][r];
        if constexpr (CROSS) {
          // the split path's FP32 factor tile is read again only by the
          // refinement sweeps after the whole factorization: non-temporal
          typedef float pv2 __attribute__((ext_vector_type(2)));
          __builtin_nontemporal_store(pv2{o.x, o.y},
                                      (pv2 *)&C[(16 * bi + li) + ld * (16 * bj + acc_row<T>(kq, r))]);
        } else {
          C[(16 * bi + li) + ld * (16 * bj + acc_row<T>(kq, r))] = o;
        }
        if (planes) Ts[wc + 16 * bj + acc_row<T>(kq, r)][wr + 16 * bi + li] = o;
      }
  if constexpr (sizeof(T) == 4) {
    if (planes) {
      // split-FP16 planes of this panel tile for the split updates
      const int parity = step % d.group, side = lower ? 0 : 1;
      const int t = lower ? ti : tj;
      write_split_planes(Ts, lower, (_Float16 *)d.hp + (long)((parity * 2 + side) * d.nt + t) * kPlaneHalves,
                         d.he + (parity * 2 + side) * d.nt + t);
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_lu_panel(const SolveDesc *__restrict__ descs, int step) {
  const SolveDesc &d = descs[blockIdx.y];
  panel_tile<T, false>(d, step - d.step0, 0, blockIdx.x);
}

// Panel of step `step` on the split-FP16 path with its cross update fused:
// the tiles of column and row `step` first take the ncross steps before it
// in its group (planes in the group's slots), then the panel.  With
// xcd_swizzle the workgroups of one solve run on one XCD (xcd_remap), so the
// planes its tiles share (the U planes of the step's column, the L planes of
// its row) are fetched into one L2.
__global__ __launch_bounds__(256) void k_lu_panel_h(const SolveDesc *__restrict__ descs, int step,
                                                      int ncross, int xcd_swizzle) {
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcd_swizzle) xcd_remap(bx, by);
  const SolveDesc &d = descs[by];
  panel_tile<float, true>(d, step - d.step0, ncross, bx);
}

// C(ti, tj) -= sum_{u < ns} L(ti, s0+u) U(s0+u, tj): one 64x64 tile per
// workgroup, a 32x32 quarter per wave.  C enters the 3M sums at the start
// (q1 = Cre, q2 = 0, q3 = Cre + Cim, products accumulated with -A), so
// C' = (q1 - q2, q3 - q1 - q2): the tile is read and written once however
// many rank-64 panels are applied.
template <class T>
__device__ __forceinline__ void update_tile(const SolveDesc &d, int ti, int tj, int s0, int ns) {
  typedef CxT<T> V;
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  const int w = threadIdx.x >> 6;
  const int wr = 32 * (w & 1), wc = 32 * (w >> 1);
  const long ld = tile_ld(d);
  GCx<T> *__restrict__ C = tile_ptr<T>(d, ti, tj) + wr + ld * wc;
  Acc3<T> q;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // C is read and written once per group: non-temporal, so that the
        // panels the other tiles read stay in L2
        typedef T cv2 __attribute__((ext_vector_type(2)));
        const cv2 c = __builtin_nontemporal_load(
            (const cv2 *)&C[(16 * bi + li) + ld * (16 * bj + acc_row<T>(kq, r))]);
        q.p[0][bj][bi][r] = c.x;
        q.p[1][bj][bi][r] = T(0);
        q.p[2][bj][bi][r] = c.x + c.y;
      }
  for (int u = 0; u < ns; u++) {
    const GCx<T> *__restrict__ A = tile_ptr<T>(d, ti, s0 + u) + wr;
    const GCx<T> *__restrict__ B = tile_ptr<T>(d, s0 + u, tj) + ld * wc;
    wave_gemm3m<T, true>([&](int i, int k) { return A[i + ld * k]; },
                         [&](int k, int j) { return B[k + ld * j]; }, q);
  }
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        V o;
        o.x = q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
        o.y = q.p[2][bj][bi][r] - q.p[0][bj][bi][r] - q.p[1][bj][bi][r];
        typedef T cv2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(cv2{o.x, o.y},
                                    (cv2 *)&C[(16 * bi + li) + ld * (16 * bj + acc_row<T>(kq, r))]);
      }
}

// Trailing update of step s: A_ij -= L_is U_sj for all i, j > s.
template <class T>
__global__ __launch_bounds__(256) void k_lu_update(const SolveDesc *__restrict__ descs, int step,
                                                   int xcd_swizzle) {
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcd_swizzle) xcd_remap(bx, by);
  const SolveDesc &d = descs[by];
  step -= d.step0;
  const int nrest = d.nt - 1 - step;
  const int x = bx;
  if (step < 0 || nrest <= 0 || x >= nrest * nrest) return;
  if (d.conv && *d.conv) return;
  update_tile<T>(d, step + 1 + x % nrest, step + 1 + x / nrest, step, 1);
}

// Grouped elimination: the steps g .. g+ns-1 of a group (ns = min(group,
// nt-1-g); step nt-1 is k_lu_last) are factored one after another, each
// after a cross update that brings its tile column (rows >= t) and tile row
// (columns > t) up to date with the group's earlier steps; then one trailing
// update applies all ns steps to the tiles beyond the group.
// Number of steps of the group starting at solve step g.
__device__ __forceinline__ int group_steps(const SolveDesc &d, int g) {
  return min(d.group, d.nt - 1 - g);
}

// Cross update for step t = g + u of the group starting at global step s:
// tile column t (rows >= t) and tile row t (columns > t) -= the products of
// steps g .. t-1.
template <class T>
__global__ __launch_bounds__(256) void k_lu_cross(const SolveDesc *__restrict__ descs, int step,
                                                  int u) {
  const SolveDesc &d = descs[blockIdx.y];
  const int g = step - d.step0, t = g + u;
  if (g < 0 || u >= group_steps(d, g)) return;
  if (d.conv && *d.conv) return;
  const int n1 = d.nt - t, x = blockIdx.x;
  if (x >= 2 * n1 - 1) return;
  if (x < n1) update_tile<T>(d, t + x, t, g, u);
  else update_tile<T>(d, t, t + 1 + (x - n1), g, u);
}

// Trailing update of the group starting at global step s: A_ij -= sum over
// the group's steps k of L_ik U_kj for all i, j beyond the group, one pass
// over C with K = 64 ns.
template <class T>
__global__ __launch_bounds__(256) void k_lu_trail(const SolveDesc *__restrict__ descs, int step,
                                                  int xcd_swizzle) {
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcd_swizzle) xcd_remap(bx, by);
  const SolveDesc &d = descs[by];
  const int g = step - d.step0;
  if (g < 0) return;
  const int ns = group_steps(d, g);
  if (ns <= 0) return;
  if (d.conv && *d.conv) return;
  const int last = g + ns - 1, nrest = d.nt - 1 - last, x = bx;
  if (x >= nrest * nrest) return;
  update_tile<T>(d, last + 1 + x % nrest, last + 1 + x / nrest, g, ns);
}

// Last tile (step nt-1): factor it with the non-central rows (first nd =
// 64 - kc) pivoting among themselves and the central rows among themselves.
// The trailing kc x kc block becomes LU(S), S the Schur complement.
//   plain FP64 path: solve S tau00 = t0 from the S factors;
//   refined paths (d.X set; FP32 mixed or FP64 refined): store the tile's
//   factors, inverses and permutation like k_lu_diag, for the refinement.
template <class T>
__global__ __launch_bounds__(256) void k_lu_last(const SolveDesc *__restrict__ descs) {
  typedef CxT<T> V;
  const SolveDesc &d = descs[blockIdx.x];
  if (d.conv && *d.conv) return;
  __shared__ V A[kTile][kTile + 1];
  __shared__ V X[kTile][kTile + 1];
  __shared__ int piv[kTile], perm[kTile];
  __shared__ int sinfo;
  __shared__ double sminpiv;
  const int tid = threadIdx.x;
  const int step = d.nt - 1;
  const int c0 = step * kTile;
  const int kc = d.kc, nd = kTile - kc;
  GCx<T> *Mt = tile_ptr<T>(d, step, step);
  const long ld = tile_ld(d);
  stage_tile<T>(A, Mt, ld, [](int r) { return r; });
  if (tid < kTile) perm[tid] = tid;
  if (tid == 0) {
    sinfo = 0;
    sminpiv = 1e300;
  }
  __syncthreads();
  tile_lu_lds<T>(A, nd, piv, perm, &sinfo, &sminpiv, c0);
  if (tid == 0) {
    if (sinfo != 0 && *d.info == 0) *d.info = sinfo;
    if (sminpiv < *d.minpiv) *d.minpiv = sminpiv;
  }
  if (d.X != nullptr) {  // refinement follows: keep the tile's factors
    for (int e = tid; e < kTile * kTile; e += 256) {
      int r = e & 63, c = e >> 6;
      Mt[r + ld * c] = A[c][r];
    }
    if (tid < kTile) d.perm[step * kTile + tid] = perm[tid];
    __syncthreads();
    tile_invert_lds<T>(A);
    GCx<T> *inv = (GCx<T> *)d.inv + step * d.inv_stride;
    for (int e = tid; e < kTile * kTile; e += 256) {
      int r = e & 63, c = e >> 6;
      inv[r + kTile * c] = A[c][r];
    }
    return;
  }
  // X = P_S t0, then unit-lower and upper solves with the S factors; one
  // thread per right-hand side.
  for (int e = tid; e < kc * kc; e += 256) {
    int r = e % kc, c = e / kc;
    double2 t = d.t0[r + (long)d.ldt * c];
    X[c][r].x = t.x;
    X[c][r].y = t.y;
  }
  __syncthreads();
  if (tid < kc) {
    const int c = tid;
    for (int j = nd; j < kTile; j++) {
      int p = piv[j];
      if (p != j) {
        V t = X[c][j - nd];
        X[c][j - nd] = X[c][p - nd];
        X[c][p - nd] = t;
      }
    }
    for (int j = 0; j < kc; j++) {
      V xj = X[c][j];
      for (int r = j + 1; r < kc; r++) {
        V l = A[nd + j][nd + r];
        X[c][r].x -= l.x * xj.x - l.y * xj.y;
        X[c][r].y -= l.x * xj.y + l.y * xj.x;
      }
    }
    for (int j = kc - 1; j >= 0; j--) {
      V u = A[nd + j][nd + j];
      V xj = (u.x != T(0) || u.y != T(0)) ? cmul2(X[c][j], crecip(u)) : X[c][j];
      X[c][j] = xj;
      for (int r = 0; r < j; r++) {
        V uu = A[nd + j][nd + r];
        X[c][r].x -= uu.x * xj.x - uu.y * xj.y;
        X[c][r].y -= uu.x * xj.y + uu.y * xj.x;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < kc * kc; e += 256) {
    int r = e % kc, c = e / kc;
    d.tau00[r + kc * c] = make_double2(X[c][r].x, X[c][r].y);
  }
}

template __global__ void k_lu_diag<double>(const SolveDesc *, int);
template __global__ void k_lu_diag<float>(const SolveDesc *, int);
template __global__ void k_lu_panel<double>(const SolveDesc *, int);
template __global__ void k_lu_panel<float>(const SolveDesc *, int);
template __global__ void k_lu_update<double>(const SolveDesc *, int, int);
template __global__ void k_lu_update<float>(const SolveDesc *, int, int);
template __global__ void k_lu_cross<double>(const SolveDesc *, int, int);
template __global__ void k_lu_cross<float>(const SolveDesc *, int, int);
template __global__ void k_lu_trail<double>(const SolveDesc *, int, int);
template __global__ void k_lu_trail<float>(const SolveDesc *, int, int);
template __global__ void k_lu_last<double>(const SolveDesc *);
template __global__ void k_lu_last<float>(const SolveDesc *);

}  // namespace lsmsx
