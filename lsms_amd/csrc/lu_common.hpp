// lu_common.hpp -- device helpers shared by the LU, refinement and inverse
// kernels: complex types per real type, global-address-space pointers, the
// 16x16x4 MFMA wrappers with their accumulator layout, and the 3M complex
// wave product used by every trailing update.
#pragma once
#include "lsmsx_device.hpp"

namespace lsmsx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T>
struct Cplx;
template <>
struct Cplx<double> {
  typedef double2 type;
  typedef d4 acc;
};
template <>
struct Cplx<float> {
  typedef float2 type;
  typedef f4 acc;
};
template <class T>
using CxT = typename Cplx<T>::type;
template <class T>
using Acc4 = typename Cplx<T>::acc;

// Pointers read from the solve descriptors are generic; casting them to the
// global address space lets the compiler emit global (not flat) loads/stores.
// (The host pass only type-checks kernel bodies, so it sees plain types.)
#if defined(__HIP_DEVICE_COMPILE__)
#define LSMSX_GLOBAL __attribute__((address_space(1)))
#else
#define LSMSX_GLOBAL
#endif
template <class T>
using GCx = LSMSX_GLOBAL CxT<T>;

// Load of one complex value from global memory, field by field.  A struct
// copy out of a global-address-space pointer into an array element becomes a
// memcpy that keeps the array in scratch memory; two scalar loads (merged
// into one wide load) keep it in registers.
template <class T>
__device__ __forceinline__ CxT<T> ldg(const GCx<T> *p) {
  CxT<T> v;
  v.x = p->x;
  v.y = p->y;
  return v;
}

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Accumulator layout of the 16x16x4 MFMAs (measured on gfx950): register r
// of lane l holds output row l/16 + 4r (f64) or 4(l/16) + r (f32), column l%16.
template <class T>
__device__ __forceinline__ int acc_row(int kq, int r);
template <>
__device__ __forceinline__ int acc_row<double>(int kq, int r) {
  return kq + 4 * r;
}
template <>
__device__ __forceinline__ int acc_row<float>(int kq, int r) {
  return 4 * kq + r;
}

template <class V>
__device__ inline auto cabs1(V z) -> decltype(z.x) {
  return fabs(z.x) + fabs(z.y);
}
template <class V>
__device__ inline V cmul2(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class V>
__device__ inline V crecip(V b) {
  auto s = 1 / (b.x * b.x + b.y * b.y);
  V r;
  r.x = b.x * s;
  r.y = -b.y * s;
  return r;
}
template <class V>
__device__ inline V czero() {
  V r;
  r.x = 0;
  r.y = 0;
  return r;
}

// Tile (ti, tj) of the matrix the LU kernels factor, and the stride between
// the columns of a tile: column-major with leading dimension ld, or
// tile-major (d.tiled, tiled_index) with stride 64.  Every LU, refinement and
// update kernel addresses the factored matrix tile by tile through these.
template <class T>
__device__ __forceinline__ GCx<T> *tile_ptr(const SolveDesc &d, int ti, int tj) {
  GCx<T> *A = (GCx<T> *)d.A;
  return d.tiled ? A + ((long)tj * d.nt + ti) * (kTile * kTile)
                 : A + (long)kTile * ti + (long)d.ld * (kTile * tj);
}
__device__ __forceinline__ long tile_ld(const SolveDesc &d) { return d.tiled ? kTile : d.ld; }

// Wave-wide argmax of v (ties -> smallest index); all lanes get the result.
template <class T>
__device__ inline void wave_argmax(T &v, int &idx) {
  for (int off = 32; off >= 1; off >>= 1) {
    T ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// ---------------------------------------------------------------------------
// Complex 32x32x64 wave product on MFMA 16x16x4, 3M form:
//   P1 = Ar Br, P2 = Ai Bi, P3 = (Ar+Ai)(Br+Bi);  re = P1 - P2, im = P3 - P1 - P2
// (three real MFMAs per complex 16x16x4 block instead of four).
// The MFMA is applied to the transposed product, D'(j,i) = sum_k B(k,j) A(i,k),
// so the accumulator of a lane covers consecutive rows i of C and every C and
// A access is a run of 16 consecutive rows.  Operands are read straight from
// L2/LDS into registers in MFMA layout (a-operand lane value B(k0+kq, j0+li),
// b-operand lane value A(i0+li, k0+kq)), loaded ahead of the MFMAs; no LDS
// staging and no barriers, so the waves of a CU run independently.
template <class T>
struct Acc3 {
  Acc4<T> p[3][2][2];
};

// One k-step (12 MFMAs) of the 3M product from operands already in registers.
template <class T, bool NEGA>
__device__ __forceinline__ void gemm3m_step(CxT<T> (&av)[2], CxT<T> (&bv)[2], Acc3<T> &q) {
  T as[2], bs[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    if (NEGA) {
      av[t].x = -av[t].x;
      av[t].y = -av[t].y;
    }
    as[t] = av[t].x + av[t].y;
    bs[t] = bv[t].x + bv[t].y;
  }
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++) q.p[0][bj][bi] = mfma4(bv[bj].x, av[bi].x, q.p[0][bj][bi]);
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++) q.p[1][bj][bi] = mfma4(bv[bj].y, av[bi].y, q.p[1][bj][bi]);
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int bi = 0; bi < 2; bi++) q.p[2][bj][bi] = mfma4(bs[bj], as[bi], q.p[2][bj][bi]);
}

// Operand loads run PF k-steps ahead of the MFMAs: one step (768 MFMA
// cycles) covers an L2 round trip in FP64; FP32 steps (384 cycles) with both
// operands in L2 load three steps ahead (callers whose occupancy matters more
// pass PF = 1).
template <class T, bool NEGA, int PF = (sizeof(T) == 4 ? 3 : 1), class FA, class FB>
__device__ __forceinline__ void wave_gemm3m(FA fa, FB fb, Acc3<T> &q) {
  typedef CxT<T> V;
  const int lane = threadIdx.x & 63, li = lane & 15, kq = lane >> 4;
  if constexpr (PF == 1) {
    V a[2], b[2], an[2], bn[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      a[t] = fa(16 * t + li, kq);
      b[t] = fb(kq, 16 * t + li);
    }
#pragma unroll 2
    for (int k0 = 0; k0 < kTile; k0 += 4) {
      if (k0 + 4 < kTile) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
          an[t] = fa(16 * t + li, k0 + 4 + kq);
          bn[t] = fb(k0 + 4 + kq, 16 * t + li);
        }
      }
      gemm3m_step<T, NEGA>(a, b, q);
#pragma unroll
      for (int t = 0; t < 2; t++) {
        a[t] = an[t];
        b[t] = bn[t];
      }
    }
    return;
  }
  constexpr int NS = kTile / 4;
  V a[PF + 1][2], b[PF + 1][2];
#pragma unroll
  for (int p = 0; p < PF; p++)
#pragma unroll
    for (int t = 0; t < 2; t++) {
      a[p][t] = fa(16 * t + li, 4 * p + kq);
      b[p][t] = fb(4 * p + kq, 16 * t + li);
    }
#pragma unroll
  for (int ks = 0; ks < NS; ks++) {
    if (ks + PF < NS) {
      const int p = (ks + PF) % (PF + 1);
#pragma unroll
      for (int t = 0; t < 2; t++) {
        a[p][t] = fa(16 * t + li, 4 * (ks + PF) + kq);
        b[p][t] = fb(4 * (ks + PF) + kq, 16 * t + li);
      }
    }
    const int c = ks % (PF + 1);
    V av[2] = {a[c][0], a[c][1]}, bv[2] = {b[c][0], b[c][1]};
    gemm3m_step<T, NEGA>(av, bv, q);
  }
}

template <class T>
__device__ __forceinline__ void zero_acc(Acc3<T> &q) {
#pragma unroll
  for (int m = 0; m < 3; m++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int r = 0; r < 4; r++) q.p[m][bj][bi][r] = T(0);
}

// Split-FP16 operands: eight halves of a 32-wide k chunk per lane, the
// FP16 MFMA with FP32 accumulation, and a sign flip of eight halves (four
// 32-bit XORs of the packed sign bits).
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const h8 GH8;
#else
typedef const h8 GH8;
#endif
__device__ __forceinline__ f4 mfma_h(h8 a, h8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
typedef unsigned u4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ h8 neg8(h8 v) {
  return __builtin_bit_cast(h8, __builtin_bit_cast(u4v, v) ^ 0x80008000u);
}

// LDS k group of row r holding logical group g (8 halves): XOR swizzle of the
// four groups of a 64-byte row by (r >> 1) ^ (r >> 2), so that any 8 (and any
// 16) consecutive rows read at one group cover distinct 16-byte bank groups:
// rows of equal parity among 8 (equal r mod 4 among 16) get distinct masks.
__device__ __forceinline__ int swz_group(int r, int g) { return g ^ (((r >> 1) ^ (r >> 2)) & 3); }

// XCD-aware workgroup order.  Workgroups are dispatched round-robin over
// the 8 XCDs by linear id, so consecutive tiles of one solve would land on
// 8 different L2 caches.  The remap gives each XCD a contiguous range of
// the (tile, solve) grid -- whole solves, in a batched launch -- so the L
// and U panels a solve's tiles share are fetched into one L2.  A bijection
// whenever the grid size is a multiple of 8 (otherwise the order is kept).
constexpr int kNumXcd = 8;
__device__ __forceinline__ void xcd_remap(int &bx, int &by) {
  const int gx = gridDim.x;
  const long total = (long)gx * gridDim.y;
  if (total % kNumXcd) return;
  const long lin = bx + (long)gx * by;
  const long per = total / kNumXcd;
  const long l2 = (lin % kNumXcd) * per + lin / kNumXcd;
  bx = (int)(l2 % gx);
  by = (int)(l2 / gx);
}

}  // namespace lsmsx
