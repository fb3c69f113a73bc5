// Column pass on the matrix cores: the 4096-point column transforms of the search as
// radix-16 DFT matrix products on v_mfma_f32_16x16x32_{f16,bf16} (gfx950).
//
// A 4096-point column is a 16 x 16 x 16 array, position pos = d0 + 16 d1 + 256 d2.  Every
// stage is a dense 16-point DFT along one axis for 256 batch columns at once: one MFMA
// group = 16 batch columns x 16 points, the complex DFT written as the 32 x 32 real
// matrix [[Fr, -Fi], [Fi, Fr]] acting on stacked (re, im) inputs and split into two
// 16-row halves (real and imaginary outputs).  The DFT matrix is split once per block
// into hi + lo parts of the element type; the data are rounded to the element type once
// per stage (two products), or with SPLIT also split into hi + lo (three products: the
// lo x lo term is left out).  Accumulation is FP32.
//
// MFMA lane layout (16x16x32): lane l = j + 16 g (j = l & 15, g = l >> 4) holds row j of
// A and column j of B for the same 8 K values (slots s = 0..7), and rows 4 g + r
// (r = 0..3) of column j of the result.  The K order is fixed by the data: slot s is the
// real (s even) or imaginary (s odd) part of input point p = 4 g + s / 2, i.e. one
// packed (re, im) pair per 32-bit word; the A matrices are built in that order, so a
// lane's four result rows feed the next DFT along the same axis as its B operand with
// no exchange (the forward transform's last stage and the inverse transform's first).
//
// Forward (input x[n] non-zero for n < 512, i.e. d2 in {0, 1}: one MFMA input group),
// with n = a + 16 b + 256 c and X[m0 + 16 m1 + 256 m2]:
//   F1  over c (slow axis) -> m0,   twiddle W_4096^{m0 (a + 16 b)}
//   F2  over b (middle)    -> m1,   twiddle W_256^{a m1}
//   F3  over a (fast)      -> m2    (output digit-reversed: position (m2, m1, m0))
// multiply by the image spectrum (conjugate of the template spectrum), then inverse:
//   I1  over m2 (fast)     -> n0,   twiddle W_256^{-m1 n0}
//   I2  over m1 (middle)   -> n1,   twiddle W_4096^{-m0 (n0 + 16 n1)}
//   I3  over m0 (slow)     -> n2    (natural order n0 + 16 n1 + 256 n2)
// Wave q of the block runs MFMA group q of every stage for both columns of its tile
// column, so every twiddle a lane applies depends on its lane and q only: all three
// twiddle sets are built once per block in registers.  The inverse stages use the forward matrices on conjugated
// data, IDFT(y) = conj(DFT(conj(y))): every inverse stage's input is stored conjugated (the
// previous stage writes conj(T w) = D conj(w) for its own conjugated result D), and the
// last stage's output is conjugated back.
//
// Between stages the data go through an LDS array of one float32 complex (8 bytes, two
// banks) per position, stored at swz(pos): bit 2 XOR bit 5, bit 3 XOR bit 7, bit 4 XOR
// bits 6 and 10.  With 32 two-bank slots per 256 bytes, every half wave's accesses along
// the middle axis (lanes: d0 and bit 2 of d1) and the slow axis (d0 and bit 2 of d2), and
// 32 consecutive positions (the tile stores), fall into distinct slots; the 16-byte
// accesses along the fast axis (16 lanes: d1) meet at most two per slot.  Positions keep
// bits 0-1, so the four consecutive elements of a fast-axis access stay in order.
//
// Ranges (fp16): the column input is scaled by 2^-6 and the image spectrum by 2^6.  For
// a projection normalised to sum(p^2) = H W (normalize_template_projection) every
// partial DFT sum of a column is bounded by sqrt(h^2 H W) (Cauchy-Schwarz), i.e. below
// 2^15 after the scaling for h = 512 and H W <= 2^25; the products of the two scales is
// one, so the output needs no rescaling.  bf16 has the float32 exponent range.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace tdtm {

typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));
typedef _Float16 mf_h2 __attribute__((ext_vector_type(2)));
typedef __bf16 mf_b2 __attribute__((ext_vector_type(2)));

constexpr int MF_N = 4096;          // column length carried by the matrix-core pass
constexpr int MF_COLS = 2;          // columns per block (one tile column)
constexpr int MF_BLOCK = 16 * 64;   // one wave per MFMA group q = 0..15, both columns
constexpr int MF_CW = 2;            // columns per tile of the intermediate
constexpr int MF_TR = 4;            // rows per tile of the intermediate

// element type traits: the 8-element operand vector, its 2-element pack and the MFMA
template <class ET>
struct MfType;
template <>
struct MfType<_Float16> {
  using v8 = mf_h8;
  using v2 = mf_h2;
  static __device__ __forceinline__ mf_f4 mfma(v8 a, v8 b, mf_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct MfType<__bf16> {
  using v8 = mf_b8;
  using v2 = mf_b2;
  static __device__ __forceinline__ mf_f4 mfma(v8 a, v8 b, mf_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

__device__ __forceinline__ int mf_swz(int p) {
  return p ^ (((p >> 5) & 1) << 2) ^ (((p >> 7) & 1) << 3) ^ ((((p >> 6) ^ (p >> 10)) & 1) << 4);
}
__device__ __forceinline__ int mf_pos(int d0, int d1, int d2) { return d0 + 16 * d1 + 256 * d2; }

// The two 16-row halves (real, imaginary outputs) of one DFT matrix, hi and lo parts,
// in a lane's A-operand order.
template <class ET>
struct MfMat {
  typename MfType<ET>::v8 re_hi, re_lo, im_hi, im_lo;
};

// DFT-16 matrix rows of lane row j: forward W_16^{jp} (INV: conjugate), from the table
// tw[m] = e^{-2 pi i m / TW_N}.
template <class ET, bool INV>
__device__ __forceinline__ MfMat<ET> mf_matrix(const float2* __restrict__ tw, int j, int g) {
  MfMat<ET> m;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int p = 4 * g + s / 2;
    const float2 w = tw[(TW_N / 16) * ((j * p) & 15)];
    const float fr = w.x, fi = INV ? -w.y : w.y;
    // real output row: Fr re - Fi im; imaginary output row: Fi re + Fr im
    const float vre = (s & 1) ? -fi : fr;
    const float vim = (s & 1) ? fr : fi;
    const ET hre = (ET)vre, him = (ET)vim;
    m.re_hi[s] = hre;
    m.re_lo[s] = (ET)(vre - (float)hre);
    m.im_hi[s] = him;
    m.im_lo[s] = (ET)(vim - (float)him);
  }
  return m;
}


// B operand of four complex points (slot 2i = re, 2i + 1 = im of point i): rounded to
// the element type (hi), and with SPLIT also the residual (lo = v - hi, rounded).
template <class ET, bool SPLIT>
struct MfOp {
  typename MfType<ET>::v8 hi, lo;
};

template <class ET, bool SPLIT>
__device__ __forceinline__ MfOp<ET, SPLIT> mf_operand(const float2 (&v)[4]) {
  MfOp<ET, SPLIT> o;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const ET hr = (ET)v[i].x, hi = (ET)v[i].y;
    o.hi[2 * i] = hr;
    o.hi[2 * i + 1] = hi;
    if constexpr (SPLIT) {
      o.lo[2 * i] = (ET)(v[i].x - (float)hr);
      o.lo[2 * i + 1] = (ET)(v[i].y - (float)hi);
    }
  }
  return o;
}

// One DFT-16 group: A (hi + lo) x B (hi [+ lo]); the lo x lo product is left out.
template <class ET, bool SPLIT>
__device__ __forceinline__ void mf_dft(const MfMat<ET>& A, const MfOp<ET, SPLIT>& b, mf_f4& dre, mf_f4& dim) {
  const mf_f4 z = {0.f, 0.f, 0.f, 0.f};
  dre = MfType<ET>::mfma(A.re_hi, b.hi, z);
  dre = MfType<ET>::mfma(A.re_lo, b.hi, dre);
  dim = MfType<ET>::mfma(A.im_hi, b.hi, z);
  dim = MfType<ET>::mfma(A.im_lo, b.hi, dim);
  if constexpr (SPLIT) {
    dre = MfType<ET>::mfma(A.re_hi, b.lo, dre);
    dim = MfType<ET>::mfma(A.im_hi, b.lo, dim);
  }
}

// Column pass on the matrix cores for H = 4096: block = one tile column (two image
// columns), MF_NWC waves per column; persistent like k_colpass (blocks b and b + 8 take
// tile columns 2m and 2m + 1 of an equal share of the m-major (tile-column pair,
// correlation) items).  Output: the same tiled half-spectrum intermediate as k_colpass.
// The exchange array holds float32 complex values; each stage rounds (and with SPLIT
// splits) its B operand when it reads them.  The last stage writes its result in place,
// which is the natural order at swz(n), and the tile stores read it from there.
template <class ET, bool SPLIT>
__global__ __launch_bounds__(MF_BLOCK) void k_colpass_mfma(const float2* __restrict__ Rb,
                                                          const float2* __restrict__ Xt, int hs, int G, int Gs, int K,
                                                          const float2* __restrict__ tw, float2* __restrict__ Gb,
                                                          int persist) {
  constexpr int HH = MF_N;
  __shared__ __attribute__((aligned(16))) float2 xch[MF_COLS][HH];
  const int tid = threadIdx.x;
  const int q = tid >> 6;  // this wave's MFMA group
  const int l = tid & 63, j = l & 15, g = l >> 4;
  const MfMat<ET> Af = mf_matrix<ET, false>(tw, j, g);
  // twiddles of this lane (W_L^e = tw[e TW_N / L]):
  //   F1: W_4096^{m0 (a + 16 b)}, m0 = 4 g + r, a = j, b = q
  //   F2 / I1: W_256^{j (4 g + r)}
  //   I2: W_4096^{m0 (n0 + 16 n1)}, m0 = q, n0 = j, n1 = 4 g + r (applied to conjugated data)
  float2 t1[4], t2[4], t3[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    t1[r] = tw[(TW_N / 4096) * (((4 * g + r) * (j + 16 * q)) & 4095)];
    t2[r] = tw[(TW_N / 256) * ((j * (4 * g + r)) & 255)];
    t3[r] = tw[(TW_N / 4096) * ((q * (j + 16 * (4 * g + r))) & 4095)];
  }
  // exact powers of two that keep fp16 values in range (see the header); harmless for bf16
  const float in_scale = 0.015625f, img_scale = 64.f;

  auto segment = [&](int gg, int k0, int k1) {
    // image columns gg, c in the F3 output order: ky = q + 16 j + 256 (4 g + r)
    float2 xi[MF_COLS][4];
#pragma unroll
    for (int c = 0; c < MF_COLS; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ky = q + 16 * j + 256 * (4 * g + r);
        xi[c][r] = c_scale(Xt[((long)gg * HH + ky) * MF_CW + c], img_scale);
      }
    const bool dc = gg == 0 && q == 0 && j == 0 && g == 0;  // (ky, kx) = (0, 0) in column 0, row r = 0
    // column inputs x[j + 16 q + 256 cd], cd = 0, 1 (lanes g = 0 only)
    float2 inx[MF_COLS][2];
    auto load_in = [&](int k) {
      const float2* rin = Rb + ((long)k * G + gg) * hs * MF_CW;
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c)
#pragma unroll
        for (int cd = 0; cd < 2; ++cd) {
          const int n = j + 16 * q + 256 * cd;
          inx[c][cd] = (g == 0 && n < hs) ? c_scale(rin[(long)n * MF_CW + c], in_scale) : make_float2(0.f, 0.f);
        }
    };
    load_in(k0);
    for (int k = k0; k < k1; ++k) {
      // F1 over d2 (pruned: points 0, 1 of 16)
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c) {
        float2 v[4] = {inx[c][0], inx[c][1], make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        mf_f4 dre, dim;
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
#pragma unroll
        for (int r = 0; r < 4; ++r) xch[c][mf_swz(mf_pos(j, q, 4 * g + r))] = c_mul(make_float2(dre[r], dim[r]), t1[r]);
      }
      if (k + 1 < k1) load_in(k + 1);  // next correlation's inputs, in flight over the stages below
      __syncthreads();
      // F2 over d1: batch (a = j, m0 = q), twiddle W_256^{a m1}
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c) {
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = xch[c][mf_swz(mf_pos(j, 4 * g + t, q))];
        mf_f4 dre, dim;
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
#pragma unroll
        for (int r = 0; r < 4; ++r) xch[c][mf_swz(mf_pos(j, 4 * g + r, q))] = c_mul(make_float2(dre[r], dim[r]), t2[r]);
      }
      __syncthreads();
      // F3 over d0 -> X[q + 16 j + 256 (4 g + r)]; times the image conjugate; I1 over d0;
      // twiddle W_256^{-m1 n0}, m1 = j
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c) {
        float4* p4 = reinterpret_cast<float4*>(&xch[c][mf_swz(mf_pos(4 * g, j, q))]);
        float2 v[4];
        {
          const float4 a = p4[0], b = p4[1];
          v[0] = make_float2(a.x, a.y);
          v[1] = make_float2(a.z, a.w);
          v[2] = make_float2(b.x, b.y);
          v[3] = make_float2(b.z, b.w);
        }
        mf_f4 dre, dim;
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // conj(image x conj(X)) = X conj(image): the inverse input, conjugated
          v[r] = c_mulc(make_float2(dre[r], dim[r]), xi[c][r]);
          if (c == 0 && r == 0 && dc) v[r] = make_float2(0.f, 0.f);
        }
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
        // D = conj(T); the next stage takes conj(T W_256^{-m1 n0}) = D W_256^{m1 n0}
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = c_mul(make_float2(dre[r], dim[r]), t2[r]);
        p4[0] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
        p4[1] = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
      }
      __syncthreads();
      // I2 over d1: batch (n0 = j, m0 = q); conj(T W_4096^{-m0 (n0 + 16 n1)}) = D t3
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c) {
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = xch[c][mf_swz(mf_pos(j, 4 * g + t, q))];
        mf_f4 dre, dim;
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
#pragma unroll
        for (int r = 0; r < 4; ++r) xch[c][mf_swz(mf_pos(j, 4 * g + r, q))] = c_mul(make_float2(dre[r], dim[r]), t3[r]);
      }
      __syncthreads();
      // I3 over d2: batch (n0 = j, n1 = q) -> y[j + 16 q + 256 (4 g + r)], written in place
      // (natural order n at swz(n)) and conjugated back
#pragma unroll
      for (int c = 0; c < MF_COLS; ++c) {
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = xch[c][mf_swz(mf_pos(j, q, 4 * g + t))];
        mf_f4 dre, dim;
        mf_dft<ET, SPLIT>(Af, mf_operand<ET, SPLIT>(v), dre, dim);
#pragma unroll
        for (int r = 0; r < 4; ++r) xch[c][mf_swz(mf_pos(j, q, 4 * g + r))] = make_float2(dre[r], -dim[r]);
      }
      __syncthreads();
      // 16-byte pieces (rows 2h, 2h+1 of one column) of the column-major tiles of the
      // intermediate: 4 consecutive threads fill a tile
      float4* gout = reinterpret_cast<float4*>(Gb + (long)k * HH * Gs * MF_CW) + (long)gg * (MF_TR * MF_CW / 2);
#pragma unroll
      for (int jj = 0; jj < HH / MF_BLOCK; ++jj) {
        const int u = tid + jj * MF_BLOCK;  // piece: tile row u / 4, column (u / 2) % 2, rows 2 (u % 2) + 0, 1
        const int trow = u / (MF_TR / 2 * MF_CW), col = (u / (MF_TR / 2)) % MF_CW, h2 = u % (MF_TR / 2);
        const int y = trow * MF_TR + 2 * h2;
        const float2 a = xch[col][mf_swz(y)], b = xch[col][mf_swz(y + 1)];
        gout[(long)trow * Gs * (MF_TR * MF_CW / 2) + (col * MF_TR + 2 * h2) / 2] = make_float4(a.x, a.y, b.x, b.y);
      }
      // the next F1 writes the exchange arrays: every thread's tile reads come first
      __syncthreads();
    }
  };
  const int b = blockIdx.x, qb = b >> 3;
  const int half = qb & 1, pair = (qb >> 1) * 8 + (b & 7);
  const long total = (long)((G + 1) / 2) * K;
  long item = (long)pair * persist;
  const long iend = min(total, item + persist);
  while (item < iend) {
    const int m = (int)(item / K), k0 = (int)(item % K);
    const int k1 = (int)min((long)K, k0 + (iend - item));
    if (2 * m + half < G) segment(2 * m + half, k0, k1);  // uniform over the block
    item += k1 - k0;
  }
}

}  // namespace tdtm
