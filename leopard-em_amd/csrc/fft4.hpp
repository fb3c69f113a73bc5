// Four-step workgroup FFT for N = 512 * NW (NW = 1, 2, 4, 8 waves of 64 lanes).
//
// N = 8 x S with S = 64 * NW.  Thread t = 64 w + l holds V = 8 values.
//   stage 1   radix-8 over x[t + T i] in registers, twiddle W_N^{k1 t}
//   E1        block-wide LDS exchange: wave w takes sub-transforms k1 = w*SUB + u
//             (SUB = 8/NW), lane l holds their samples n2 = l + 64 j (j < NW)
//   stage 2   radix-NW over j in registers, twiddle W_S^{m1 l}
//   E2 / E3   two exchanges *inside the wave's own LDS rows* (no workgroup barrier):
//             64-point transforms over the lanes as 8 x 8
//   stage 3/4 radix-8, twiddle W_64^{p r}, radix-8
// nat2perm takes natural order in (v[i] = x[t + T i]) and leaves a fixed permutation
// out (v[s] = X[fs_index(t, s)]); perm2nat runs the transposed stage sequence, taking the
// permuted order in and leaving natural order out.  Between the two the data can be
// multiplied element-wise in registers.  Only E1 needs workgroup barriers (one for
// nat2perm, two for perm2nat); E2 / E3 are ordered by the wave's in-order LDS queue.
//
// LDS (per transform, separate real / imaginary float arrays): 8 rows of RS = 72 NW
// floats; wave w's rows [w SUB, (w+1) SUB) are 576 contiguous floats it also uses for
// E2 (row stride 72) and E3 (group stride 72, row stride 9) -- strides chosen so every
// exchange is free of bank conflicts for 32- and 64-bank LDS.
#pragma once
#include "fft.hpp"

namespace tdtm {

// sizes carried by the four-step transform (smaller ones use the Stockham passes)
template <int N>
struct UseFour {
  static constexpr bool value = N >= 512 && N <= 4096;
};

template <int N>
struct FS {
  static_assert(N == 512 || N == 1024 || N == 2048 || N == 4096, "four-step sizes");
  static constexpr int NW = N / 512;
  static constexpr int T = 64 * NW;
  static constexpr int S = N / 8;
  static constexpr int SUB = 8 / NW;
  static constexpr int RS = 72 * NW;
  static constexpr int floats = 8 * RS;  // per component
};

// natural index of slot s held by thread t after nat2perm (and expected by perm2nat)
template <int N>
__device__ __forceinline__ int fs_index(int t, int s) {
  constexpr int NW = FS<N>::NW, SUB = FS<N>::SUB;
  const int w = t >> 6, l = t & 63;
  const int grp = l >> 3, r = l & 7;
  const int u = grp / NW, m1 = grp % NW;
  const int k1 = w * SUB + u;
  return k1 + 8 * (m1 + NW * r + 8 * NW * s);
}

__device__ __forceinline__ void wave_lds_order() {
  // LDS accesses of one wave execute in order; this only stops the compiler from
  // moving them across the exchange
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

template <int R, bool INV>
__device__ __forceinline__ void twiddle_mul(float2 (&a)[R], const float2* __restrict__ tw, int k, int step) {
  if constexpr (R > 1) {
    float2 w[R];
    twiddles<R, INV>(w, tw, k, step);
#pragma unroll
    for (int j = 1; j < R; ++j) a[j] = c_mul(a[j], w[j]);
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2 (&a)[R]) {
  if constexpr (R > 1) dft<R, INV>(a);
}

template <int N, bool INV, int NZJ>
__device__ __forceinline__ void fs_nat2perm(float2 (&v)[8], float* __restrict__ re, float* __restrict__ im, int t,
                                            const float2* __restrict__ tw) {
  using F = FS<N>;
  constexpr int NW = F::NW, SUB = F::SUB, RS = F::RS;
  const int w = t >> 6, l = t & 63;
  const float2* twp = tw;
  asm volatile("" : "+s"(twp));
  // stage 1: radix-8 over i, twiddle W_N^{k1 t}
  dft_sparse<8, NZJ, INV>(v, twp);
  twiddle_mul<8, INV>(v, twp, t, TW_N / N);
  // E1: row k1, column t
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    re[k1 * RS + t] = v[k1].x;
    im[k1 * RS + t] = v[k1].y;
  }
  lds_barrier();
  float2 b[8];
#pragma unroll
  for (int u = 0; u < SUB; ++u)
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const int a = (w * SUB + u) * RS + l + 64 * j;
      b[u * NW + j] = make_float2(re[a], im[a]);
    }
  // stage 2: radix-NW over j, twiddle W_S^{m1 l}
#pragma unroll
  for (int u = 0; u < SUB; ++u) {
    float2 a[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) a[j] = b[u * NW + j];
    dft_any<NW, INV>(a);
    twiddle_mul<NW, INV>(a, twp, l, TW_N / F::S);
#pragma unroll
    for (int j = 0; j < NW; ++j) b[u * NW + j] = a[j];
  }
  // E2 (wave-local): group grp = u NW + m1 at row grp, column l
  float* wr = re + w * 576;
  float* wi = im + w * 576;
  wave_lds_order();
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    wr[g * 72 + l] = b[g].x;
    wi[g * 72 + l] = b[g].y;
  }
  wave_lds_order();
  const int grp = l >> 3, p = l & 7;
#pragma unroll
  for (int q = 0; q < 8; ++q) b[q] = make_float2(wr[grp * 72 + p + 8 * q], wi[grp * 72 + p + 8 * q]);
  // stage 3: radix-8 over q, twiddle W_64^{p r}
  dft<8, INV>(b);
  twiddle_mul<8, INV>(b, twp, p, TW_N / 64);
  // E3 (wave-local): [grp][r][p]
  wave_lds_order();
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    wr[grp * 72 + r * 9 + p] = b[r].x;
    wi[grp * 72 + r * 9 + p] = b[r].y;
  }
  wave_lds_order();
  const int rr = l & 7;
#pragma unroll
  for (int pp = 0; pp < 8; ++pp) v[pp] = make_float2(wr[grp * 72 + rr * 9 + pp], wi[grp * 72 + rr * 9 + pp]);
  // stage 4: radix-8 over p -> slot s
  dft<8, INV>(v);
}

// Transposed sequence: v[s] = Y[fs_index(t, s)] in, v[i] = y[t + T i] out.  Ends with a
// workgroup barrier, so the LDS region is free again on return.
template <int N, bool INV>
__device__ __forceinline__ void fs_perm2nat(float2 (&v)[8], float* __restrict__ re, float* __restrict__ im, int t,
                                            const float2* __restrict__ tw) {
  using F = FS<N>;
  constexpr int NW = F::NW, SUB = F::SUB, RS = F::RS;
  const int w = t >> 6, l = t & 63;
  const float2* twp = tw;
  asm volatile("" : "+s"(twp));
  float* wr = re + w * 576;
  float* wi = im + w * 576;
  const int grp = l >> 3, rr = l & 7;
  // stage 4^T: radix-8 over s -> p
  dft<8, INV>(v);
  // E3^T: lane (grp, r) holds p; lane (grp, p) needs r
  wave_lds_order();
#pragma unroll
  for (int pp = 0; pp < 8; ++pp) {
    wr[grp * 72 + rr * 9 + pp] = v[pp].x;
    wi[grp * 72 + rr * 9 + pp] = v[pp].y;
  }
  wave_lds_order();
  const int p = l & 7;
  float2 b[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) b[r] = make_float2(wr[grp * 72 + r * 9 + p], wi[grp * 72 + r * 9 + p]);
  // stage 3^T: twiddle W_64^{p r}, radix-8 over r -> q
  twiddle_mul<8, INV>(b, twp, p, TW_N / 64);
  dft<8, INV>(b);
  // E2^T: lane (grp, p) holds column p + 8 q of row grp
  wave_lds_order();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    wr[grp * 72 + p + 8 * q] = b[q].x;
    wi[grp * 72 + p + 8 * q] = b[q].y;
  }
  wave_lds_order();
#pragma unroll
  for (int g = 0; g < 8; ++g) b[g] = make_float2(wr[g * 72 + l], wi[g * 72 + l]);
  // stage 2^T: twiddle W_S^{m1 l}, radix-NW over m1 -> j
#pragma unroll
  for (int u = 0; u < SUB; ++u) {
    float2 a[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) a[j] = b[u * NW + j];
    twiddle_mul<NW, INV>(a, twp, l, TW_N / F::S);
    dft_any<NW, INV>(a);
#pragma unroll
    for (int j = 0; j < NW; ++j) b[u * NW + j] = a[j];
  }
  // E1^T: own rows, then all rows after the barrier
  wave_lds_order();
#pragma unroll
  for (int u = 0; u < SUB; ++u)
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const int a = (w * SUB + u) * RS + l + 64 * j;
      re[a] = b[u * NW + j].x;
      im[a] = b[u * NW + j].y;
    }
  lds_barrier();
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) v[k1] = make_float2(re[k1 * RS + t], im[k1 * RS + t]);
  lds_barrier();
  // stage 1^T: twiddle W_N^{k1 t}, radix-8 over k1 -> i
  twiddle_mul<8, INV>(v, twp, t, TW_N / N);
  dft<8, INV>(v);
}

}  // namespace tdtm
