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


template <int R, bool INV, class TWT>
__device__ __forceinline__ void twiddle_mul(float2 (&a)[R], TWT tw, int k, int step) {
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

// NT independent transforms at once (separate LDS regions re[n] / im[n]): twiddles are
// generated once per stage for all of them and the workgroup barriers are shared.
template <int R, bool INV, int NT, class TWT>
__device__ __forceinline__ void twiddle_mul_n(float2 (&a)[NT][R], TWT tw, int k, int step) {
  if constexpr (R > 1) {
    float2 w[R];
    twiddles<R, INV>(w, tw, k, step);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int j = 1; j < R; ++j) a[n][j] = c_mul(a[n][j], w[j]);
  }
}

// Lane-ordered twiddle bases of the four-step transform of size N, for a copy in LDS.
// Each stage needs W^{k}, W^{2k}, W^{4k} (radix 8; fewer for radix NW < 8) for its own
// per-lane k, so they are stored as arrays indexed by k -- lanes read consecutive
// entries (or, in stage 3, broadcast 8 of them), free of bank conflicts:
//   stage 1  A1[e][t] = W_N^{2^e t}   (e < 3, t < T)
//   stage 2  A2[e][l] = W_S^{2^e l}   (e < log2 NW, l < 64)
//   stage 3  A3[e][p] = W_64^{2^e p}  (e < 3, p < 8)
template <int N>
struct FsTwLayout {
  static constexpr int T = FS<N>::T, NW = FS<N>::NW, S = FS<N>::S;
  static constexpr int LNW = NW == 1 ? 0 : NW == 2 ? 1 : NW == 4 ? 2 : 3;
  static constexpr int O1 = 0, O2 = 3 * T, O3 = O2 + LNW * 64;
  static constexpr int entries = O3 + 3 * 8;
  static constexpr int floats = 2 * entries;
  // table index of entry j (every index is a multiple of TW_N / N below TW_N / 2)
  __device__ static int index(int j) {
    if (j < O2) return ((j % T) << (j / T)) * (TW_N / N);
    if (j < O3) return (((j - O2) % 64) << ((j - O2) / 64)) * (TW_N / S);
    return (((j - O3) % 8) << ((j - O3) / 8)) * (TW_N / 64);
  }
};

template <int N>
struct FsTw {
  const float2* p;  // LDS copy in FsTwLayout<N> order
};

template <int N>
__device__ __forceinline__ FsTw<N> opaque_offset(FsTw<N> t) {
  return FsTw<N>{opaque_offset(t.p)};
}

// Fill an FsTwLayout<N> copy from the global table (all threads of the block; the caller
// synchronises before use).
template <int N>
__device__ __forceinline__ void fs_tw_fill(float2* dst, const float2* __restrict__ tw, int tid, int nthreads) {
  for (int j = tid; j < FsTwLayout<N>::entries; j += nthreads) dst[j] = tw[FsTwLayout<N>::index(j)];
}

template <class TWT>
struct IsFsTw {
  static constexpr bool value = false;
};
template <int N>
struct IsFsTw<FsTw<N>> {
  static constexpr bool value = true;
};

// The twiddles of the three twiddled stages of one lane, held in registers (w[1..7] of
// stage 1, w[1..NW-1] of stage 2, w[1..7] of stage 3; w[0] = 1).  They depend on the
// lane only, so a kernel that runs many transforms builds them once (fs_tw_regs) and
// saves the per-transform loads and products; the inverse uses their conjugates.
template <int N>
struct FsTwReg {
  float2 w1[7], w2[7], w3[7];
};

template <int N>
__device__ __forceinline__ FsTwReg<N> opaque_offset(const FsTwReg<N>& t) {
  return t;
}

template <class TWT>
struct IsFsTwReg {
  static constexpr bool value = false;
};
template <int N>
struct IsFsTwReg<FsTwReg<N>> {
  static constexpr bool value = true;
};

// Twiddles of four-step stage STAGE (1: W_N^{j t}, 2: W_S^{j l}, 3: W_64^{j p}) for lane
// index k: w[j] for j < R, from a lane-ordered LDS copy or any indexable table.
template <int STAGE, int N, int R, bool INV, class TWT>
__device__ __forceinline__ void fs_tw(float2 (&w)[R], TWT tw, int k) {
  using L = FsTwLayout<N>;
  constexpr int step = STAGE == 1 ? TW_N / N : STAGE == 2 ? TW_N / L::S : TW_N / 64;
  if constexpr (IsFsTwReg<TWT>::value) {
    (void)k;
    const float2* src = STAGE == 1 ? tw.w1 : STAGE == 2 ? tw.w2 : tw.w3;
    w[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int j = 1; j < R; ++j) w[j] = INV ? c_conj(src[j - 1]) : src[j - 1];
  } else if constexpr (IsFsTw<TWT>::value) {
    static_assert(STAGE != 2 || (R == L::NW), "stage 2 is radix NW");
    constexpr int off = STAGE == 1 ? L::O1 : STAGE == 2 ? L::O2 : L::O3;
    constexpr int cnt = STAGE == 1 ? L::T : STAGE == 2 ? 64 : 8;
    auto ld = [&](int e) {
      const float2 v = tw.p[off + e * cnt + k];
      return INV ? c_conj(v) : v;
    };
    w[0] = make_float2(1.f, 0.f);
    if constexpr (R >= 2) w[1] = ld(0);
    if constexpr (R >= 4) {
      w[2] = ld(1);
      w[3] = c_mul(w[1], w[2]);
    }
    if constexpr (R >= 8) {
      w[4] = ld(2);
      w[5] = c_mul(w[1], w[4]);
      w[6] = c_mul(w[2], w[4]);
      w[7] = c_mul(w[3], w[4]);
    }
  } else {
    twiddles<R, INV>(w, tw, k, step);
  }
}

// Registers of lane t from any table fs_tw reads (the same loads and products, so the
// transforms round identically with either form).
template <int N, class TWT>
__device__ __forceinline__ FsTwReg<N> fs_tw_regs(TWT tw, int t) {
  constexpr int NW = FS<N>::NW;
  const int l = t & 63, p = l & 7;
  FsTwReg<N> r;
  float2 w[8];
  fs_tw<1, N, 8, false>(w, tw, t);
#pragma unroll
  for (int j = 1; j < 8; ++j) r.w1[j - 1] = w[j];
  float2 w2[NW];
  fs_tw<2, N, NW, false>(w2, tw, l);
#pragma unroll
  for (int j = 1; j < 8; ++j) r.w2[j - 1] = j < NW ? w2[j < NW ? j : 0] : make_float2(1.f, 0.f);
  fs_tw<3, N, 8, false>(w, tw, p);
#pragma unroll
  for (int j = 1; j < 8; ++j) r.w3[j - 1] = w[j];
  return r;
}

template <int R, int NT>
__device__ __forceinline__ void apply_tw_n(float2 (&a)[NT][R], const float2 (&w)[R]) {
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int j = 1; j < R; ++j) a[n][j] = c_mul(a[n][j], w[j]);
}

// LDS element access of transform n: split real / imaginary float arrays re[n] / im[n],
// or (IL) one interleaved float2 array starting at re[n] (im unused) -- same indices, one
// 8-byte access per element instead of two 4-byte ones.
template <bool IL, int NT>
struct FsLds {
  float* const (&re)[NT];
  float* const (&im)[NT];
  __device__ __forceinline__ void st(int n, int i, float2 v) const {
    if constexpr (IL) {
      reinterpret_cast<float2*>(re[n])[i] = v;
    } else {
      re[n][i] = v.x;
      im[n][i] = v.y;
    }
  }
  __device__ __forceinline__ float2 ld(int n, int i) const {
    if constexpr (IL)
      return reinterpret_cast<const float2*>(re[n])[i];
    else
      return make_float2(re[n][i], im[n][i]);
  }
};

template <int N, bool INV, int NZJ, int NT, bool IL = false, class TWT>
__device__ __forceinline__ void fs_nat2perm_n(float2 (&v)[NT][8], float* const (&re)[NT], float* const (&im)[NT],
                                              int t, TWT tw) {
  using F = FS<N>;
  constexpr int NW = F::NW, SUB = F::SUB, RS = F::RS;
  const int w = t >> 6, l = t & 63;
  const FsLds<IL, NT> L{re, im};
  const auto twp = opaque_offset(tw);
  // stage 1: radix-8 over i, twiddle W_N^{k1 t}
#pragma unroll
  for (int n = 0; n < NT; ++n) dft_sparse<8, NZJ, INV>(v[n], twp);
  {
    float2 w1[8];
    fs_tw<1, N, 8, INV>(w1, twp, t);
    apply_tw_n<8, NT>(v, w1);
  }
  // E1: row k1, column t
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
      L.st(n, k1 * RS + t, v[n][k1]);
    }
  lds_barrier();
  float2 b[NT][8];
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int u = 0; u < SUB; ++u)
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const int a = (w * SUB + u) * RS + l + 64 * j;
        b[n][u * NW + j] = L.ld(n, a);
      }
  // stage 2: radix-NW over j, twiddle W_S^{m1 l}
  {
    float2 wt[NW];
    fs_tw<2, N, NW, INV>(wt, twp, l);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int u = 0; u < SUB; ++u) {
        float2 a[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) a[j] = b[n][u * NW + j];
        dft_any<NW, INV>(a);
#pragma unroll
        for (int j = 1; j < NW; ++j) a[j] = c_mul(a[j], wt[j]);
#pragma unroll
        for (int j = 0; j < NW; ++j) b[n][u * NW + j] = a[j];
      }
  }
  // E2 (wave-local): group grp = u NW + m1 at row grp, column l
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int g = 0; g < 8; ++g) L.st(n, w * 576 + g * 72 + l, b[n][g]);
  }
  wave_lds_order();
  const int grp = l >> 3, p = l & 7;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int q = 0; q < 8; ++q) b[n][q] = L.ld(n, w * 576 + grp * 72 + p + 8 * q);
  }
  // stage 3: radix-8 over q, twiddle W_64^{p r}
#pragma unroll
  for (int n = 0; n < NT; ++n) dft<8, INV>(b[n]);
  {
    float2 w3[8];
    fs_tw<3, N, 8, INV>(w3, twp, p);
    apply_tw_n<8, NT>(b, w3);
  }
  // E3 (wave-local): [grp][r][p]
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int r = 0; r < 8; ++r) L.st(n, w * 576 + grp * 72 + r * 9 + p, b[n][r]);
  }
  wave_lds_order();
  const int rr = l & 7;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int pp = 0; pp < 8; ++pp) v[n][pp] = L.ld(n, w * 576 + grp * 72 + rr * 9 + pp);
  }
  // stage 4: radix-8 over p -> slot s
#pragma unroll
  for (int n = 0; n < NT; ++n) dft<8, INV>(v[n]);
}

// TAIL = false leaves out the closing barrier after the E1 loads.  A caller may do that
// when the next LDS writes of the transform area are the next transform's E1 stores:
// thread t stores exactly the addresses it loaded here (row k1, column t), and every
// other write (the wave-local E2 / E3 rows) comes after that transform's E1 barrier.
template <int N, bool INV, int NT, bool IL = false, bool TAIL = true, class TWT>
__device__ __forceinline__ void fs_perm2nat_n(float2 (&v)[NT][8], float* const (&re)[NT], float* const (&im)[NT],
                                              int t, TWT tw) {
  using F = FS<N>;
  constexpr int NW = F::NW, SUB = F::SUB, RS = F::RS;
  const int w = t >> 6, l = t & 63;
  const FsLds<IL, NT> L{re, im};
  const auto twp = opaque_offset(tw);
  const int grp = l >> 3, rr = l & 7;
  // stage 4^T: radix-8 over s -> p
#pragma unroll
  for (int n = 0; n < NT; ++n) dft<8, INV>(v[n]);
  // E3^T: lane (grp, r) holds p; lane (grp, p) needs r
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int pp = 0; pp < 8; ++pp) L.st(n, w * 576 + grp * 72 + rr * 9 + pp, v[n][pp]);
  }
  wave_lds_order();
  const int p = l & 7;
  float2 b[NT][8];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int r = 0; r < 8; ++r) b[n][r] = L.ld(n, w * 576 + grp * 72 + r * 9 + p);
  }
  // stage 3^T: twiddle W_64^{p r}, radix-8 over r -> q
  {
    float2 w3[8];
    fs_tw<3, N, 8, INV>(w3, twp, p);
    apply_tw_n<8, NT>(b, w3);
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) dft<8, INV>(b[n]);
  // E2^T: lane (grp, p) holds column p + 8 q of row grp
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int q = 0; q < 8; ++q) L.st(n, w * 576 + grp * 72 + p + 8 * q, b[n][q]);
  }
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int g = 0; g < 8; ++g) b[n][g] = L.ld(n, w * 576 + g * 72 + l);
  }
  // stage 2^T: twiddle W_S^{m1 l}, radix-NW over m1 -> j
  {
    float2 wt[NW];
    fs_tw<2, N, NW, INV>(wt, twp, l);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int u = 0; u < SUB; ++u) {
        float2 a[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) a[j] = b[n][u * NW + j];
#pragma unroll
        for (int j = 1; j < NW; ++j) a[j] = c_mul(a[j], wt[j]);
        dft_any<NW, INV>(a);
#pragma unroll
        for (int j = 0; j < NW; ++j) b[n][u * NW + j] = a[j];
      }
  }
  // E1^T: own rows, then all rows after the barrier
  wave_lds_order();
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int u = 0; u < SUB; ++u)
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const int a = (w * SUB + u) * RS + l + 64 * j;
        L.st(n, a, b[n][u * NW + j]);
      }
  lds_barrier();
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) v[n][k1] = L.ld(n, k1 * RS + t);
  if constexpr (TAIL) lds_barrier();
  // stage 1^T: twiddle W_N^{k1 t}, radix-8 over k1 -> i
  {
    float2 w1[8];
    fs_tw<1, N, 8, INV>(w1, twp, t);
    apply_tw_n<8, NT>(v, w1);
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) dft<8, INV>(v[n]);
}

// single-transform forms
template <int N, bool INV, int NZJ, bool IL = false, class TWT>
__device__ __forceinline__ void fs_nat2perm(float2 (&v)[8], float* __restrict__ re, float* __restrict__ im, int t,
                                            TWT tw) {
  float2(&vv)[1][8] = *reinterpret_cast<float2(*)[1][8]>(&v);
  float* const rr[1] = {re};
  float* const ii[1] = {im};
  fs_nat2perm_n<N, INV, NZJ, 1, IL>(vv, rr, ii, t, tw);
}

template <int N, bool INV, bool IL = false, class TWT>
__device__ __forceinline__ void fs_perm2nat(float2 (&v)[8], float* __restrict__ re, float* __restrict__ im, int t,
                                            TWT tw) {
  float2(&vv)[1][8] = *reinterpret_cast<float2(*)[1][8]>(&v);
  float* const rr[1] = {re};
  float* const ii[1] = {im};
  fs_perm2nat_n<N, INV, 1, IL>(vv, rr, ii, t, tw);
}

}  // namespace tdtm
