// Workgroup-level Stockham FFT building blocks for gfx950.
//
// A size-N transform is carried by T = N/V threads, each holding V complex values in
// registers.  Between radix passes the values are exchanged through LDS, stored as
// separate real / imaginary float arrays with one pad word per 16 (index i -> i + i/16),
// which keeps the Stockham scatter of every pass free of LDS bank conflicts.
//
// Register contract (natural order on both sides):
//   on entry  v[i] = x[t + i*T],   on exit  v[i] = X[t + i*T],   i < V.
// so the last pass of one transform and the first pass of the next need no exchange.
//
// Sign convention: forward X[k] = sum x[n] e^{-2 pi i nk/N}; inverse uses e^{+...}
// and is unnormalised.  Twiddles come from a table tw[m] = e^{-2 pi i m / TW_N}
// (computed in double on the host), so every twiddle is either a table entry or the
// product of at most three of them.
#pragma once
#include <hip/hip_runtime.h>

namespace tdtm {

constexpr int TW_N = 8192;  // twiddle table length (largest supported row length)

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Products are written with explicit fused multiply-adds and the library is built with
// -ffp-contract=off: every code shape (launch options, paired transforms) then rounds
// identically, so results do not depend on the launch shape.
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 c_mulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// multiply by the quarter-turn root: -i for the forward transform, +i for the inverse
template <bool INV>
__device__ __forceinline__ float2 c_rot(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// multiply by W8^1 (forward e^{-i pi/4}, inverse e^{+i pi/4})
template <bool INV>
__device__ __forceinline__ float2 c_w8(float2 a) {
  const float r = 0.70710678118654752440f;
  return INV ? make_float2((a.x - a.y) * r, (a.x + a.y) * r) : make_float2((a.x + a.y) * r, (a.y - a.x) * r);
}
// multiply by W8^3
template <bool INV>
__device__ __forceinline__ float2 c_w83(float2 a) {
  const float r = 0.70710678118654752440f;
  return INV ? make_float2(-(a.x + a.y) * r, (a.x - a.y) * r) : make_float2((a.y - a.x) * r, -(a.x + a.y) * r);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = c_add(t, a1);
  a1 = c_sub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
  float2 t2 = c_add(a1, a3), t3 = c_rot<INV>(c_sub(a1, a3));
  a0 = c_add(t0, t2);
  a2 = c_sub(t0, t2);
  a1 = c_add(t1, t3);
  a3 = c_sub(t1, t3);
}

// radix-8, decimation in time: two radix-4 on even / odd inputs, then W8^k
template <bool INV>
__device__ __forceinline__ void dft8(float2 (&a)[8]) {
  dft4<INV>(a[0], a[2], a[4], a[6]);
  dft4<INV>(a[1], a[3], a[5], a[7]);
  // W8^1 a3 = r u1 and W8^3 a7 = r u3 with the unscaled u formed first; y = base +- r u
  // as explicit fused multiply-adds
  const float r = 0.70710678118654752440f;
  const float2 u1 = INV ? make_float2(a[3].x - a[3].y, a[3].x + a[3].y)
                        : make_float2(a[3].x + a[3].y, a[3].y - a[3].x);
  const float2 u3 = INV ? make_float2(-(a[7].x + a[7].y), a[7].x - a[7].y)
                        : make_float2(a[7].y - a[7].x, -(a[7].x + a[7].y));
  float2 o2 = c_rot<INV>(a[5]);
  float2 y0 = c_add(a[0], a[1]), y4 = c_sub(a[0], a[1]);
  float2 y1 = make_float2(fmaf(u1.x, r, a[2].x), fmaf(u1.y, r, a[2].y));
  float2 y5 = make_float2(fmaf(-u1.x, r, a[2].x), fmaf(-u1.y, r, a[2].y));
  float2 y2 = c_add(a[4], o2), y6 = c_sub(a[4], o2);
  float2 y3 = make_float2(fmaf(u3.x, r, a[6].x), fmaf(u3.y, r, a[6].y));
  float2 y7 = make_float2(fmaf(-u3.x, r, a[6].x), fmaf(-u3.y, r, a[6].y));
  a[0] = y0; a[1] = y1; a[2] = y2; a[3] = y3;
  a[4] = y4; a[5] = y5; a[6] = y6; a[7] = y7;
}

// radix-16 as 4 x 4 with internal twiddles W16^{rs}
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&a)[16]) {
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
  const float2 w1 = INV ? make_float2(c1, s1) : make_float2(c1, -s1);
  const float2 w3 = INV ? make_float2(s1, c1) : make_float2(s1, -c1);
#pragma unroll
  for (int r = 0; r < 4; ++r) dft4<INV>(a[r], a[r + 4], a[r + 8], a[r + 12]);
  // B_r[s] lives at a[r + 4 s]
  a[5] = c_mul(a[5], w1);            // r1 s1 -> W^1
  a[9] = c_w8<INV>(a[9]);            // r1 s2 -> W^2
  a[13] = c_mul(a[13], w3);          // r1 s3 -> W^3
  a[6] = c_w8<INV>(a[6]);            // r2 s1 -> W^2
  a[10] = c_rot<INV>(a[10]);         // r2 s2 -> W^4
  a[14] = c_w83<INV>(a[14]);         // r2 s3 -> W^6
  a[7] = c_mul(a[7], w3);            // r3 s1 -> W^3
  a[11] = c_w83<INV>(a[11]);         // r3 s2 -> W^6
  a[15] = c_mul(a[15], make_float2(-w1.x, -w1.y));  // r3 s3 -> W^9 = -W^1
#pragma unroll
  for (int s = 0; s < 4; ++s) dft4<INV>(a[4 * s], a[4 * s + 1], a[4 * s + 2], a[4 * s + 3]);
  // y[s + 4 q] = a[4 s + q]  -> transpose the 4 x 4 register block
  float2 t[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) t[i] = a[i];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) a[s + 4 * q] = t[4 * s + q];
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&a)[R]) {
  if constexpr (R == 2) {
    dft2<INV>(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    dft8<INV>(a);
  } else {
    static_assert(R == 16, "radix");
    dft16<INV>(a);
  }
}

// twiddle w^j (j = 1..R-1) for w = W_L^k, L = TW_N / step; products of <= 3 table entries
template <int R, bool INV, class TWT>
__device__ __forceinline__ void twiddles(float2 (&w)[R], TWT tw, int k, int step) {
  auto ld = [&](int m) {
    float2 v = tw[m];
    return INV ? c_conj(v) : v;
  };
  w[0] = make_float2(1.f, 0.f);
  if constexpr (R >= 2) w[1] = ld(k * step);
  if constexpr (R >= 4) {
    w[2] = ld(2 * k * step);
    w[3] = c_mul(w[1], w[2]);
  }
  if constexpr (R >= 8) {
    w[4] = ld(4 * k * step);
    w[5] = c_mul(w[1], w[4]);
    w[6] = c_mul(w[2], w[4]);
    w[7] = c_mul(w[3], w[4]);
  }
  if constexpr (R >= 16) {
    w[8] = ld(8 * k * step);
#pragma unroll
    for (int j = 1; j < 8; ++j) w[8 + j] = c_mul(w[j], w[8]);
  }
}

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }

// p + 0 through an empty asm: the compiler cannot hoist loads through the result out of
// the caller's loop (which would keep their values live in registers), while the
// pointer keeps its provenance (a kernel argument, i.e. global memory), so the loads are
// global_load and not flat_load (a flat load counts against lgkmcnt and would make every
// LDS wait also wait for it).
template <class T>
__device__ __forceinline__ const T* opaque_offset(const T* p) {
  int off = 0;
  asm volatile("" : "+s"(off));
  return p + off;
}

// Twiddle table in LDS: p[j] = tw[j << SH], for transforms whose table indices are all
// multiples of 1 << SH (a W-point row pass uses multiples of TW_N / W).
// log2(TW_N / n) for a transform length n that divides TW_N
constexpr int tw_shift(int n) {
  int s = 0;
  while ((n << s) < TW_N) ++s;
  return s;
}

template <int SH>
struct TwLds {
  const float2* p;
  __device__ __forceinline__ float2 operator[](int m) const { return p[m >> SH]; }
};

template <int SH>
__device__ __forceinline__ TwLds<SH> opaque_offset(TwLds<SH> t) {
  return TwLds<SH>{opaque_offset(t.p)};
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads (e.g.
// asynchronous global->LDS prefetches of the next correlation) stay in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders the LDS accesses of one wave: they execute in order, so this only stops the
// compiler from moving them across an exchange done by the threads of a single wave
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

template <int N>
struct LdsSize {
  static constexpr int floats = N + N / 16 + 1;  // per component (index N is valid)
};

// out[m] = sum_{j < NZJ} a[j] W_R^{jm}: first pass when only the leading NZJ inputs of
// every butterfly can be non-zero (zero-padded transforms)
template <int R, int NZJ, bool INV, class TWT>
__device__ __forceinline__ void dft_sparse(float2 (&a)[R], TWT tw) {
  if constexpr (NZJ >= R || NZJ > 2) {
    dft<R, INV>(a);
  } else if constexpr (NZJ == 1) {
#pragma unroll
    for (int m = 1; m < R; ++m) a[m] = a[0];
  } else {
    // a[m] = a0 + a1 W_R^m with the R-th roots of unity (R <= 8) as exact constants
    static_assert(R <= 8, "two-input sparse butterflies up to radix 8");
    (void)tw;
    const float2 a0 = a[0], a1 = a[1];
    constexpr float H = 0.70710678118654752f;
    constexpr float WR[8] = {1.f, H, 0.f, -H, -1.f, -H, 0.f, H};   // cos(2 pi m / 8)
    constexpr float WI[8] = {0.f, -H, -1.f, -H, 0.f, H, 1.f, H};   // -sin(2 pi m / 8)
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int e = m * (8 / R);
      const float2 w = make_float2(WR[e], INV ? -WI[e] : WI[e]);
      a[m] = c_add(a0, c_mul(a1, w));
    }
  }
}

// One Stockham pass of a size-N transform (span NS -> NS*R).  v holds x[t + i*T].
// WAVE: all T threads of the transform lie in one wave, whose in-order LDS accesses need
// no workgroup barrier.
template <int N, int V, int NS, bool INV, int NZJ, bool WAVE = false, class TWT>
__device__ __forceinline__ void stockham_pass(float2 (&v)[V], float* __restrict__ lre, float* __restrict__ lim,
                                              int t, TWT tw) {
  constexpr int T = N / V;
  constexpr int R = (N / NS < V) ? N / NS : V;
  constexpr int NB = V / R;
  constexpr bool LAST = (NS * R == N);
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int b = t + q * T;
    const int k = b & (NS - 1);
    float2 a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = v[q + j * NB];
    if constexpr (NS > 1) {
      // opaque copy of the table pointer: keeps the twiddle loads inside the caller's
      // loops instead of hoisting them (and their registers) out of the loop
      const auto twp = opaque_offset(tw);
      float2 w[R];
      twiddles<R, INV>(w, twp, k, TW_N / (NS * R));
#pragma unroll
      for (int j = 1; j < R; ++j) a[j] = c_mul(a[j], w[j]);
      dft<R, INV>(a);
    } else {
      dft_sparse<R, NZJ, INV>(a, tw);
    }
    if constexpr (LAST) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[q + j * NB] = a[j];
    } else {
      const int base = (b / NS) * NS * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int p = lds_pad(base + j * NS);
        lre[p] = a[j].x;
        lim[p] = a[j].y;
      }
    }
  }
  if constexpr (!LAST) {
    if constexpr (WAVE) wave_lds_order(); else lds_barrier();
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const int p = lds_pad(t + i * T);
      v[i] = make_float2(lre[p], lim[p]);
    }
    if constexpr (WAVE) wave_lds_order(); else lds_barrier();
  }
}

// Full transform; NZJ describes the zero padding of the input (first pass only).
template <int N, int V, bool INV, int NZJ = 16, int NS = 1, bool WAVE = false, class TWT>
__device__ __forceinline__ void fft_wg(float2 (&v)[V], float* __restrict__ lre, float* __restrict__ lim, int t,
                                       TWT tw) {
  constexpr int R = (N / NS < V) ? N / NS : V;
  stockham_pass<N, V, NS, INV, (NS == 1 ? NZJ : 16), WAVE>(v, lre, lim, t, tw);
  if constexpr (NS * R < N) fft_wg<N, V, INV, NZJ, NS * R, WAVE>(v, lre, lim, t, tw);
}

}  // namespace tdtm
