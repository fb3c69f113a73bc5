// Once-per-search input preparation on gfx950: the transforms and filters the manager
// builds before a search (match_template_manager.py:160-199) as device kernels.
//
//   k_rfft_rows    r2c along x of real rows (half-length complex FFT + split), optionally
//                  reading a cubic volume through a full fftshift
//   k_fft_cols     in-place forward FFT along a strided axis (columns of a row-major
//                  array, or the slowest axis of a volume), optionally fftshifting the
//                  output along that axis
//   k_radial_power radial power spectrum of an rfft, gathered per frequency bin
//   k_whiten_curve 1/sqrt(mean power) per bin, normalised to a maximum of 1
//   k_whiten_eval  the curve interpolated onto an rfft grid (the whitening filter)
//   k_prep_filter  image spectrum x whitening + partial sums of its energy
//   k_prep_scale   normalisation to unit real-space variance
//   k_idft_cols / k_idft_rows
//                  inverse 2D real DFT of any size (direct sums in double), to move an
//                  image spectrum whose size is not a power of two onto a padded grid
//
// Everything is HBM-bound streaming work over the (H, W/2+1) spectrum or the n^3 volume;
// the per-bin and per-block reductions run in a fixed order, so results do not depend
// on scheduling (reruns are bit-identical).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "tdtm.h"

namespace tdtm {
namespace {

constexpr int PV = 8;            // values per thread in the transforms
constexpr int MAX_DEVICES = 64;  // device ordinals with a cached twiddle table
constexpr int FILTER_BLOCK = 256;
constexpr int FILTER_PER_THREAD = 16;

// Twiddle table per device (uploaded on first use, kept until the process ends).
std::mutex g_tw_mutex;
float2* g_tw_dev[MAX_DEVICES] = {};

int device_twiddles(const float2** out) {
  int dev = 0;
  TDTM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  if (!g_tw_dev[dev]) {
    float2* p = nullptr;
    TDTM_HIP(hipMalloc(&p, TW_N * sizeof(float2)));
    const std::vector<float2> tw = twiddle_table();
    const hipError_t e = hipMemcpy(p, tw.data(), TW_N * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return fail(std::string("twiddle upload: ") + hipGetErrorString(e));
    }
    g_tw_dev[dev] = p;
  }
  *out = g_tw_dev[dev];
  return 0;
}

template <int M>
struct RowsCfg {
  static constexpr int T = M / PV;                  // threads per row
  static constexpr int RPB = T >= 256 ? 1 : 256 / T;  // rows per block
  static constexpr int BLOCK = RPB * T;
};

// r2c along x (unnormalised forward) of `rows` rows of W = 2M real samples:
//   z[n] = x[2n] + i x[2n+1]  ->  M-point FFT Z  ->
//   X[k] = E[k] + W_W^k O[k],  E[k] = (Z[k] + conj Z[M-k]) / 2,
//   O[k] = -i (Z[k] - conj Z[M-k]) / 2,  k = 0 .. M  (Z[M] = Z[0]).
// Output rows have M + 1 complex values.  SHIFT (cubic volume of side n = W, rows =
// n^2): row r = z n + y reads source row ((z + n/2) mod n, (y + n/2) mod n) and sample
// (x + n/2) mod n -- the fftshift of all three axes before the transform
// (utils/image_processing.py:196).
template <int M, bool SHIFT>
__global__ __launch_bounds__(RowsCfg<M>::BLOCK) void k_rfft_rows(const float* __restrict__ x, long rows,
                                                                 const float2* __restrict__ tw,
                                                                 float2* __restrict__ out) {
  using C = RowsCfg<M>;
  constexpr int T = C::T;
  constexpr int W = 2 * M;
  __shared__ float lds[C::RPB][2][LdsSize<M>::floats];
  const int rl = threadIdx.x / T, t = threadIdx.x % T;
  const long r = (long)blockIdx.x * C::RPB + rl;
  const bool live = r < rows;  // every thread takes part in the barriers
  long src = live ? r : 0;
  if (SHIFT) {
    const long z = src / W, y = src % W;
    src = ((z + W / 2) % W) * W + (y + W / 2) % W;
  }
  const float2* xr = reinterpret_cast<const float2*>(x + src * W);
  float2 v[PV];
#pragma unroll
  for (int i = 0; i < PV; ++i) {
    const int n = t + i * T;
    // sample pair (2n, 2n+1) of the shifted row is the pair starting at 2n + W/2 (mod W)
    const int m = SHIFT ? (n + M / 2) % M : n;
    v[i] = live ? xr[m] : make_float2(0.f, 0.f);
  }
  float* lre = lds[rl][0];
  float* lim = lds[rl][1];
  fft_wg<M, PV, false>(v, lre, lim, t, tw);
  // every pass of fft_wg but the last ends with a barrier after its LDS reads, and the
  // last pass does not touch LDS: the arrays are free for the exchange
#pragma unroll
  for (int i = 0; i < PV; ++i) {
    const int p = lds_pad(t + i * T);
    lre[p] = v[i].x;
    lim[p] = v[i].y;
  }
  __syncthreads();
  if (!live) return;
  float2* o = out + r * (M + 1);
  auto split = [&](int k) {
    const int pa = lds_pad(k % M), pb = lds_pad((M - k) % M);
    const float ar = lre[pa], ai = lim[pa];
    const float br = lre[pb], bi = -lim[pb];  // conj Z[M-k]
    const float2 e = make_float2(0.5f * (ar + br), 0.5f * (ai + bi));
    const float2 od = make_float2(0.5f * (ai - bi), -0.5f * (ar - br));
    o[k] = c_add(e, c_mul(tw[k * (TW_N / W)], od));
  };
#pragma unroll
  for (int i = 0; i < PV; ++i) split(t + i * T);
  if (t == 0) split(M);
}

template <int N>
struct ColsCfg {
  static constexpr int T = N / PV;
  static constexpr int CB = T >= 128 ? 2 : 256 / T;  // columns per block
  static constexpr int BLOCK = CB * T;
};

// In-place forward FFT (unnormalised) along a strided axis: column c < ncols of batch
// b = blockIdx.y holds data[b * bstride + c + j * estride], j < N.  CB adjacent columns
// per block, interleaved over the lanes (thread -> column tid % CB), so a wave's
// accesses cover CB contiguous complex values per row; the 4 blocks with ids 8 apart
// (the same XCD: workgroups are dispatched to the 8 XCDs round robin) take 4
// consecutive column groups and complete each other's 64-byte segments in its L2.
// SHIFT writes element k at (k + N/2) mod N (fftshift along the axis).  Every block
// reads its whole columns before the first barrier and writes them after the last, and
// no two blocks share a column, so the transform can run in place.
template <int N, bool SHIFT>
__global__ __launch_bounds__(ColsCfg<N>::BLOCK) void k_fft_cols(float2* data, int ncols, long estride, long bstride,
                                                                const float2* __restrict__ tw) {
  using C = ColsCfg<N>;
  constexpr int T = C::T;
  __shared__ float lds[C::CB][2][LdsSize<N>::floats];
  const int L = blockIdx.x;
  const int q = (L / 32) * 32 + (L % 8) * 4 + (L / 8) % 4;
  const int nq = (ncols + C::CB - 1) / C::CB;
  if (q >= nq) return;  // uniform over the block
  const int cl = threadIdx.x % C::CB, t = threadIdx.x / C::CB;
  const int c = q * C::CB + cl;
  const bool live = c < ncols;
  float2* col = data + (long)blockIdx.y * bstride + (live ? c : 0);
  float2 v[PV];
#pragma unroll
  for (int i = 0; i < PV; ++i) v[i] = live ? col[(long)(t + i * T) * estride] : make_float2(0.f, 0.f);
  fft_wg<N, PV, false>(v, lds[cl][0], lds[cl][1], t, tw);
  if (!live) return;
#pragma unroll
  for (int i = 0; i < PV; ++i) {
    const int k = t + i * T;
    const int kk = SHIFT ? (k + N / 2) % N : k;
    col[(long)kk * estride] = v[i];
  }
}

__global__ void k_zero_dc(float2* X) { X[0] = make_float2(0.f, 0.f); }

__global__ void k_fill_ones(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1.0;
}

// fftfreq / rfftfreq of index i for an n-point grid, as torch computes them: the signed
// index times the rounded reciprocal 1/n (exact for powers of two)
__device__ __forceinline__ double fftfreq(int i, int n) { return (double)(i < (n + 1) / 2 ? i : i - n) * (1.0 / n); }
__device__ __forceinline__ double rfftfreq(int j, int n) { return (double)j * (1.0 / n); }

// Radial power spectrum of an unshifted (H, W/2+1) rfft (config/correlation_filters.py:
// 97-108 -> torch_fourier_filter whitening_filter): for every element with
// f = sqrt(ky^2 + kx^2) <= fm (cycles / pixel), bin = rint(f / fm * (nb - 1)),
// sum[bin] += |X|^2 (|X| rounded to float first, as torch's abs of a complex64 is),
// cnt[bin] += 1.  One block per bin: it visits only the rows and the kx range its
// annulus can reach (widened by a sample on each side; membership is decided by the
// exact bin formula), and reduces in a fixed order.
__global__ __launch_bounds__(256) void k_radial_power(const float2* __restrict__ X, int H, int W, int nb, double fm,
                                                      double* __restrict__ sums, double* __restrict__ cnts) {
  __shared__ double ss[256], sc[256];
  const int b = blockIdx.x;
  const int wr = W / 2 + 1;
  const double scale = (double)(nb - 1) / fm;
  const double rlo = fmax((b - 0.5) / scale, 0.0);
  const double rhi = fmin((b + 0.5) / scale, fm);
  const int ry = min(H / 2, (int)ceil(rhi * H) + 1);
  // signed row frequencies i (ky = i / H): [-H/2, H/2 - 1] for even H (row H/2 is -1/2,
  // visited here as i = H/2), [-(H-1)/2, (H-1)/2] for odd H; each row once
  const int i0 = max(-ry, (H & 1) ? -(H / 2) : -(H / 2) + 1), i1 = ry;
  double s = 0.0, c = 0.0;
  for (int i = i0 + (int)threadIdx.x; i <= i1; i += blockDim.x) {
    const int iy = i < 0 ? i + H : i;
    const double ky = fftfreq(iy, H);
    const double k2 = ky * ky;
    // kx range of the annulus in this row, widened by one sample on each side (a row
    // that only touches the annulus, e.g. ky = -1/2 with f = 1/2 at kx = 0, still
    // checks kx = 0 and 1/W)
    const double jl = sqrt(fmax(rlo * rlo - k2, 0.0)) * W;
    const double jh = sqrt(fmax(rhi * rhi - k2, 0.0)) * W;
    const int j0 = max(0, (int)floor(jl) - 1), j1 = min(W / 2, (int)ceil(jh) + 1);
    const float2* row = X + (long)iy * wr;
    for (int j = j0; j <= j1; ++j) {
      const double kx = rfftfreq(j, W);
      const double f = sqrt(ky * ky + kx * kx);
      if (f <= fm && (int)rint(f / fm * (double)(nb - 1)) == b) {
        const float2 v = row[j];
        const double a = (double)hypotf(v.x, v.y);
        s += a * a;
        c += 1.0;
      }
    }
  }
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[b] = ss[0];
    cnts[b] = sc[0];
  }
}

// curve[b] = (mean power of bin b)^-1/2 (empty or zero bins -> 0), divided by its maximum
__global__ __launch_bounds__(1024) void k_whiten_curve(const double* __restrict__ sums, const double* __restrict__ cnts,
                                                       int nb, double* __restrict__ curve) {
  __shared__ double mx[1024];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double avg = sums[b] / fmax(cnts[b], 1.0);
    const double cv = avg > 0.0 ? 1.0 / sqrt(avg) : 0.0;
    curve[b] = cv;
    m = fmax(m, cv);
  }
  mx[threadIdx.x] = m;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  const double top = mx[0];
  for (int b = threadIdx.x; b < nb; b += blockDim.x) curve[b] = curve[b] / top;
}

// whitening value at frequency (ky, kx): linear interpolation of the curve at
// f / fm * (nb - 1), 1 beyond fm
__device__ __forceinline__ double whiten_at_d(double ky, double kx, const double* __restrict__ curve, int nb,
                                              double fm) {
  const double f = sqrt(ky * ky + kx * kx);
  if (!(f <= fm)) return 1.0;
  const double pos = f / fm * (double)(nb - 1);
  const int i0 = min(max((int)floor(pos), 0), nb - 1);
  const int i1 = min(i0 + 1, nb - 1);
  const double t = pos - (double)i0;
  return curve[i0] * (1.0 - t) + curve[i1] * t;
}

// Bandpass and arbitrary-curve filters of the pre-processing (config/correlation_filters.py:
// 164-334): bandpass (on) = 1 on lo <= |k| <= hi, a raised-cosine edge 0.5 (1 + cos(pi d / fo))
// at distance 0 < d <= fo outside the band, 0 beyond (hard edges for fo = 0); curve (nc knots,
// increasing frequencies) = linear interpolation at |k|, 1 below the first / above the last
// knot.  |k| in cycles / pixel.
constexpr int MAX_CURVE = 64;
struct RadialFilters {
  int bp_on;
  double lo, hi, fo;
  int nc;
  double cf[MAX_CURVE], ca[MAX_CURVE];
};

__device__ __forceinline__ double bandpass_at(const RadialFilters& rf, double f) {
  const double d = fmax(rf.lo - f, 0.0) + fmax(f - rf.hi, 0.0);
  if (rf.fo > 0.0) return d <= rf.fo ? 0.5 * (1.0 + cos(M_PI * fmin(d / rf.fo, 1.0))) : 0.0;
  return d == 0.0 ? 1.0 : 0.0;
}

__device__ __forceinline__ double curve_at(const RadialFilters& rf, double f) {
  if (!(f >= rf.cf[0] && f <= rf.cf[rf.nc - 1])) return 1.0;
  if (rf.nc == 1) return rf.ca[0];
  int j = 1;  // first knot above f (the last segment for f at the top knot)
  while (j < rf.nc - 1 && rf.cf[j] <= f) ++j;
  const double t = (f - rf.cf[j - 1]) / (rf.cf[j] - rf.cf[j - 1]);
  return rf.ca[j - 1] + t * (rf.ca[j] - rf.ca[j - 1]);
}

// product of the enabled radial filters at |k| = f
__device__ __forceinline__ double radial_at(const RadialFilters& rf, double f) {
  double v = 1.0;
  if (rf.bp_on) v *= bandpass_at(rf, f);
  if (rf.nc > 0) v *= curve_at(rf, f);
  return v;
}

__device__ __forceinline__ float whiten_at(double ky, double kx, const double* __restrict__ curve, int nb,
                                           double fm) {
  return (float)whiten_at_d(ky, kx, curve, nb, fm);
}

// whitening x radial filters (rf.bp_on == 0 and rf.nc == 0: the whitening value alone)
__device__ __forceinline__ float filter_at(double ky, double kx, const double* __restrict__ curve, int nb, double fm,
                                           const RadialFilters& rf) {
  double v = whiten_at_d(ky, kx, curve, nb, fm);
  if (rf.bp_on || rf.nc > 0) v *= radial_at(rf, sqrt(ky * ky + kx * kx));
  return (float)v;
}

__global__ void k_whiten_eval(const double* __restrict__ curve, int nb, double fm, int h, int w,
                              RadialFilters rf, float* __restrict__ out) {
  const int wr = w / 2 + 1;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)h * wr) return;
  const int iy = (int)(e / wr), jx = (int)(e % wr);
  out[e] = filter_at(fftfreq(iy, h), rfftfreq(jx, w), curve, nb, fm, rf);
}

// X *= filter(ky, kx) in place (complex64 x float32, as utils/image_processing.py:49; the
// filter is whitening x bandpass x curve) and per-block partial sums of |X|^2 with the inner
// columns 1 .. W/2-1 counted twice (the energy of the full spectrum, :54-58); with the bandpass
// on (part_dim != nullptr) also partial sums of the bandpass alone, inner columns twice (its
// dimensionality, :66).
__global__ __launch_bounds__(FILTER_BLOCK) void k_prep_filter(float2* __restrict__ X, int H, int W,
                                                              const double* __restrict__ curve, int nb, double fm,
                                                              RadialFilters rf, double* __restrict__ part,
                                                              double* __restrict__ part_dim) {
  __shared__ double ss[FILTER_BLOCK], sd[FILTER_BLOCK];
  const int wr = W / 2 + 1;
  const long n = (long)H * wr;
  const long base = (long)blockIdx.x * FILTER_BLOCK * FILTER_PER_THREAD;
  double s = 0.0, sdim = 0.0;
#pragma unroll 4
  for (int i = 0; i < FILTER_PER_THREAD; ++i) {
    const long e = base + (long)i * FILTER_BLOCK + threadIdx.x;
    if (e < n) {
      const int iy = (int)(e / wr), jx = (int)(e % wr);
      const double ky = fftfreq(iy, H), kx = rfftfreq(jx, W);
      const float wv = filter_at(ky, kx, curve, nb, fm, rf);
      float2 v = X[e];
      v = make_float2(v.x * wv, v.y * wv);
      X[e] = v;
      const double a = (double)v.x * v.x + (double)v.y * v.y;
      const bool inner = jx >= 1 && jx <= W / 2 - 1;
      s += inner ? 2.0 * a : a;
      if (part_dim) {
        const double b = (double)(float)bandpass_at(rf, sqrt(ky * ky + kx * kx));
        sdim += inner ? 2.0 * b : b;
      }
    }
  }
  ss[threadIdx.x] = s;
  sd[threadIdx.x] = sdim;
  __syncthreads();
  for (int w = FILTER_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ss[0];
    if (part_dim) part_dim[blockIdx.x] = sd[0];
  }
}

// X *= sqrt(dimensionality / energy) with dimensionality sum(bp) + sum(bp[:, 1:-1])
// (:60-67): H W for an all-pass bandpass, else the sum of the partial sums part_dim; box
// factor 1 (full image, :69-73).  Every block reduces the partial sums itself, in the same
// order.
__global__ __launch_bounds__(FILTER_BLOCK) void k_prep_scale(float2* __restrict__ X, long n,
                                                             const double* __restrict__ part, int npart,
                                                             double dimensionality,
                                                             const double* __restrict__ part_dim) {
  __shared__ double ss[FILTER_BLOCK], sd[FILTER_BLOCK];
  double s = 0.0, sdim = 0.0;
  for (int i = threadIdx.x; i < npart; i += FILTER_BLOCK) {
    s += part[i];
    if (part_dim) sdim += part_dim[i];
  }
  ss[threadIdx.x] = s;
  sd[threadIdx.x] = sdim;
  __syncthreads();
  for (int w = FILTER_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (part_dim) dimensionality = sd[0];
  const float scale = (float)(sqrt(dimensionality) / sqrt(ss[0]));
  const long base = (long)blockIdx.x * FILTER_BLOCK * FILTER_PER_THREAD;
  for (int i = 0; i < FILTER_PER_THREAD; ++i) {
    const long e = base + (long)i * FILTER_BLOCK + threadIdx.x;
    if (e < n) {
      const float2 v = X[e];
      X[e] = make_float2(v.x * scale, v.y * scale);
    }
  }
}

// Any-size inverse DFTs (once per micrograph, to re-grid a spectrum): direct sums in
// double.  Each thread's rotation e^{2 pi i m / N} is advanced by complex products
// (rounding drift ~N ulp of double, far below float precision).
constexpr int DFT_BLOCK = 256;
constexpr int DFT_MAX_H = 4096;  // column length (LDS: DFT_MAX_H double2)
constexpr int DFT_MAX_W = 8192;  // row length (LDS: DFT_MAX_W / 2 + 1 double2)

// Y[y][kx] = sum_ky X[ky][kx] e^{+2 pi i ky y / H} for every column kx < Wr of an (H, Wr)
// complex64 spectrum; one block per column, the column staged in LDS.
__global__ __launch_bounds__(DFT_BLOCK) void k_idft_cols(const float2* __restrict__ X, int H, int Wr,
                                                        double2* __restrict__ Y) {
  __shared__ double2 col[DFT_MAX_H];
  const int kx = blockIdx.x;
  for (int j = threadIdx.x; j < H; j += DFT_BLOCK) {
    const float2 v = X[(long)j * Wr + kx];
    col[j] = make_double2(v.x, v.y);
  }
  __syncthreads();
  for (int y = threadIdx.x; y < H; y += DFT_BLOCK) {
    double sn, cs;
    sincospi(2.0 * (double)y / (double)H, &sn, &cs);
    double wr = 1.0, wi = 0.0, ar = 0.0, ai = 0.0;
    for (int ky = 0; ky < H; ++ky) {
      const double2 c = col[ky];
      ar = fma(c.x, wr, fma(-c.y, wi, ar));
      ai = fma(c.x, wi, fma(c.y, wr, ai));
      const double nr = fma(wr, cs, -wi * sn);
      wi = fma(wr, sn, wi * cs);
      wr = nr;
    }
    Y[(long)y * Wr + kx] = make_double2(ar, ai);
  }
}

// x[y][n] = s (Re Y[0] + (-1)^n Re Y[W/2] + 2 sum_{k=1}^{W/2-1} Re(Y[k] e^{+2 pi i k n / W}))
// for every row y < H (the imaginary parts of DC and Nyquist are ignored, as pocketfft's
// c2r does; s = 1 / (H W), torch's default normalisation), written as float into row y
// of a (rows, ld) array; one block per row, the row staged in LDS.
__global__ __launch_bounds__(DFT_BLOCK) void k_idft_rows(const double2* __restrict__ Y, int H, int W,
                                                        float* __restrict__ out, long ld) {
  __shared__ double2 row[DFT_MAX_W / 2 + 1];
  const int y = blockIdx.x;
  const int M = W / 2, Wr = M + 1;
  for (int j = threadIdx.x; j < Wr; j += DFT_BLOCK) row[j] = Y[(long)y * Wr + j];
  __syncthreads();
  const double scale = 1.0 / ((double)H * (double)W);
  for (int n = threadIdx.x; n < W; n += DFT_BLOCK) {
    double sn, cs;
    sincospi(2.0 * (double)n / (double)W, &sn, &cs);
    double wr = cs, wi = sn, acc = 0.0;  // k = 1
    for (int k = 1; k < M; ++k) {
      const double2 c = row[k];
      acc = fma(c.x, wr, fma(-c.y, wi, acc));
      const double nr = fma(wr, cs, -wi * sn);
      wi = fma(wr, sn, wi * cs);
      wr = nr;
    }
    const double v = row[0].x + ((n & 1) ? -row[M].x : row[M].x) + 2.0 * acc;
    out[(long)y * ld + n] = (float)(v * scale);
  }
}

// Y[y][kx] = sum_x x[y][x] e^{-2 pi i kx x / W}, kx <= W/2, for every row y of a real
// (H, W) image (any even W <= DFT_MAX_W); one block per row, the row staged in LDS.
__global__ __launch_bounds__(DFT_BLOCK) void k_dft_rows_fwd(const float* __restrict__ img, int W,
                                                           double2* __restrict__ Y) {
  __shared__ double row[DFT_MAX_W];
  const int y = blockIdx.x;
  const int Wr = W / 2 + 1;
  for (int j = threadIdx.x; j < W; j += DFT_BLOCK) row[j] = img[(long)y * W + j];
  __syncthreads();
  for (int kx = threadIdx.x; kx < Wr; kx += DFT_BLOCK) {
    double sn, cs;
    sincospi(-2.0 * (double)kx / (double)W, &sn, &cs);  // e^{-2 pi i kx / W}
    double wr = 1.0, wi = 0.0, ar = 0.0, ai = 0.0;
    for (int x = 0; x < W; ++x) {
      const double v = row[x];
      ar = fma(v, wr, ar);
      ai = fma(v, wi, ai);
      const double nr = fma(wr, cs, -wi * sn);
      wi = fma(wr, sn, wi * cs);
      wr = nr;
    }
    Y[(long)y * Wr + kx] = make_double2(ar, ai);
  }
}

// X[ky][kx] = sum_y Y[y][kx] e^{-2 pi i ky y / H} for every column kx < Wr (any H <=
// DFT_MAX_H), written as complex64; one block per column, the column staged in LDS.
__global__ __launch_bounds__(DFT_BLOCK) void k_dft_cols_fwd(const double2* __restrict__ Y, int H, int Wr,
                                                           float2* __restrict__ X) {
  __shared__ double2 col[DFT_MAX_H];
  const int kx = blockIdx.x;
  for (int j = threadIdx.x; j < H; j += DFT_BLOCK) col[j] = Y[(long)j * Wr + kx];
  __syncthreads();
  for (int ky = threadIdx.x; ky < H; ky += DFT_BLOCK) {
    double sn, cs;
    sincospi(-2.0 * (double)ky / (double)H, &sn, &cs);
    double wr = 1.0, wi = 0.0, ar = 0.0, ai = 0.0;
    for (int y = 0; y < H; ++y) {
      const double2 c = col[y];
      ar = fma(c.x, wr, fma(-c.y, wi, ar));
      ai = fma(c.x, wi, fma(c.y, wr, ai));
      const double nr = fma(wr, cs, -wi * sn);
      wi = fma(wr, sn, wi * cs);
      wr = nr;
    }
    X[(long)ky * Wr + kx] = make_float2((float)ar, (float)ai);
  }
}

// ------------------------------------------------------------------ launchers
template <int M, bool SHIFT>
int launch_rows_m(const float* x, long rows, const float2* tw, float2* out, hipStream_t st) {
  using C = RowsCfg<M>;
  const long nblk = (rows + C::RPB - 1) / C::RPB;
  if (nblk > 0x7fffffffL) return fail("too many rows");
  k_rfft_rows<M, SHIFT><<<dim3((unsigned)nblk), C::BLOCK, 0, st>>>(x, rows, tw, out);
  TDTM_LAUNCHED();
  return 0;
}

template <bool SHIFT>
int launch_rows(int w, const float* x, long rows, const float2* tw, float2* out, hipStream_t st) {
  switch (w) {
    case 32: return launch_rows_m<16, SHIFT>(x, rows, tw, out, st);
    case 64: return launch_rows_m<32, SHIFT>(x, rows, tw, out, st);
    case 128: return launch_rows_m<64, SHIFT>(x, rows, tw, out, st);
    case 256: return launch_rows_m<128, SHIFT>(x, rows, tw, out, st);
    case 512: return launch_rows_m<256, SHIFT>(x, rows, tw, out, st);
    case 1024: return launch_rows_m<512, SHIFT>(x, rows, tw, out, st);
    case 2048: return launch_rows_m<1024, SHIFT>(x, rows, tw, out, st);
    case 4096: return launch_rows_m<2048, SHIFT>(x, rows, tw, out, st);
    case 8192: return launch_rows_m<4096, SHIFT>(x, rows, tw, out, st);
    default: return fail("unsupported row length");
  }
}

template <int N, bool SHIFT>
int launch_cols_n(float2* data, int ncols, long estride, long bstride, int nbatch, const float2* tw, hipStream_t st) {
  using C = ColsCfg<N>;
  const long nq = (ncols + C::CB - 1) / C::CB;
  const long nblk = (nq + 31) / 32 * 32;
  if (nblk > 0x7fffffffL || nbatch > 65535) return fail("column grid too large");
  k_fft_cols<N, SHIFT><<<dim3((unsigned)nblk, nbatch), C::BLOCK, 0, st>>>(data, ncols, estride, bstride, tw);
  TDTM_LAUNCHED();
  return 0;
}

template <bool SHIFT>
int launch_cols(int n, float2* data, int ncols, long estride, long bstride, int nbatch, const float2* tw,
                hipStream_t st) {
  switch (n) {
    case 32: return launch_cols_n<32, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 64: return launch_cols_n<64, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 128: return launch_cols_n<128, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 256: return launch_cols_n<256, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 512: return launch_cols_n<512, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 1024: return launch_cols_n<1024, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 2048: return launch_cols_n<2048, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    case 4096: return launch_cols_n<4096, SHIFT>(data, ncols, estride, bstride, nbatch, tw, st);
    default: return fail("unsupported column length");
  }
}

int rfft2_tw(const float* image, int H, int W, float2* out, const float2* tw, hipStream_t st) {
  if (launch_rows<false>(W, image, H, tw, out, st)) return 1;
  return launch_cols<false>(H, out, W / 2 + 1, W / 2 + 1, 0, 1, tw, st);
}

int rfft2_impl(const float* image, int H, int W, float2* out, hipStream_t st) {
  const float2* tw = nullptr;
  if (device_twiddles(&tw)) return 1;
  return rfft2_tw(image, H, W, out, tw, st);
}

struct PrepWorkspace {
  double *sums, *cnts, *curve, *part;
  double* part_dim;  // bandpass partial sums (filtered preparation with the bandpass on), or null
  double2* dft;  // any-size images: the row transforms (H, W/2+1)
  int nb, npart;
};

bool prep_pow2(int H, int W) { return is_pow2(H) && is_pow2(W) && H >= 32 && W >= 32; }

bool prep_size_ok(int H, int W) {
  return H >= 2 && W >= 2 && !(W & 1) && H <= DFT_MAX_H && W <= DFT_MAX_W;
}

int num_bins(int H, int W, int requested) { return requested > 0 ? requested : (H < W ? H : W) / 2 + 1; }

long filter_blocks(int H, int W) {
  const long n = (long)H * (W / 2 + 1);
  return (n + FILTER_BLOCK * FILTER_PER_THREAD - 1) / (FILTER_BLOCK * FILTER_PER_THREAD);
}

// Workspace doubles: sums, counts, curve (nb each), energy partials (npart), with `dim`
// the bandpass partials (npart), then (sizes that are not powers of two) the row transforms.
PrepWorkspace carve(void* ws, int H, int W, int requested_bins, bool dim = false) {
  PrepWorkspace w;
  w.nb = num_bins(H, W, requested_bins);
  w.npart = (int)filter_blocks(H, W);
  double* p = static_cast<double*>(ws);
  w.sums = p;
  w.cnts = p + w.nb;
  w.curve = p + 2 * w.nb;
  w.part = p + 3 * w.nb;
  w.part_dim = dim ? w.part + w.npart : nullptr;
  const long used = 3L * w.nb + (dim ? 2L : 1L) * w.npart;
  w.dft = reinterpret_cast<double2*>(p + ((used + 1) & ~1L));  // 16-byte aligned
  return w;
}

// The preparation chain on an existing twiddle table and workspace (shared by the
// allocating entry point and the scratch-array one).
// rf: bandpass / curve filters (both off in the unfiltered entry points).
int prepare_image_impl(const float* image, int H, int W, int h, int whitening, int num_freq_bins, double max_freq,
                       float2* X, float* whitening_template, void* workspace, const float2* tw, hipStream_t st,
                       const RadialFilters& rf) {
  const PrepWorkspace w = carve(workspace, H, W, num_freq_bins, rf.bp_on != 0);
  const double fm = max_freq;
  if (prep_pow2(H, W)) {
    if (rfft2_tw(image, H, W, X, tw, st)) return 1;
  } else {  // direct double-precision DFTs for sizes the FFT kernels do not take
    k_dft_rows_fwd<<<H, DFT_BLOCK, 0, st>>>(image, W, w.dft);
    TDTM_LAUNCHED();
    k_dft_cols_fwd<<<W / 2 + 1, DFT_BLOCK, 0, st>>>(w.dft, H, W / 2 + 1, X);
    TDTM_LAUNCHED();
  }
  k_zero_dc<<<1, 1, 0, st>>>(X);
  TDTM_LAUNCHED();
  if (whitening) {
    k_radial_power<<<w.nb, 256, 0, st>>>(X, H, W, w.nb, fm, w.sums, w.cnts);
    TDTM_LAUNCHED();
    k_whiten_curve<<<1, 1024, 0, st>>>(w.sums, w.cnts, w.nb, w.curve);
    TDTM_LAUNCHED();
  } else {
    k_fill_ones<<<(w.nb + 255) / 256, 256, 0, st>>>(w.curve, w.nb);  // flat curve: every filter value is 1
    TDTM_LAUNCHED();
  }
  const long nt = (long)h * (h / 2 + 1);
  k_whiten_eval<<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(w.curve, w.nb, fm, h, h, rf, whitening_template);
  TDTM_LAUNCHED();
  k_prep_filter<<<w.npart, FILTER_BLOCK, 0, st>>>(X, H, W, w.curve, w.nb, fm, rf, w.part, w.part_dim);
  TDTM_LAUNCHED();
  k_prep_scale<<<w.npart, FILTER_BLOCK, 0, st>>>(X, (long)H * (W / 2 + 1), w.part, w.npart, (double)H * (double)W,
                                                  w.part_dim);
  TDTM_LAUNCHED();
  return 0;
}

int prepare_image_checks(const float* image, int H, int W, int h, int num_freq_bins, double max_freq,
                         const void* image_dft, const float* whitening_template) {
  if (!image || !image_dft || !whitening_template) return fail("null argument");
  if (!prep_size_ok(H, W))
    return fail("image size must be H <= 4096 by an even W <= 8192, got " + std::to_string(H) + "x" +
                std::to_string(W));
  if (h < 2 || (h & 1)) return fail("template size must be even and >= 2");
  if (num_freq_bins == 1 || num_freq_bins > 65536) return fail("num_freq_bins must be 0 (default) or in [2, 65536]");
  if (!(max_freq > 0.0)) return fail("max_freq must be positive");
  return 0;
}

// ------------------------------------------------- preparation without allocation
// tdtm_prepare_image_noalloc keeps its twiddle table and its workspace (per-bin power
// sums and counts, the whitening curve, per-block energy partials) in `scratch`, one
// static device array per GPU that lives until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current device (the
// array's name on the host is not a device address).  Every call rewrites all of it
// before reading it: the twiddle table first, then each bin / block slot is written
// by one block before the next kernel reads it, so nothing has to start at zero.  Calls
// that overlap on one GPU share the array and race.
constexpr int NOALLOC_MAX_BINS = 8192;
constexpr long NOALLOC_TW_BYTES = (long)TW_N * sizeof(float2);  // 64 KiB, 256-byte aligned
constexpr long NOALLOC_WS_DOUBLES = 3L * NOALLOC_MAX_BINS + (4096L * 4097 + FILTER_BLOCK * FILTER_PER_THREAD - 1) /
                                                                (FILTER_BLOCK * FILTER_PER_THREAD);
constexpr long NOALLOC_SCRATCH_BYTES = NOALLOC_TW_BYTES + ((NOALLOC_WS_DOUBLES * 8 + 255) / 256) * 256;
static_assert(NOALLOC_SCRATCH_BYTES <= TDTM_PREPARE_NOALLOC_SCRATCH_BYTES, "the scratch array stays within the size the header states");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
void* g_scratch_dev[MAX_DEVICES] = {};

// Device address of this GPU's scratch array (looked up once per device).
int scratch_address(void** out) {
  int dev = 0;
  TDTM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

// The twiddle table of twiddle_table() (common.hpp), built in double on the device by the
// same first-octant mapping (exact quarter / eighth turns).
__global__ void k_twiddle_fill(float2* __restrict__ tw) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= TW_N) return;
  constexpr int Q = TW_N / 4;
  const int q = m / Q, r = m % Q;
  double c, sn;
  if (2 * r <= Q) {
    const double a = 2.0 * M_PI * (double)r / (double)TW_N;
    c = cos(a);
    sn = sin(a);
  } else {
    const double a = 2.0 * M_PI * (double)(Q - r) / (double)TW_N;
    c = sin(a);
    sn = cos(a);
  }
  double cq = c, sq = sn;
  if (q == 1) cq = -sn, sq = c;
  if (q == 2) cq = -c, sq = -sn;
  if (q == 3) cq = sn, sq = -c;
  tw[m] = make_float2((float)cq, (float)-sq);
}

}  // namespace
}  // namespace tdtm

using namespace tdtm;

extern "C" {

int tdtm_rfft2(const float* image, int H, int W, void* out, void* stream) {
  if (!image || !out) return fail("null argument");
  if (!is_pow2(H) || !is_pow2(W) || H < 32 || W < 32 || H > 4096 || W > 8192)
    return fail("image size must be a power of two in [32, 4096] x [32, 8192], got " + std::to_string(H) + "x" +
                std::to_string(W));
  return rfft2_impl(image, H, W, static_cast<float2*>(out), static_cast<hipStream_t>(stream));
}

int tdtm_volume_rfft(const float* volume, int n, void* out, void* stream) {
  if (!volume || !out) return fail("null argument");
  if (!is_pow2(n) || n < 32 || n > 512) return fail("volume side must be a power of two in [32, 512]");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const float2* tw = nullptr;
  if (device_twiddles(&tw)) return 1;
  float2* o = static_cast<float2*>(out);
  const int wr = n / 2 + 1;
  // x: r2c of the fftshifted rows; y: per z-plane, columns kx, output fftshifted;
  // z: all (y, kx) columns of the volume, output fftshifted
  if (launch_rows<true>(n, volume, (long)n * n, tw, o, st)) return 1;
  if (launch_cols<true>(n, o, wr, wr, (long)n * wr, n, tw, st)) return 1;
  return launch_cols<true>(n, o, n * wr, (long)n * wr, 0, 1, tw, st);
}

int64_t tdtm_irfft2_any_workspace_bytes(int H, int W) {
  if (H < 1 || H > DFT_MAX_H || W < 2 || W > DFT_MAX_W || (W & 1)) return -1;
  return (int64_t)H * (W / 2 + 1) * (int64_t)sizeof(double2);
}

int tdtm_irfft2_any(const void* spectrum, int H, int W, float* image, int64_t ld, void* workspace, void* stream) {
  if (!spectrum || !image || !workspace) return fail("null argument");
  if (H < 1 || H > DFT_MAX_H || W < 2 || W > DFT_MAX_W || (W & 1))
    return fail("irfft2_any: need 1 <= H <= 4096 and an even W <= 8192, got " + std::to_string(H) + "x" +
                std::to_string(W));
  if (ld < W) return fail("irfft2_any: row stride below the image width");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  double2* Y = static_cast<double2*>(workspace);
  k_idft_cols<<<W / 2 + 1, DFT_BLOCK, 0, st>>>(static_cast<const float2*>(spectrum), H, W / 2 + 1, Y);
  TDTM_LAUNCHED();
  k_idft_rows<<<H, DFT_BLOCK, 0, st>>>(Y, H, W, image, (long)ld);
  TDTM_LAUNCHED();
  return 0;
}

int64_t tdtm_prepare_image_workspace_bytes(int H, int W, int num_freq_bins) {
  if (!prep_size_ok(H, W)) return -1;
  if (num_freq_bins == 1 || num_freq_bins > 65536) return -1;
  int64_t doubles = 3L * num_bins(H, W, num_freq_bins) + filter_blocks(H, W);
  if (!prep_pow2(H, W)) doubles = ((doubles + 1) & ~1L) + 2L * H * (W / 2 + 1);
  return doubles * (int64_t)sizeof(double);
}

int64_t tdtm_prepare_image_filtered_workspace_bytes(int H, int W, int num_freq_bins) {
  if (!prep_size_ok(H, W)) return -1;
  if (num_freq_bins == 1 || num_freq_bins > 65536) return -1;
  int64_t doubles = 3L * num_bins(H, W, num_freq_bins) + 2 * filter_blocks(H, W);
  if (!prep_pow2(H, W)) doubles = ((doubles + 1) & ~1L) + 2L * H * (W / 2 + 1);
  return doubles * (int64_t)sizeof(double);
}

int tdtm_prepare_image(const float* image, int H, int W, int h, int whitening, int num_freq_bins, double max_freq,
                       void* image_dft, float* whitening_template, void* workspace, void* stream) {
  if (prepare_image_checks(image, H, W, h, num_freq_bins, max_freq, image_dft, whitening_template)) return 1;
  if (!workspace) return fail("null argument");
  const float2* tw = nullptr;
  if (prep_pow2(H, W) && device_twiddles(&tw)) return 1;
  return prepare_image_impl(image, H, W, h, whitening, num_freq_bins, max_freq, static_cast<float2*>(image_dft),
                            whitening_template, workspace, tw, static_cast<hipStream_t>(stream), RadialFilters{});
}

int tdtm_prepare_image_filtered(const float* image, int H, int W, int h, int whitening, int num_freq_bins,
                                double max_freq, int bandpass, double bp_low, double bp_high, double bp_falloff,
                                int n_curve, const double* curve_freq, const double* curve_amp, void* image_dft,
                                float* template_filter, void* workspace, void* stream) {
  if (prepare_image_checks(image, H, W, h, num_freq_bins, max_freq, image_dft, template_filter)) return 1;
  if (!workspace) return fail("null argument");
  RadialFilters rf{};
  if (bandpass) {
    if (!(bp_low >= 0.0) || !(bp_high >= 0.0) || !(bp_falloff >= 0.0))
      return fail("bandpass cut-offs and falloff must be >= 0");
    rf.bp_on = 1;
    rf.lo = bp_low;
    rf.hi = bp_high;
    rf.fo = bp_falloff;
  }
  if (n_curve < 0 || n_curve > MAX_CURVE)
    return fail("the arbitrary curve takes at most " + std::to_string(MAX_CURVE) + " points");
  if (n_curve > 0) {
    if (!curve_freq || !curve_amp) return fail("null argument");
    for (int i = 0; i < n_curve; ++i) {
      if (!std::isfinite(curve_freq[i]) || !std::isfinite(curve_amp[i])) return fail("curve values must be finite");
      if (i > 0 && !(curve_freq[i] > curve_freq[i - 1])) return fail("curve frequencies must increase");
      rf.cf[i] = curve_freq[i];
      rf.ca[i] = curve_amp[i];
    }
    rf.nc = n_curve;
  }
  const float2* tw = nullptr;
  if (prep_pow2(H, W) && device_twiddles(&tw)) return 1;
  return prepare_image_impl(image, H, W, h, whitening, num_freq_bins, max_freq, static_cast<float2*>(image_dft),
                            template_filter, workspace, tw, static_cast<hipStream_t>(stream), rf);
}

int tdtm_prepare_image_noalloc(const float* image, int H, int W, int h, int whitening, int num_freq_bins,
                               double max_freq, void* image_dft, float* whitening_template, void* stream) {
  if (prepare_image_checks(image, H, W, h, num_freq_bins, max_freq, image_dft, whitening_template)) return 1;
  if (!prep_pow2(H, W) || H > 4096 || W > 8192)
    return fail("tdtm_prepare_image_noalloc: image edges must be powers of two in [32, 4096] x [32, 8192], got " +
                std::to_string(H) + "x" + std::to_string(W));
  if (num_bins(H, W, num_freq_bins) > NOALLOC_MAX_BINS)
    return fail("tdtm_prepare_image_noalloc: at most " + std::to_string(NOALLOC_MAX_BINS) + " frequency bins");
  if (tdtm_prepare_image_workspace_bytes(H, W, num_freq_bins) > NOALLOC_SCRATCH_BYTES - NOALLOC_TW_BYTES)
    return fail("tdtm_prepare_image_noalloc: workspace exceeds the scratch array");
  void* base = nullptr;
  if (scratch_address(&base)) return 1;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  float2* tw = static_cast<float2*>(base);
  k_twiddle_fill<<<TW_N / 256, 256, 0, st>>>(tw);
  TDTM_LAUNCHED();
  void* ws = static_cast<unsigned char*>(base) + NOALLOC_TW_BYTES;
  return prepare_image_impl(image, H, W, h, whitening, num_freq_bins, max_freq, static_cast<float2*>(image_dft),
                            whitening_template, ws, tw, st, RadialFilters{});
}

}  // extern "C"
