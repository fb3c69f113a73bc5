// Whole-orientation 2DTM search on gfx950: kernels and C ABI.
//
// One search step (B orientations x D defocus x C pixel sizes = K correlations):
//   k_slice        trilinear Fourier central slice, Hermitian fold, DC zero, contrast
//                  flip and the projection's quadrant swap folded in as (-1)^(ky+kx)
//   k_proj_cols    filter multiply + inverse FFT along y of every slice column
//   k_proj_rows    c2r along x -> real projection rows + normalisation partial sums
//   k_proj_norm    per-projection edge mean / scale (affine normalisation)
//   k_rowfwd       normalised projection rows -> zero-padded r2c along x (W points)
//   k_colpass      per column group: pruned forward FFT along y, multiply by the image
//                  spectrum conj, inverse FFT along y -> tiled half-spectrum intermediate
//   k_rowpass      per row group: c2r along x fused with MIP / argmax / sum / sum-sq
//                  update; the per-correlation correlation maps never reach HBM
// The intermediate between k_colpass and k_rowpass (one complex half spectrum per
// correlation) is the only full-size array written and read per correlation.  It is
// stored in 4-row x 2-column tiles (64 bytes) so that the column pass writes, and
// the row pass reads, whole 64-byte segments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "fft4.hpp"
#include "fft_mfma.hpp"
#include "tdtm.h"

namespace tdtm {

constexpr int V = 8;   // values per thread in every FFT
constexpr int CW = 2;  // columns per column-pass block (tile width)
constexpr int TR = 4;  // rows per tile (= rows per row-pass block)
// Position of (row r, column c) inside a tile of the intermediate, in complex elements: the
// tile is column-major, each column's TR rows contiguous (32 bytes), so that a column-pass
// block of either width writes whole 16-byte pieces (two rows of one column).
__host__ __device__ constexpr int tile_off(int r, int c) { return c * TR + r; }
// float4 index of the piece (column c, rows 2 h2, 2 h2 + 1) inside a tile
__host__ __device__ constexpr int tile_f4(int c, int h2) { return (c * TR + 2 * h2) / 2; }
static_assert(MF_CW == CW && MF_TR == TR, "the matrix-core column pass writes the same tiles");
constexpr int TDTM_LDS_BYTES = 160 * 1024;  // LDS one workgroup may allocate on gfx950
constexpr int REFINE_CHUNKS_MAX = 64;        // correlation chunks of a refinement row pass


// ------------------------------------------------------------------------- kernels
__global__ void k_tile_image(const float2* __restrict__ X, int H, int Wr, int G, float scale,
                             float2* __restrict__ Xt) {
  const long total = (long)G * H * CW;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int c = e % CW;
  const int ky = (e / CW) % H;
  const int g = e / ((long)CW * H);
  const int kx = g * CW + c;
  float2 v = make_float2(0.f, 0.f);
  if (kx < Wr) v = c_scale(X[(long)ky * Wr + kx], scale);
  Xt[e] = v;
}

__global__ void k_filters(const float* __restrict__ ctf, const float* __restrict__ wf, long plane, long total,
                          float* __restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  out[e] = ctf[e] * wf[e % plane];
}

// Trilinear central slice of the fftshifted 3D rfft; output unshifted in y.
__device__ __forceinline__ void rotmat_zyz(const float* e, float (&r)[9]) {
  const float d2r = 0.017453292519943295f;
  float sph, cph, sth, cth, sps, cps;
  sincosf(e[0] * d2r, &sph, &cph);
  sincosf(e[1] * d2r, &sth, &cth);
  sincosf(e[2] * d2r, &sps, &cps);
  r[0] = cph * cth * cps - sph * sps;
  r[1] = -cph * cth * sps - sph * cps;
  r[2] = cph * sth;
  r[3] = sph * cth * cps + cph * sps;
  r[4] = -sph * cth * sps + cph * cps;
  r[5] = sph * sth;
  r[6] = -sth * cps;
  r[7] = sth * sps;
  r[8] = cth;
}

// Value of the central slice of the fftshifted 3D rfft `vol` (n, n, n/2+1) at slice-plane
// frequency (fx, fy) (cycles / sample) for the rotation with columns 0, 1 = (r00, r10, r20),
// (r01, r11, r21): trilinear sample at R (fx, fy, 0), conjugated where the rotated
// frequency has x < 0 (Hermitian fold), zero outside the volume.
__device__ __forceinline__ float2 slice_value(const float2* __restrict__ vol, int n, const float (&ro)[9], float fx,
                                              float fy) {
  const int wr = n / 2 + 1;
  float rx = ro[0] * fx + ro[1] * fy, ry = ro[3] * fx + ro[4] * fy, rz = ro[6] * fx + ro[7] * fy;
  const bool flip = rx < 0.f;
  if (flip) {
    rx = -rx;
    ry = -ry;
    rz = -rz;
  }
  const float cz = rz * n + (float)(n / 2), cy = ry * n + (float)(n / 2), cx = rx * n;
  float2 val = make_float2(0.f, 0.f);
  if (cz >= 0.f && cz <= (float)(n - 1) && cy >= 0.f && cy <= (float)(n - 1) && cx >= 0.f &&
      cx <= (float)(wr - 1)) {
    const int z0 = (int)floorf(cz), y0 = (int)floorf(cy), x0 = (int)floorf(cx);
    const float fz = cz - z0, fyy = cy - y0, fxx = cx - x0;
    const int z1 = min(z0 + 1, n - 1), y1 = min(y0 + 1, n - 1), x1 = min(x0 + 1, wr - 1);
    const long sz = (long)n * wr;
    const float2 v000 = vol[z0 * sz + (long)y0 * wr + x0], v001 = vol[z0 * sz + (long)y0 * wr + x1];
    const float2 v010 = vol[z0 * sz + (long)y1 * wr + x0], v011 = vol[z0 * sz + (long)y1 * wr + x1];
    const float2 v100 = vol[z1 * sz + (long)y0 * wr + x0], v101 = vol[z1 * sz + (long)y0 * wr + x1];
    const float2 v110 = vol[z1 * sz + (long)y1 * wr + x0], v111 = vol[z1 * sz + (long)y1 * wr + x1];
    const float w000 = (1 - fz) * (1 - fyy) * (1 - fxx), w001 = (1 - fz) * (1 - fyy) * fxx;
    const float w010 = (1 - fz) * fyy * (1 - fxx), w011 = (1 - fz) * fyy * fxx;
    const float w100 = fz * (1 - fyy) * (1 - fxx), w101 = fz * (1 - fyy) * fxx;
    const float w110 = fz * fyy * (1 - fxx), w111 = fz * fyy * fxx;
    val.x = w000 * v000.x + w001 * v001.x + w010 * v010.x + w011 * v011.x + w100 * v100.x + w101 * v101.x +
            w110 * v110.x + w111 * v111.x;
    val.y = w000 * v000.y + w001 * v001.y + w010 * v010.y + w011 * v011.y + w100 * v100.y + w101 * v101.y +
            w110 * v110.y + w111 * v111.y;
  }
  if (flip) val.y = -val.y;
  return val;
}

// Anisotropic magnification of the slices (a 2 x 2 matrix A on (x, y) frequency column
// vectors; `on` = 0: none).  Follows torch_fourier_slice.transform_slice_2d as called at
// backend/cross_correlation.py:81-88: the slice value at k is the slice's bilinear sample at
// A k on its own (fftshifted-in-y) grid, folded (negated and conjugated) where x' < 0, zero
// outside the grid.
struct SliceMag {
  float a00, a01, a10, a11;
  int on;
};

// R = Rz(phi) Ry(theta) Rz(psi) of orientation o0 + b, right-multiplied by the rotation
// of `base` (refinement: offset composed with the particle's pose) when base != nullptr.
// o0p != nullptr: the first orientation is read from device memory (graph replays of a
// captured step advance it on the device).
__global__ void k_slice(const float2* __restrict__ vol, int n, const float* __restrict__ euler, int o0, int nb,
                        const float* __restrict__ base, float2* __restrict__ S, const int* __restrict__ o0p,
                        SliceMag mag) {
  if (o0p != nullptr) o0 = *o0p;
  const int wr = n / 2 + 1;
  const long total = (long)nb * n * wr;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int kx = idx % wr;
  const int r = (idx / wr) % n;
  const int b = idx / ((long)wr * n);
  float ro[9];
  rotmat_zyz(euler + 3 * (o0 + b), ro);
  if (base != nullptr) {
    float rb[9], rc[9];
    rotmat_zyz(base, rb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) rc[3 * i + j] = ro[3 * i] * rb[j] + ro[3 * i + 1] * rb[3 + j] + ro[3 * i + 2] * rb[6 + j];
#pragma unroll
    for (int i = 0; i < 9; ++i) ro[i] = rc[i];
  }
  const float fx = (float)kx / (float)n;
  const float fy = (float)(r < n / 2 ? r : r - n) / (float)n;
  float2 val;
  if (!mag.on) {
    val = slice_value(vol, n, ro, fx, fy);
  } else {
    // bilinear sample of the slice grid (rows iy <-> fy = (iy - n/2)/n, columns ix <->
    // fx = ix/n) at A k
    float tx = mag.a00 * fx + mag.a01 * fy, ty = mag.a10 * fx + mag.a11 * fy;
    const bool fold = tx < 0.f;
    if (fold) {
      tx = -tx;
      ty = -ty;
    }
    const float cy = ty * n + (float)(n / 2), cx = tx * n;
    val = make_float2(0.f, 0.f);
    if (cy >= 0.f && cy <= (float)(n - 1) && cx >= 0.f && cx <= (float)(wr - 1)) {
      const int y0 = (int)floorf(cy), x0 = (int)floorf(cx);
      const float wy = cy - y0, wx = cx - x0;
      const int y1 = min(y0 + 1, n - 1), x1 = min(x0 + 1, wr - 1);
      const float inv = 1.f / (float)n;
      const float fy0 = (float)(y0 - n / 2) * inv, fy1 = (float)(y1 - n / 2) * inv;
      const float fx0 = (float)x0 * inv, fx1 = (float)x1 * inv;
      const float2 s00 = slice_value(vol, n, ro, fx0, fy0), s01 = slice_value(vol, n, ro, fx1, fy0);
      const float2 s10 = slice_value(vol, n, ro, fx0, fy1), s11 = slice_value(vol, n, ro, fx1, fy1);
      const float w00 = (1 - wy) * (1 - wx), w01 = (1 - wy) * wx, w10 = wy * (1 - wx), w11 = wy * wx;
      val.x = w00 * s00.x + w01 * s01.x + w10 * s10.x + w11 * s11.x;
      val.y = w00 * s00.y + w01 * s01.y + w10 * s10.y + w11 * s11.y;
    }
    if (fold) val.y = -val.y;
  }
  if (r == 0 && kx == 0) val = make_float2(0.f, 0.f);
  // contrast flip (-1) times (-1)^(ky+kx): the spatial quadrant swap of the projection
  const float sgn = ((r + kx) & 1) ? 1.f : -1.f;
  S[idx] = c_scale(val, sgn);
}

// Inverse FFT along y of (slice * filter) for CP columns per block.
// Block (column group, slice b, filter chunk): the slice columns are loaded once and
// go through the filters f0 .. f1-1 of the chunk (correlation k = b + nb f, filter f =
// ci nd + d), the next filter's values loaded while the current transform runs.
template <int HS, int CP>
__global__ __launch_bounds__(CP* HS / V) void k_proj_cols(const float2* __restrict__ S,
                                                          const float* __restrict__ filt, int nb, int nf,
                                                          int fchunk, const float2* __restrict__ tw,
                                                          float2* __restrict__ Tb) {
  constexpr int T = HS / V;
  constexpr int WR = HS / 2 + 1;
  constexpr int LF = LdsSize<HS>::floats;
  constexpr int SH = tw_shift(HS);
  __shared__ float lds[CP * 2 * LF];
  // the transform's twiddles (table entries that are multiples of TW_N / HS) in LDS: the
  // loop below then waits on global memory only for the next filter's values, not for
  // twiddle loads queued behind the previous filter's stores
  __shared__ float2 ltw[HS / 2];
  for (int j = threadIdx.x; j < HS / 2; j += CP * T) ltw[j] = tw[j << SH];
  __syncthreads();
  const TwLds<SH> twl{ltw};
  const int c = threadIdx.x % CP, t = threadIdx.x / CP;
  const int kx = blockIdx.x * CP + c;
  const int b = blockIdx.y;
  const int f0 = blockIdx.z * fchunk;
  const int f1 = min(nf, f0 + fchunk);
  const bool ok = kx < WR;
  const float2* s = S + (long)b * HS * WR;
  float2 sv[V];
  float fv[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const long e = (long)(t + i * T) * WR + kx;
    sv[i] = ok ? s[e] : make_float2(0.f, 0.f);
    fv[i] = ok ? filt[(long)f0 * HS * WR + e] : 0.f;
  }
  for (int f = f0; f < f1; ++f) {
    float2 v[V];
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = c_scale(sv[i], fv[i]);
    if (f + 1 < f1) {
#pragma unroll
      for (int i = 0; i < V; ++i) fv[i] = ok ? filt[(long)(f + 1) * HS * WR + (long)(t + i * T) * WR + kx] : 0.f;
    }
    fft_wg<HS, V, true>(v, lds + c * 2 * LF, lds + c * 2 * LF + LF, t, twl);
    if (ok) {
      float2* o = Tb + (long)(b + (long)nb * f) * HS * WR;
#pragma unroll
      for (int i = 0; i < V; ++i) o[(long)(t + i * T) * WR + kx] = v[i];
    }
  }
}

// c2r along x of each row (imaginary parts of DC / Nyquist ignored, as pocketfft's
// c2r does) + per-row partial sums for the projection normalisation.
template <int HS, int RPB>
__global__ __launch_bounds__(RPB*(HS / 2) / V) void k_proj_rows(const float2* __restrict__ Tb,
                                                                const float2* __restrict__ tw,
                                                                float* __restrict__ P, float* __restrict__ part) {
  constexpr int M = HS / 2;
  constexpr int T = M / V;
  constexpr int WR = M + 1;
  constexpr int LF = LdsSize<M>::floats;
  // one region per row: first the staged half-spectrum row, then the transform's split
  // re / im arrays (the row's own wave reads the staged row before it writes them)
  __shared__ __attribute__((aligned(16))) float lds[RPB * 2 * LF];
  static_assert(2 * WR <= 2 * LF, "a staged row fits its transform region");
  const int rl = threadIdx.x / T, t = threadIdx.x % T;
  const int y = blockIdx.x * RPB + rl;
  const int k = blockIdx.y;
  const float2* row = Tb + ((long)k * HS + y) * WR;
  static_assert(T <= 64 && 64 % T == 0, "a row's threads lie in one wave");
  // every exchange below stays inside the row's threads, i.e. inside one wave
  float2* stg = reinterpret_cast<float2*>(lds + rl * 2 * LF);
  {
    // all of the thread's loads first (one memory latency instead of one per element)
    float2 ld[V];
#pragma unroll
    for (int i = 0; i < V; ++i) ld[i] = row[t + i * T];
    const float2 nyq = t == 0 ? row[M] : make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < V; ++i) stg[t + i * T] = ld[i];
    if (t == 0) stg[M] = nyq;
  }
  wave_lds_order();
  float2 v[V];
  const float2* g = stg;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int m = t + i * T;
    float2 A = g[m];
    float2 Bc = g[M - m];
    if (m == 0) {
      A.y = 0.f;
      Bc.y = 0.f;
    }
    Bc = c_conj(Bc);
    const float2 w = c_conj(tw[m * (TW_N / HS)]);  // e^{+2 pi i m / HS}
    const float2 dsum = c_add(A, Bc), ddif = c_sub(A, Bc);
    const float2 od = c_mul(w, ddif);
    v[i] = make_float2(dsum.x - od.y, dsum.y + od.x);  // dsum + i*od
  }
  wave_lds_order();
  fft_wg<M, V, true, 16, 1, true>(v, lds + rl * 2 * LF, lds + rl * 2 * LF + LF, t, tw);
  const float scale = 1.0f / ((float)HS * (float)HS);
  float s1 = 0.f, s2 = 0.f;
  float* prow = P + ((long)k * HS + y) * HS;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int nn = t + i * T;
    const float2 p = c_scale(v[i], scale);
    reinterpret_cast<float2*>(prow)[nn] = p;
    s1 += p.x + p.y;
    s2 += p.x * p.x + p.y * p.y;
  }
  float se = 0.f;
  if (t == 0) se += v[0].x * scale;            // p[0]
  if (t == T - 1) se += v[V - 1].y * scale;    // p[HS-1]
#pragma unroll
  for (int off = T / 2; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, T);
    s2 += __shfl_xor(s2, off, T);
    se += __shfl_xor(se, off, T);
  }
  if (t == 0) {
    float* pp = part + ((long)k * HS + y) * 3;
    pp[0] = s1;
    pp[1] = s2;
    pp[2] = (y == 0 || y == HS - 1) ? s1 : se;
  }
}

// Edge mean e and scale a so that a*(p - e) is the reference's normalised projection
// (backend/utils.py:154-187, incl. the doubled relative-size factor on the mean).
__global__ void k_proj_norm(const float* __restrict__ part, int hs, int H, int W, float2* __restrict__ ae) {
  __shared__ double red[3][256];
  const int k = blockIdx.x;
  double s1 = 0, s2 = 0, se = 0;
  for (int y = threadIdx.x; y < hs; y += blockDim.x) {
    const float* pp = part + ((long)k * hs + y) * 3;
    s1 += pp[0];
    s2 += pp[1];
    se += pp[2];
  }
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = se;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      red[0][threadIdx.x] += red[0][threadIdx.x + off];
      red[1][threadIdx.x] += red[1][threadIdx.x + off];
      red[2][threadIdx.x] += red[2][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)hs * hs;
    const double e = red[2][0] / (4.0 * hs - 4.0);
    const double rel = n / ((double)H * W);
    const double mu = (red[0][0] / n - e) * rel * rel;
    const double cc = e + mu;
    double var = red[1][0] - 2.0 * cc * red[0][0] + n * cc * cc;
    var += ((double)H - hs) * ((double)W - hs) * mu * mu;
    const double a = sqrt((double)H * W) / sqrt(var > 1e-8 ? var : 1e-8);
    ae[k] = make_float2((float)a, (float)e);
  }
}

// Rows per row-forward block: 4 for the long rows (1024 threads at W = 4096), up to 16
// for short rows so that small images (refinement crops) do not launch one small block
// per 4 rows of every correlation.
template <int WS>
struct RowFwdCfg {
  static constexpr int T = (WS / 2) / V;
  static constexpr int RR = T >= 512 ? 2 : (T >= 256 ? 4 : (T >= 128 ? 8 : 16));
  static constexpr int BLOCK = RR * T;
};

// Zero-padded r2c along x (WS points, HS non-zero) of the normalised projection rows.
// A block transforms RR consecutive rows at once (one row per T threads), so every
// tile column gets RR x 16 contiguous bytes from one block.  With RR < 8 the 8 / RR
// blocks of each 8-row group (a 128-byte line of Rb per tile column) get ids 8 apart, so
// they run on the same XCD (workgroups go to the 8 XCDs round robin) and complete the
// lines in its L2; with RR >= 8 a block writes whole lines.
template <int WS, int NZJ>
__global__ __launch_bounds__(RowFwdCfg<WS>::BLOCK) void k_rowfwd(const float* __restrict__ P,
                                                                const float2* __restrict__ ae, int hs, int G,
                                                                const float2* __restrict__ tw,
                                                                float2* __restrict__ Rb) {
  constexpr int RR = RowFwdCfg<WS>::RR;
  constexpr int M = WS / 2;
  constexpr int T = M / V;
  constexpr bool FOUR = UseFour<M>::value;
  constexpr int MF = FOUR ? M : 512;
  constexpr int LF = FOUR ? FS<MF>::floats : LdsSize<M>::floats;
  __shared__ float lds[RR * 2 * LF];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  float* lre = lds + r * 2 * LF;
  float* lim = lre + LF;
  int y0 = blockIdx.x * RR;
  if (RR < 8 && hs % 64 == 0) {
    // the 8 / RR blocks of an 8-row line get ids 8 apart (one XCD)
    constexpr int NB = 8 / RR;
    const int b = blockIdx.x, q = b >> 3;
    y0 = 8 * ((q / NB) * 8 + (b & 7)) + RR * (q % NB);
  }
  const int y = y0 + r;
  const int k = blockIdx.y;
  // twiddles e^{-2 pi i kx / WS} of the split below, loaded (after the transform) before
  // any store is issued: a wait for a load also waits for the stores issued before it
  constexpr int GW = WS / 4 + 1;                    // tile columns of a WS-point row (= G)
  constexpr int GSTEP = RowFwdCfg<WS>::BLOCK / RR;  // tile columns per round of the block
  constexpr int NGR = (GW + GSTEP - 1) / GSTEP;     // rounds
  float2 wsp[NGR][CW];
  const float2 s = ae[k];
  const float2* prow = reinterpret_cast<const float2*>(P + ((long)k * hs + y) * hs);
  float2 v[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int nn = t + i * T;
    float2 z = make_float2(0.f, 0.f);
    if (i < NZJ && nn < hs / 2) {
      const float2 p = prow[nn];
      z = make_float2(s.x * (p.x - s.y), s.x * (p.y - s.y));
    }
    v[i] = z;
  }
  if constexpr (FOUR) {
    fs_nat2perm<MF, false, NZJ, true>(v, lre, lim, t, tw);
    __syncthreads();  // other waves may still use their rows for the wave-local exchanges
  } else {
    fft_wg<M, V, false, NZJ>(v, lre, lim, t, tw);
  }
#pragma unroll
  for (int j = 0; j < NGR; ++j)
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const int kx = CW * (threadIdx.x / RR + j * GSTEP) + c;
      wsp[j][c] = kx <= M ? tw[kx * (TW_N / WS)] : make_float2(1.f, 0.f);
    }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    int m = t + i * T;
    if constexpr (FOUR) m = fs_index<MF>(t, i);
    const int p = lds_pad(m);
    lre[p] = v[i].x;
    lim[p] = v[i].y;
  }
  __syncthreads();
  // split into the W-point spectrum and store: thread -> (row r2 = tid % RR, tile column
  // g), so that a wave's 16-byte stores (both columns of a tile column for one row) cover
  // RR consecutive rows, i.e. RR x 16 contiguous bytes per tile column
  float2* out = Rb + (long)k * G * hs * CW;
  const int r2 = threadIdx.x % RR;
  const float* sre = lds + r2 * 2 * LF;
  const float* sim = sre + LF;
  auto split = [&](int kx, float2 wk) {
    if (kx > M) return make_float2(0.f, 0.f);
    const int a = kx & (M - 1), bi = (M - kx) & (M - 1);
    const float2 A = make_float2(sre[lds_pad(a)], sim[lds_pad(a)]);
    const float2 Bc = make_float2(sre[lds_pad(bi)], -sim[lds_pad(bi)]);
    const float2 E = c_scale(c_add(A, Bc), 0.5f);
    const float2 O = c_scale(c_sub(A, Bc), 0.5f);
    const float2 wo = c_mul(wk, O);                // e^{-2 pi i kx/WS} * O
    return make_float2(E.x + wo.y, E.y - wo.x);    // E - i*wo
  };
  (void)G;  // G == GW for every plan (tile columns of the row length)
#pragma unroll
  for (int j = 0; j < NGR; ++j) {
    const int g = threadIdx.x / RR + j * GSTEP;
    if (g < GW) {
      const float2 X0 = split(CW * g, wsp[j][0]), X1 = split(CW * g + 1, wsp[j][1]);
      *reinterpret_cast<float4*>(out + ((long)g * hs + y0 + r2) * CW) = make_float4(X0.x, X0.y, X1.x, X1.y);
    }
  }
}

// Inputs of NT correlations k, k+1, .. (clamped to klast, so a prefetch past the end of
// the chunk re-reads its last correlation) for the four-step column transform: the
// leading NL slots (rows t + i*T < hs) of column c; the remaining slots are zero.
template <int HH, int NZJ, int NT>
__device__ __forceinline__ void colpass_load(float2 (&in)[NT][NZJ < V ? NZJ : V], const float2* __restrict__ Rb,
                                             int hs, int G, int g, int c, int t, int k, int klast) {
  constexpr int T = HH / V;
  constexpr int NL = NZJ < V ? NZJ : V;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const float2* rin = Rb + ((long)min(k + n, klast) * G + g) * hs * CW;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int yy = t + i * T;
      in[n][i] = yy < hs ? rin[(long)yy * CW + c] : make_float2(0.f, 0.f);
    }
  }
}

// Column pass, four-step path: NT correlations k .. k+NT-1 of one column through the
// transforms together (twiddles and barriers shared): pruned forward FFT of the inputs
// in `in`, multiply by conj against the image column (held in xi in the permuted
// order), inverse FFT, tiled store.  The inputs of the next iteration (NX correlations
// from k + NT, clamped to klast) are loaded into `in` right after the forward
// transforms, so their latency overlaps the inverse transforms and this iteration's
// stores.
// TS (two columns per block, one correlation per iteration): the result goes through LDS
// so that every thread stores whole rows (both columns, 16 bytes) and 4 consecutive lanes
// fill one 64-byte tile -- full-tile stores instead of 8-byte stores into tile halves.
// xs = NCOL * HH float2 of LDS apart from the transform areas (16-byte aligned).
// Barriers per iteration: the forward and inverse E1 exchanges and one before the
// transposed reads.  The inverse transform's closing barrier is left out (its E1 loads
// are followed by the next forward E1 stores to the same addresses by the same thread),
// and xs needs no barrier after its reads: the next writes to xs come after the next
// iteration's two E1 barriers, which every wave reaches only after its reads.
template <int HH, int NZJ, int NT, int NX, bool FULL, bool TS, int NCOL = CW, class TWT>
__device__ __forceinline__ void colpass_four(float2 (&in)[NT][NZJ < V ? NZJ : V], const float2* __restrict__ Rb,
                                             int hs, int G, int Gs, int g, int c, int t, int k, int klast, int ystore,
                                             bool dc, const float2 (&xi)[V], float* const (&re)[NT],
                                             float* const (&im)[NT], TWT tw, float2* __restrict__ Gb, float2* xs) {
  constexpr int T = HH / V;
  constexpr int NL = NZJ < V ? NZJ : V;
  float2 v[NT][V];
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int i = 0; i < V; ++i) v[n][i] = i < NL ? in[n][i] : make_float2(0.f, 0.f);
  fs_nat2perm_n<HH, false, NZJ, NT, true>(v, re, im, t, tw);
  if constexpr (NX > 0) {
    float2(&nx)[NX][NL] = *reinterpret_cast<float2(*)[NX][NL]>(&in);
    colpass_load<HH, NZJ, NX>(nx, Rb, hs, G, g, c, t, k + NT, klast);
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) {
#pragma unroll
    for (int i = 0; i < V; ++i) v[n][i] = c_mulc(xi[i], v[n][i]);
    if (dc) v[n][0] = make_float2(0.f, 0.f);
  }
  fs_perm2nat_n<HH, true, NT, true, false>(v, re, im, t, tw);
  if constexpr (TS) {
    static_assert(NT == 1, "transposed stores take one correlation per iteration");
    static_assert(NCOL == 1 || NCOL == CW, "columns per block");
    // the block's rows into LDS in the intermediate's own tile order -- 8 float2 per tile
    // row with two columns per block (column 0's 4 rows, then column 1's), the two halves
    // swapped where bit 3 XOR bit 4 of the row is set, so that a wave's 8-byte writes (64
    // consecutive rows of one column) fall in distinct banks; the rows in natural order with
    // one column -- then 16-byte pieces (rows 2h, 2h+1 of one column of a tile) to the
    // intermediate: thread u takes piece u of the block, and a tile row's pieces are one
    // 64-byte run of LDS (a permutation of its 16-byte chunks), so the reads are conflict
    // free; consecutive threads store consecutive pieces, a tile filled by 4 lanes (2 with
    // one column per block, the other column's block filling the rest)
    // (T >= 64 rows per step of i and T / 2 >= 32 tile rows per step of j leave bits 3
    // and 4 of the row's parity unchanged: one swizzle per thread, immediate offsets)
    static_assert(T >= 64, "swizzle constant over the store loops");
    const int cp = NCOL == 1 ? 0 : c;
    const int wslot = NCOL == 1 ? t : 8 * (t >> 2) + 4 * (cp ^ (((t >> 3) ^ (t >> 4)) & 1)) + (t & 3);
#pragma unroll
    for (int i = 0; i < V; ++i) xs[wslot + (NCOL == 1 ? 1 : 2) * i * T] = v[0][i];
    lds_barrier();
    const int q = cp * T + t;  // thread index in the block (NCOL T threads)
    constexpr int PPT = TR / 2 * NCOL;  // pieces per tile written by this block
    float4* gout = reinterpret_cast<float4*>(Gb + (long)k * HH * Gs * CW) + (long)g * (TR * CW / 2);
    const float4* xs4 = reinterpret_cast<const float4*>(xs);
    const int rslot = NCOL == 1 ? q : q ^ (((((q / PPT) >> 1) ^ ((q / PPT) >> 2)) & 1) << 1);
#pragma unroll
    for (int j = 0; j < HH / (2 * T); ++j) {
      const int u = q + j * NCOL * T;  // piece index: tile row u / PPT, piece u % PPT
      const int trow = u / PPT;
      const int col = NCOL == 1 ? c : (u % PPT) / (TR / 2);
      const int h2 = u % (TR / 2);
      const int y = trow * TR + 2 * h2;
      const float4 a = xs4[rslot + j * NCOL * T];
      if (FULL || y < ystore) gout[(long)trow * Gs * (TR * CW / 2) + tile_f4(col, h2)] = a;
    }
  } else {
    // row y = t + i*T lies in tile row y / TR = t / TR + i * (T / TR) at position t % TR;
    // tile rows are Gs tiles apart
    constexpr long ISTRIDE = (long)(T / TR) * TR * CW;  // complex per step of i, per tile column
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      float2* gout = Gb + (long)(k + n) * HH * Gs * CW + ((long)(t / TR) * Gs + g) * TR * CW + tile_off(t % TR, c);
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (FULL || t + i * T < ystore) gout[(long)i * ISTRIDE * Gs] = v[n][i];
    }
  }
}

// Column pass: pruned forward FFT along y (HS of HH rows non-zero), multiply by the
// conjugate against the image spectrum column, inverse FFT along y; tiled store.
// Sizes >= 512 use the four-step transform with NT correlations per iteration; smaller
// sizes use the Stockham passes.  CWB = columns per block (1 or the CW columns of a
// tile).  TWM = 1 reads the twiddles from an LDS copy of the table entries the
// four-step transforms of size HH use; TWM = 2 builds each lane's twiddles once, in
// registers, for all the block's correlations; FULL = every row is stored (ystore == HH).
// (launch bounds: one-column blocks at least 4 waves per SIMD, i.e. two blocks per CU and
// at most 128 VGPRs)
template <int HH, int NZJ, int CWB, int NT, int TWM, bool FULL, bool TS = false>
__global__ __launch_bounds__(CWB*(HH / V), CWB == 1 ? 4 : 1) void k_colpass(const float2* __restrict__ Rb, const float2* __restrict__ Xt,
                                                         int hs, int G, int Gs, int K, int kchunk, int ystore,
                                                         const float2* __restrict__ tw_global, float2* __restrict__ Gb,
                                                         int persist) {
  constexpr int T = HH / V;
  constexpr bool FOUR = UseFour<HH>::value;
  constexpr int HF = FOUR ? HH : 512;
  constexpr int LF = FOUR ? FS<HF>::floats : LdsSize<HH>::floats;
  constexpr int NL = NZJ < V ? NZJ : V;
  static_assert(CWB == CW || (CWB == 1 && FOUR), "one column per block needs the four-step path");
  static_assert(NT == 1 || FOUR, "paired correlations need the four-step path");
  static_assert(TWM == 0 || FOUR, "the LDS twiddle table serves the four-step path");
  // TWM == 1: lane-ordered copy of the four-step twiddle bases (FsTwLayout) after the
  // transforms' work areas
  // TS: the transposed-store planes (one of HH float2 per column) after both
  constexpr int XS0 = CWB * NT * 2 * LF + (TWM == 1 ? FsTwLayout<HF>::floats : 0);
  __shared__ __attribute__((aligned(16))) float lds[XS0 + (TS ? 2 * CWB * HH : 0)];
  // block -> (tile column g, column c): the blocks that write the same 128-byte lines of
  // Gb -- tile columns 2m and 2m+1 (and both columns of each with one column per block)
  // -- get ids 8 apart, so they run on the same XCD (workgroups go to the 8 XCDs round
  // robin) and complete the lines in its L2; the grid is padded to whole groups and the
  // blocks past the last tile column exit
  // Persistent launch (persist > 0; one correlation per iteration): one block per resident
  // slot; the blocks that write the same lines (b and b + 8 with two columns per block; b,
  // b + 8, b + 16, b + 24 with one) form a group on one XCD that takes tile columns 2m and
  // 2m + 1 (both columns) of the same work items (m, k), and group j takes the items
  // [j persist, (j + 1) persist) of the m-major list of ceil(G/2) x K items -- every group
  // gets the same number of correlations, the image column is reloaded where m changes.
  int g = 0, c = 0;
  {
    const int b = blockIdx.x, q = b >> 3;
    if constexpr (CWB == CW) {
      g = 2 * ((q >> 1) * 8 + (b & 7)) + (q & 1);
    } else {
      c = q & 1;
      g = 2 * ((q >> 2) * 8 + (b & 7)) + ((q >> 1) & 1);
    }
    if (persist == 0 && g >= G) return;
  }
  float2* ltw = reinterpret_cast<float2*>(lds + CWB * NT * 2 * LF);
  if constexpr (TWM == 1) {
    fs_tw_fill<HF>(ltw, tw_global, threadIdx.x, CWB * T);
    __syncthreads();
  }
  // four-step: the waves of one column are contiguous; Stockham: columns interleaved
  const int cl = FOUR ? threadIdx.x / T : threadIdx.x % CW;  // column within the block
  const int t = FOUR ? threadIdx.x % T : threadIdx.x / CW;
  const auto tw = [&]() {
    if constexpr (TWM == 2)
      return fs_tw_regs<HF>(tw_global, t);
    else if constexpr (TWM == 1)
      return FsTw<HF>{ltw};
    else
      return tw_global;
  }();
  if constexpr (CWB == CW) c = cl;
  float* base = lds + cl * NT * 2 * LF;
  float2* lds2 = reinterpret_cast<float2*>(lds + XS0);
  static_assert(!TS || (NT == 1 && FOUR), "transposed stores: layout");
  float2 xi[V];
  // image column g (permuted order on the four-step path)
  auto load_image = [&](int gg) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      int ky = t + i * T;
      if constexpr (FOUR) ky = fs_index<HF>(t, i);
      xi[i] = Xt[((long)gg * HH + ky) * CW + c];
    }
  };
  if constexpr (FOUR && NT == 1) {
    // correlations k0 .. k1-1 of tile column gg, the first iteration peeled so that the
    // loop is always entered with the prefetched inputs followed by one iteration's
    // stores in the memory queue (the wait for the inputs then leaves the stores in flight)
    auto segment = [&](int gg, int k0, int k1) {
      load_image(gg);
      const bool dc = (gg == 0 && c == 0 && t == 0);  // ky = 0 sits in slot 0 of thread 0
      float* const re1[1] = {base};
      float* const im1[1] = {base + LF};
      float2 in1[1][NL];
      const int klast = k1 - 1;
      int k = k0;
      colpass_load<HF, NZJ, 1>(in1, Rb, hs, G, gg, c, t, k, klast);
      colpass_four<HF, NZJ, 1, 1, FULL, TS, CWB>(in1, Rb, hs, G, Gs, gg, c, t, k, klast, ystore, dc, xi, re1, im1,
                                                 tw, Gb, lds2);
      for (++k; k < k1; ++k)
        colpass_four<HF, NZJ, 1, 1, FULL, TS, CWB>(in1, Rb, hs, G, Gs, gg, c, t, k, klast, ystore, dc, xi, re1, im1,
                                                   tw, Gb, lds2);
    };
    if (persist > 0) {
      const int b = blockIdx.x, q = b >> 3;
      // group of the blocks that share lines; half = tile column 2m + half of the group
      const int half = CWB == CW ? (q & 1) : ((q >> 1) & 1);
      const int pair = CWB == CW ? (q >> 1) * 8 + (b & 7) : (q >> 2) * 8 + (b & 7);
      const long total = (long)((G + 1) / 2) * K;
      long item = (long)pair * persist;
      const long iend = min(total, item + persist);
      while (item < iend) {
        const int m = (int)(item / K), k0 = (int)(item % K);
        const int k1 = (int)min((long)K, k0 + (iend - item));
        if (2 * m + half < G) segment(2 * m + half, k0, k1);  // uniform over the block
        item += k1 - k0;
      }
    } else {
      const int k0 = blockIdx.y * kchunk;
      const int k1 = min(K, k0 + kchunk);
      if (k0 < k1) segment(g, k0, k1);
    }
    return;
  }
  const int k0 = blockIdx.y * kchunk;
  const int k1 = min(K, k0 + kchunk);
  load_image(g);
  const bool dc = (g == 0 && c == 0 && t == 0);  // ky = 0 sits in slot 0 of thread 0 in both orders
  if constexpr (FOUR) {
    float* const re1[1] = {base};
    float* const im1[1] = {base + LF};
    float2 in[NT][NL];
    float2(&in1)[1][NL] = *reinterpret_cast<float2(*)[1][NL]>(&in);
    int k = k0;
    if constexpr (NT > 1) {
      float* const re2[2] = {base, base + 2 * LF};
      float* const im2[2] = {base + LF, base + 3 * LF};
      const int klast = k1 - 1;
      // the first iteration is peeled so that the loop is always entered with the
      // prefetched inputs followed by one iteration's stores in the memory queue (the
      // wait for the inputs then leaves the stores in flight)
      if (k + 1 < k1) {
        colpass_load<HF, NZJ, 2>(in, Rb, hs, G, g, c, t, k, klast);
        colpass_four<HF, NZJ, 2, 2, FULL, TS, CWB>(in, Rb, hs, G, Gs, g, c, t, k, klast, ystore, dc, xi, re2, im2, tw, Gb, lds2);
        k += 2;
      }
      for (; k + 1 < k1; k += 2)
        colpass_four<HF, NZJ, 2, 2, FULL, TS, CWB>(in, Rb, hs, G, Gs, g, c, t, k, klast, ystore, dc, xi, re2, im2, tw, Gb, lds2);
      // a single last correlation: its input was prefetched into in[0], unless the chunk
      // holds only this one
      if (k < k1) {
        if (k == k0) colpass_load<HF, NZJ, 1>(in1, Rb, hs, G, g, c, t, k, klast);
        colpass_four<HF, NZJ, 1, 0, FULL, TS, CWB>(in1, Rb, hs, G, Gs, g, c, t, k, klast, ystore, dc, xi, re1, im1, tw, Gb, lds2);
      }
    } else {
      const int klast = k1 - 1;
      if (k < k1) {  // peeled first iteration (see the paired loop)
        colpass_load<HF, NZJ, 1>(in1, Rb, hs, G, g, c, t, k, klast);
        colpass_four<HF, NZJ, 1, 1, FULL, TS, CWB>(in1, Rb, hs, G, Gs, g, c, t, k, klast, ystore, dc, xi, re1, im1, tw, Gb, lds2);
        ++k;
      }
      for (; k < k1; ++k)
        colpass_four<HF, NZJ, 1, 1, FULL, TS, CWB>(in1, Rb, hs, G, Gs, g, c, t, k, klast, ystore, dc, xi, re1, im1, tw, Gb, lds2);
    }
  } else {
    float* lre = base;
    float* lim = base + LF;
    for (int k = k0; k < k1; ++k) {
      const float2* rin = Rb + ((long)k * G + g) * hs * CW;
      float2 v[V];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const int yy = t + i * T;
        v[i] = (i < NZJ && yy < hs) ? rin[(long)yy * CW + c] : make_float2(0.f, 0.f);
      }
      fft_wg<HH, V, false, NZJ>(v, lre, lim, t, tw);
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = c_mulc(xi[i], v[i]);
      if (dc) v[0] = make_float2(0.f, 0.f);
      fft_wg<HH, V, true>(v, lre, lim, t, tw);
      float2* gout = Gb + (long)k * HH * Gs * CW;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const int yy = t + i * T;
        if (FULL || yy < ystore) gout[((long)(yy / TR) * Gs + g) * TR * CW + tile_off(yy % TR, c)] = v[i];
      }
    }
  }
}

__device__ __forceinline__ void two_sum(float& hi, float& lo, float b) {
  const float s = hi + b;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (b - bb);
  hi = s;
  lo += err;
}

// Row-pass geometry for a W-point row (M = W/2 complex points) with RB of the TR rows of
// a tile row per block.
template <int WS, int RB = TR>
struct RowCfg {
  static_assert(RB == 1 || RB == 2 || RB == 4, "rows per block");
  static constexpr int M = WS / 2;
  static constexpr int T = M / V;
  static constexpr int BLOCK = RB * T;                                     // threads per block
  static constexpr bool FOUR = UseFour<M>::value;
  static constexpr int MF = FOUR ? M : 512;
  static constexpr int LF = FOUR ? FS<MF>::floats : LdsSize<M>::floats;  // per component per row
  static constexpr int G = (M + 1 + CW - 1) / CW;                          // tile columns
  static constexpr int GS = G + (G & 1);                                   // tile stride (even)
  static constexpr int TILEF = GS * TR * CW;                               // complex per tile row in HBM
  static constexpr int TILE = G * RB * CW;                                 // complex staged per block
  static constexpr int RSTR = 2 * LF + 16;                                 // floats per row region
  static constexpr int REGION = RB * RSTR;                                 // floats, FFT work area
  static constexpr int BUF = REGION > 2 * TILE ? REGION : 2 * TILE;        // floats per LDS buffer
};

// Padded index of staged row samples (split re / im per row): keeps the four-step read
// order of the c2r input within 2-way bank conflicts.
__device__ __forceinline__ int stg_pad(int i) { return i + (i >> 5); }

// Asynchronous copy of one row-tile block (TILE complex, contiguous in HBM) into LDS with
// global_load_lds_dwordx4: wave w copies 64 x 16-byte chunks per round to consecutive
// LDS addresses, so LDS holds an exact copy of the HBM layout.  Completion is awaited
// with s_waitcnt vmcnt before the consuming barrier.
template <int WS>
__device__ __forceinline__ void tile_block_async(const float2* __restrict__ src, float* dst) {
  constexpr int NF4 = RowCfg<WS>::TILE / 2;
  constexpr int BLOCK = TR * RowCfg<WS>::T;             // threads per block (the launch size)
  constexpr int WL = BLOCK < 64 ? BLOCK : 64;           // active lanes per wave
  constexpr int STEP = BLOCK;                            // chunks copied per round by the block
  static_assert(BLOCK % WL == 0, "block must be whole waves or a single partial wave");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float4* s4 = reinterpret_cast<const float4*>(src);
  float4* d4 = reinterpret_cast<float4*>(dst);
  for (int base = wave * WL; base < NF4; base += STEP) {
    if (base + lane < NF4)
      __builtin_amdgcn_global_load_lds((const void*)(s4 + base + lane),
                                       (__attribute__((address_space(3))) void*)(d4 + base), 16, 0, 0);
  }
}

// Correlation loop shared by the match and refine row passes.  For every correlation k:
// wait for its tile block (prefetched into one of two LDS buffers during k-1), start the
// prefetch of k+1 into the other buffer, build the half-length complex input of the
// c2r from the staged tile (imaginary parts of DC / Nyquist dropped), run the inverse
// transform in the same buffer, and hand the natural-order result to `epi(k, v)`:
// thread (row r, t) holds v[i] = (x[2n], x[2n+1]) for n = t + i*T of row y = p*TR + r.
// Synchronous tile staging: 8-byte loads (one row of one column of a tile), stored
// de-interleaved into the per-row split re / im regions at stg_pad positions.
// `src` points at the block's first row inside the tile row's first tile; element (row
// rr, column c) of tile g is at src[g TR CW + tile_off(rr, c)].  Element e of the block
// is (tile g, column c, row rr) with rr fastest, so consecutive threads read consecutive
// addresses of a column's rows.
template <int WS, int RB>
__device__ __forceinline__ void tile_stage_one(float* dst, int e, float2 x) {
  using C = RowCfg<WS, RB>;
  const int rr = e % RB, c = (e / RB) % CW, g = e / (RB * CW);
  const int a = stg_pad(CW * g + c);
  float* re = dst + rr * C::RSTR;
  re[a] = x.x;
  re[a + C::LF] = x.y;
}

template <int WS, int RB>
__device__ __forceinline__ const float2& tile_src(const float2* __restrict__ src, int e) {
  const int rr = e % RB, c = (e / RB) % CW, g = e / (RB * CW);
  return src[g * TR * CW + tile_off(rr, c)];
}

template <int WS, int RB>
__device__ __forceinline__ void tile_block_sync(const float2* __restrict__ src, float* dst) {
  using C = RowCfg<WS, RB>;
  for (int e = threadIdx.x; e < C::TILE; e += C::BLOCK) tile_stage_one<WS, RB>(dst, e, tile_src<WS, RB>(src, e));
}

// Synchronous tile staging with every load of the thread issued before the first LDS
// store (one memory latency per tile instead of one per round).
template <int WS, int RB>
__device__ __forceinline__ void tile_block_sync_batched(const float2* __restrict__ src, float* dst) {
  using C = RowCfg<WS, RB>;
  constexpr int NE = C::TILE;
  constexpr int BLOCK = C::BLOCK;
  constexpr int ROUNDS = (NE + BLOCK - 1) / BLOCK;
  static_assert((ROUNDS - 1) * BLOCK < NE && ROUNDS * BLOCK >= NE, "only the last round is partial");
  float2 x[ROUNDS];
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j) {
    const int e = threadIdx.x + j * BLOCK;
    if (j + 1 < ROUNDS || e < NE) x[j] = tile_src<WS, RB>(src, e);
  }
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j) {
    const int e = threadIdx.x + j * BLOCK;
    if (j + 1 < ROUNDS || e < NE) tile_stage_one<WS, RB>(dst, e, x[j]);
  }
}

// Reformat of a raw staged tile block (HBM layout: whole tiles, TR rows) into the
// de-interleaved per-row regions: consecutive 16-byte reads (rows 2h, 2h+1 of one column
// of a tile), 4-byte writes into two rows.
template <int WS>
__device__ __forceinline__ void tile_block_reformat(const float* raw, float* dst) {
  using C = RowCfg<WS>;
  constexpr int NF4 = C::TILE / 2;
  constexpr int BLOCK = TR * C::T;
  constexpr int ROUNDS = (NF4 + BLOCK - 1) / BLOCK;
  constexpr int P4 = TR * CW / 2;  // float4 per tile
  const float4* s4 = reinterpret_cast<const float4*>(raw);
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j) {
    const int f = threadIdx.x + j * BLOCK;
    if (f < NF4) {
      const float4 x = s4[f];
      const int g = f / P4, c = (f % P4) / (TR / 2), r = 2 * (f % (TR / 2));
      const int a = stg_pad(CW * g + c);
      float* re0 = dst + r * C::RSTR;
      float* re1 = re0 + C::RSTR;
      re0[a] = x.x;
      re0[a + C::LF] = x.y;
      re1[a] = x.z;
      re1[a + C::LF] = x.w;
    }
  }
}

// Tile staging modes of the row pass: 0 = synchronous load into the de-interleaved
// per-row regions; 1 = asynchronous raw copy (double-buffered) read in place; 2 =
// asynchronous raw copy of the next tile while the current one is reformatted into the
// de-interleaved regions.
template <int WS, int MODE, int RB = TR>
struct RowLds {
  static_assert(RB == TR || MODE == 0 || MODE == 3, "asynchronous staging copies whole tile rows");
  static constexpr int floats = (MODE == 0 || MODE == 3) ? RowCfg<WS, RB>::BUF
                              : MODE == 1                ? 2 * RowCfg<WS>::BUF
                                                         : 2 * RowCfg<WS>::TILE + RowCfg<WS>::REGION;
};

// log2(TW_N / WS): a W-point row pass reads only table entries that are multiples of
// TW_N / W and below TW_N / 2 (pre-twiddles m < W/2, transform twiddles of the W/2-point
// four-step), so its LDS twiddle copy holds W / 2 entries.
// With the four-step transform (M >= 512) the copy is lane-ordered instead: the
// FsTwLayout<M> stage bases followed by the pre-twiddle bases PB[t] (see rowpass_loop).
template <int WS>
struct RowTw {
  static constexpr int SH = tw_shift(WS);
  static_assert((WS << SH) == TW_N, "row length must divide the twiddle table");
  static constexpr int entries = WS / 2;
  static constexpr bool FOUR = RowCfg<WS>::FOUR;
  static constexpr int MF = RowCfg<WS>::MF;
  static constexpr int floats = FOUR ? FsTwLayout<MF>::floats + 2 * RowCfg<WS>::T : 2 * entries;
};

// W_16^i = e^{-2 pi i i / 16}: the pre-twiddle of slot i of a four-step row thread is
// W_W^{fs_index(t, i)} = PB[t] W_16^i with PB[t] = W_W^{fs_index(t, 0)}, because
// fs_index(t, i) = fs_index(t, 0) + (M / 8) i and W_W^{M/8} = W_16.
__device__ __forceinline__ float2 w16_times(float2 a, int i) {
  constexpr float C1 = 0.92387953251128675613f, S1 = 0.38268343236508977173f, H = 0.70710678118654752440f;
  switch (i) {
    case 0: return a;
    case 1: return c_mul(a, make_float2(C1, -S1));
    case 2: return c_mul(a, make_float2(H, -H));
    case 3: return c_mul(a, make_float2(S1, -C1));
    case 4: return make_float2(a.y, -a.x);
    case 5: return c_mul(a, make_float2(-S1, -C1));
    case 6: return c_mul(a, make_float2(-H, -H));
    default: return c_mul(a, make_float2(-C1, -S1));
  }
}

template <int WS, int MODE, int TWM, int RB, class Epi>
__device__ __forceinline__ void rowpass_loop(const float2* __restrict__ Gb, int HH, int kbeg, int kend, int p,
                                             int r0, const float2* __restrict__ tw_global, float* lds, Epi&& epi) {
  using C = RowCfg<WS, RB>;
  constexpr int M = C::M, T = C::T, LF = C::LF, MF = C::MF;
  constexpr bool ASYNC = MODE == 1;
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  // twiddle source: the global table, or (TWM == 1) an LDS copy after the staging / work
  // area -- lane-ordered stage and pre-twiddle bases for the four-step transform, every
  // (TW_N / WS)-th table entry otherwise
  float2* ltw = reinterpret_cast<float2*>(lds + RowLds<WS, MODE, RB>::floats);
  float2* lpb = ltw + (C::FOUR ? FsTwLayout<MF>::entries : 0);  // PB[t] (four-step)
  if constexpr (TWM == 1) {
    if constexpr (C::FOUR) {
      fs_tw_fill<MF>(ltw, tw_global, threadIdx.x, C::BLOCK);
      for (int j = threadIdx.x; j < T; j += C::BLOCK) lpb[j] = tw_global[fs_index<MF>(j, 0) * (TW_N / WS)];
    } else {
      for (int j = threadIdx.x; j < RowTw<WS>::entries; j += C::BLOCK) ltw[j] = tw_global[j << RowTw<WS>::SH];
    }
    __syncthreads();
  }
  const auto tw = [&]() {
    if constexpr (TWM == 1 && C::FOUR)
      return FsTw<MF>{ltw};
    else if constexpr (TWM == 1)
      return TwLds<RowTw<WS>::SH>{ltw};
    else
      return tw_global;
  }();
  const long kstride = (long)(HH / TR) * C::TILEF;
  const float2* gsrc = Gb + (long)p * C::TILEF + tile_off(r0, 0);
  float* raw2 = lds;                  // MODE 2: raw tile
  float* work2 = lds + 2 * C::TILE;   // MODE 2: de-interleaved regions / transform work area
  const int K = kend;
  if (MODE == 1 && kbeg < K) tile_block_async<WS>(gsrc + kbeg * kstride, lds + (kbeg & 1) * C::BUF);
  if (MODE == 2 && kbeg < K) tile_block_async<WS>(gsrc + kbeg * kstride, raw2);
  for (int k = kbeg; k < K; ++k) {
    float* cur = MODE == 1 ? lds + (k & 1) * C::BUF : (MODE == 2 ? work2 : lds);
    float* nxt = lds + ((k + 1) & 1) * C::BUF;
    if (MODE == 1) {
      __builtin_amdgcn_s_waitcnt(0);  // this thread's part of tile block k has landed
      __syncthreads();                // everybody's part landed; iteration k-1 is done with nxt
      if (k + 1 < K) tile_block_async<WS>(gsrc + (k + 1) * kstride, nxt);
    } else if (MODE == 2) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();                // tile k landed in raw; iteration k-1 is done with work2
      tile_block_reformat<WS>(raw2, work2);
      __syncthreads();                // raw consumed, work2 staged
      if (k + 1 < K) tile_block_async<WS>(gsrc + (k + 1) * kstride, raw2);
    } else if (MODE == 3) {
      tile_block_sync_batched<WS, RB>(gsrc + k * kstride, cur);
      __syncthreads();
    } else {
      tile_block_sync<WS, RB>(gsrc + k * kstride, cur);
      __syncthreads();
    }
    const float2* stg = reinterpret_cast<const float2*>(cur);
    const float* sre = cur + r * C::RSTR;
    const float* sim = sre + LF;
    float2 v[V];
    const auto twp = opaque_offset(tw);  // reload the pre-twiddles per correlation (no hoisting)
    float2 pb = make_float2(1.f, 0.f);   // four-step: PB[t], the pre-twiddle base
    if constexpr (C::FOUR) {
      if constexpr (TWM == 1)
        pb = opaque_offset(static_cast<const float2*>(lpb))[t];
      else
        pb = twp[fs_index<MF>(t, 0) * (TW_N / WS)];
    }
#pragma unroll
    for (int i = 0; i < V; ++i) {
      int mm = t + i * T;
      if constexpr (C::FOUR) mm = fs_index<MF>(t, i);
      const int bm = M - mm;
      float2 A, Bc;
      if (ASYNC) {  // raw copy of the HBM tile layout
        A = stg[(mm >> 1) * TR * CW + tile_off(r, mm & 1)];
        Bc = stg[(bm >> 1) * TR * CW + tile_off(r, bm & 1)];
      } else {
        A = make_float2(sre[stg_pad(mm)], sim[stg_pad(mm)]);
        Bc = make_float2(sre[stg_pad(bm)], sim[stg_pad(bm)]);
      }
      Bc.y = -Bc.y;
      if (mm == 0) {
        A.y = 0.f;
        Bc.y = 0.f;
      }
      float2 w;
      if constexpr (C::FOUR)
        w = c_conj(w16_times(pb, i));
      else
        w = c_conj(twp[mm * (TW_N / WS)]);
      const float2 dsum = c_add(A, Bc), od = c_mul(w, c_sub(A, Bc));
      v[i] = make_float2(dsum.x - od.y, dsum.y + od.x);
    }
    lds_barrier();  // staged tile consumed: the buffer becomes the transform's work area
    float* lre = cur + r * C::RSTR;
    float* lim = lre + LF;
    if constexpr (C::FOUR) {
      fs_perm2nat<MF, true, true>(v, lre, lim, t, tw);
    } else {
      fft_wg<M, V, true>(v, lre, lim, t, tw);
      lds_barrier();
    }
    epi(k, v);
  }
}

// Row pass of the search: fused MIP / argmax / sum / sum-of-squares over the K
// correlations of the step, then one merge into the running maps.
// Block -> (tile row p, row group sub): with RB < TR rows per block, the TR / RB blocks of
// one tile row get ids 8 apart so they run on the same XCD (workgroups are dispatched to
// the 8 XCDs round robin) and share its L2 for the 64-byte tiles they split.
template <int RB>
__device__ __forceinline__ void rowpass_block(int& p, int& sub) {
  constexpr int NS = TR / RB;
  const int b = blockIdx.x;
  if constexpr (NS == 1) {
    p = b;
    sub = 0;
  } else {
    const int rest = b >> 3;
    sub = rest % NS;
    p = (rest / NS) * 8 + (b & 7);
  }
}

// (launch bounds: at least 4 waves per SIMD, i.e. at most 128 VGPRs, so that with RB < TR
// several blocks share a CU)
template <int WS, int MODE, int TWM, int RB>
__global__ __launch_bounds__(RB*(WS / 2) / V, 4) void k_rowpass(const float2* __restrict__ Gb, int HH, int G, int K,
                                                            int nb, int nd, int O, int o0,
                                                            const float2* __restrict__ tw, float* __restrict__ mip,
                                                            int* __restrict__ bidx, float* __restrict__ shi,
                                                            float* __restrict__ slo, float* __restrict__ qhi,
                                                            float* __restrict__ qlo, const int* __restrict__ o0p) {
  if (o0p != nullptr) o0 = *o0p;  // device-side first orientation (graph replays)
  using C = RowCfg<WS>;
  constexpr int T = C::T;
  __shared__ __attribute__((aligned(16))) float lds[RowLds<WS, MODE, RB>::floats + (TWM ? RowTw<WS>::floats : 0)];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  int p, sub;  // tile row, row group within it
  rowpass_block<RB>(p, sub);
  const int y = p * TR + sub * RB + r;
  (void)G;
  // running max / sum / sum of squares per pixel; the correlation index of the max is
  // kept as a 16-bit step-local index, two pixels per register (K < 65536)
  float m[2 * V], s[2 * V], q[2 * V];
  unsigned ix2[V];
#pragma unroll
  for (int j = 0; j < 2 * V; ++j) {
    m[j] = -INFINITY;
    s[j] = 0.f;
    q[j] = 0.f;
  }
#pragma unroll
  for (int i = 0; i < V; ++i) ix2[i] = 0xffffffffu;
  rowpass_loop<WS, MODE, TWM, RB>(Gb, HH, 0, K, p, sub * RB, tw, lds, [&](int k, float2 (&v)[V]) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const float a0 = v[i].x, a1 = v[i].y;
      if (a0 > m[2 * i]) {
        m[2 * i] = a0;
        ix2[i] = (ix2[i] & 0xffff0000u) | (unsigned)k;
      }
      if (a1 > m[2 * i + 1]) {
        m[2 * i + 1] = a1;
        ix2[i] = (ix2[i] & 0x0000ffffu) | ((unsigned)k << 16);
      }
      s[2 * i] += a0;
      s[2 * i + 1] += a1;
      q[2 * i] = fmaf(a0, a0, q[2 * i]);
      q[2 * i + 1] = fmaf(a1, a1, q[2 * i + 1]);
    }
  });
  const long rowbase = (long)y * WS;
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const long px = rowbase + 2 * (t + i * T);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const long e = px + h;
      const int j = 2 * i + h;
      const float mo = mip[e];
      if (m[j] > mo) {
        const int kk = (int)((ix2[i] >> (16 * h)) & 0xffffu);
        const int bb = kk % nb, dd = (kk / nb) % nd, ci = kk / (nb * nd);
        mip[e] = m[j];
        bidx[e] = ci * nd * O + dd * O + o0 + bb;
      }
      float sh = shi[e], sl = slo[e];
      two_sum(sh, sl, s[j]);
      shi[e] = sh;
      slo[e] = sl;
      float qh = qhi[e], ql = qlo[e];
      two_sum(qh, ql, q[j]);
      qhi[e] = qh;
      qlo[e] = ql;
    }
  }
}

// Device-side first orientation of the next captured step.
__global__ void k_set_start(int* o0p, int v) { *o0p = v; }
__global__ void k_advance_start(int* o0p, int step) { *o0p += step; }
// Batched refinement: state = {first offset of the next batch, its result slot, particle}.
__global__ void k_refine_state(int* state, int o0, int slot, int part, const float* __restrict__ pose,
                               float* __restrict__ pose_dst) {
  state[0] = o0;
  state[1] = slot;
  state[2] = part;
  pose_dst[0] = pose[0];
  pose_dst[1] = pose[1];
  pose_dst[2] = pose[2];
}
__global__ void k_refine_advance(int* state, int step) {
  state[0] += step;
  state[1] += 1;
}

__global__ void k_stats_init(long n, float* mip, int* idx, float* a, float* b, float* c, float* d) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  mip[e] = -INFINITY;
  idx[e] = -1;
  a[e] = 0.f;
  b[e] = 0.f;
  c[e] = 0.f;
  d[e] = 0.f;
}

__global__ void k_stats_merge(long n, float* mip, int* idx, float* shi, float* slo, float* qhi, float* qlo,
                              const float* mip_b, const int* idx_b, const float* shi_b, const float* slo_b,
                              const float* qhi_b, const float* qlo_b) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (mip_b[e] > mip[e]) {
    mip[e] = mip_b[e];
    idx[e] = idx_b[e];
  }
  float h = shi[e], l = slo[e];
  two_sum(h, l, shi_b[e]);
  l += slo_b[e];
  shi[e] = h;
  slo[e] = l;
  h = qhi[e];
  l = qlo[e];
  two_sum(h, l, qhi_b[e]);
  l += qlo_b[e];
  qhi[e] = h;
  qlo[e] = l;
}

__global__ void k_finalize(long n, double total, const float* __restrict__ mip, const int* __restrict__ bidx,
                           const float* __restrict__ shi, const float* __restrict__ slo,
                           const float* __restrict__ qhi, const float* __restrict__ qlo,
                           const float* __restrict__ euler, int O, const float* __restrict__ dvals, int D,
                           float* zs, float* mean_o, float* std_o, float* phi, float* theta, float* psi,
                           float* df) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double mean = ((double)shi[e] + (double)slo[e]) / total;
  double var = ((double)qhi[e] + (double)qlo[e]) / total - mean * mean;
  const double sd = sqrt(var > 0.0 ? var : 0.0);
  const double z = sd != 0.0 ? ((double)mip[e] - mean) / sd : 0.0;
  zs[e] = (float)z;
  mean_o[e] = (float)mean;
  std_o[e] = (float)sd;
  const long stride = (long)D * O;
  long rem = (long)bidx[e] % stride;
  if (rem < 0) rem += stride;
  const int d = (int)(rem / O), o = (int)(rem % O);
  phi[e] = euler[3 * o];
  theta[e] = euler[3 * o + 1];
  psi[e] = euler[3 * o + 2];
  df[e] = dvals[d];
}


// ------------------------------------------------------------------ refinement
// Magnification matrix of the CTF frequencies (on = 0: none).
struct CtfMag {
  double a00, a01, a10, a11;
  int on;
};

// CTF x B-envelope stack for one particle: (C, D, h, h/2+1), torch_ctf sign convention.
// `mag` (on): the CTF at frequency k is evaluated at A k (anisotropic magnification,
// utils/ctf_utils.py:66-100), the B-envelope at k (:103-111).
__global__ void k_ctf_stack(int h, int D, int C, double df_mean, double astig, double astig_ang,
                            const float* __restrict__ dfo, const float* __restrict__ pxo, double lam, double cs_a,
                            double px, double wac, double phase, double bfac, CtfMag mag, float* __restrict__ out) {
  const int wr = h / 2 + 1;
  const long plane = (long)h * wr;
  const long total = plane * D * C;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int kx = e % wr;
  const int r = (e / wr) % h;
  const int d = (e / plane) % D;
  const int c = e / (plane * D);
  const double ky0 = (double)(r < (h + 1) / 2 ? r : r - h) / (double)h / px;
  const double kx0 = (double)kx / (double)h / px;
  const double ky = mag.on ? mag.a10 * kx0 + mag.a11 * ky0 : ky0;
  const double kxx = mag.on ? mag.a00 * kx0 + mag.a01 * ky0 : kx0;
  const double k2 = ky * ky + kxx * kxx;
  const double ang = atan2(ky, kxx);
  const double ratio = (px + (double)pxo[c]) / px;
  const double cs = cs_a / (ratio * ratio * ratio * ratio);
  const double df = (double)dfo[d] + df_mean + astig * cos(2.0 * (ang - astig_ang));
  const double chi = -M_PI * lam * k2 * df + 0.5 * M_PI * cs * lam * lam * lam * k2 * k2 - wac - phase;
  const double k2e = ky0 * ky0 + kx0 * kx0;
  out[e] = (float)(-sin(chi) * exp(-bfac * k2e / 4.0));
}

struct RefineBest {
  float z;
  float cc;
  int pix;
  int k;
};

// better candidate: larger z; equal z -> smaller flat pixel index (the first maximum of
// the per-pixel maximum map); same pixel -> smaller correlation index (torch.max over
// the correlation axis keeps the first maximum)
__device__ __forceinline__ bool ref_better(float z, int pix, int k, float zb, int pixb, int kb) {
  return z > zb || (z == zb && (pix < pixb || (pix == pixb && k < kb)));
}

// Row pass of a refinement batch: c2r along x for the rows of the valid region and the
// z-score epilogue z = (cc - mean) / std on the valid crop.  Per pixel the first
// correlation reaching the pixel's maximum z is kept (torch.max over the flattened
// (pixel size, defocus, offset) axis); each block then reduces to its best
// (z, pixel, correlation) and its maximum raw cc.  For a fixed pixel z is increasing in
// cc (std > 0), so the loop tracks each pixel's maximum cc and its first correlation, and
// z is evaluated once per pixel after the loop.  Block (tile row, chunk): the K
// correlations are split into chunks of kchunk so that the grid fills the GPU even for
// small crops (a 512 x 512 crop with a 256^3 template has 65 tile rows).
// partp != nullptr (graph replays of a batched refinement): mean / stdv are the bases of
// per-particle maps `plane` floats apart and *partp is the particle.
template <int WS>
__global__ __launch_bounds__(TR*(WS / 2) / V) void k_rowpass_refine(const float2* __restrict__ Gb, int HH, int K,
                                                                   int kchunk, int crop_h, int crop_w,
                                                                   const float* __restrict__ mean,
                                                                   const float* __restrict__ stdv, long plane,
                                                                   const int* __restrict__ partp,
                                                                   const float2* __restrict__ tw,
                                                                   RefineBest* __restrict__ part) {
  if (partp != nullptr) {
    mean += (long)*partp * plane;
    stdv += (long)*partp * plane;
  }
  using C = RowCfg<WS>;
  constexpr int T = C::T;
  __shared__ __attribute__((aligned(16))) float lds[RowLds<WS, 3>::floats + RowTw<WS>::floats];
  __shared__ RefineBest red[TR * T];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int p = blockIdx.x;
  const int y = p * TR + r;
  const int k0 = blockIdx.y * kchunk;
  const int k1 = min(K, k0 + kchunk);
  float cb[2 * V];
  unsigned kb2[V];
  float ccmax = -INFINITY;
#pragma unroll
  for (int j = 0; j < 2 * V; ++j) cb[j] = -INFINITY;
#pragma unroll
  for (int i = 0; i < V; ++i) kb2[i] = 0xffffffffu;
  rowpass_loop<WS, 3, 1, TR>(Gb, HH, k0, k1, p, 0, tw, lds, [&](int k, float2 (&v)[V]) {
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int x = 2 * (t + i * T) + hh;
        if (y < crop_h && x < crop_w) {
          const float c = hh ? v[i].y : v[i].x;
          ccmax = fmaxf(ccmax, c);
          if (c > cb[2 * i + hh]) {
            cb[2 * i + hh] = c;
            kb2[i] = hh ? ((kb2[i] & 0x0000ffffu) | ((unsigned)k << 16)) : ((kb2[i] & 0xffff0000u) | (unsigned)k);
          }
        }
      }
  });
  RefineBest b{-INFINITY, ccmax, 0x7fffffff, 0x7fffffff};
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int x = 2 * (t + i * T) + hh;
      const unsigned kk = (kb2[i] >> (16 * hh)) & 0xffffu;
      if (y < crop_h && x < crop_w && kk != 0xffffu) {
        const int pix = y * crop_w + x;
        const float z = (cb[2 * i + hh] - mean[pix]) / stdv[pix];
        if (ref_better(z, pix, (int)kk, b.z, b.pix, b.k)) {
          b.z = z;
          b.pix = pix;
          b.k = (int)kk;
        }
      }
    }
  red[threadIdx.x] = b;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      RefineBest o = red[threadIdx.x + off];
      RefineBest sb = red[threadIdx.x];
      sb.cc = fmaxf(sb.cc, o.cc);
      if (ref_better(o.z, o.pix, o.k, sb.z, sb.pix, sb.k)) {
        sb.z = o.z;
        sb.pix = o.pix;
        sb.k = o.k;
      }
      red[threadIdx.x] = sb;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Reduce the row-pass blocks of one batch into slot bi of `slots` (slotp != nullptr:
// slot *slotp, graph replays).
__global__ void k_refine_reduce(const RefineBest* __restrict__ part, int n, RefineBest* __restrict__ slots, int bi,
                                const int* __restrict__ slotp) {
  RefineBest* slot = slots + (slotp != nullptr ? *slotp : bi);
  __shared__ RefineBest red[256];
  RefineBest b{-INFINITY, -INFINITY, 0x7fffffff, 0x7fffffff};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    RefineBest o = part[i];
    b.cc = fmaxf(b.cc, o.cc);
    if (ref_better(o.z, o.pix, o.k, b.z, b.pix, b.k)) {
      b.z = o.z;
      b.pix = o.pix;
      b.k = o.k;
    }
  }
  red[threadIdx.x] = b;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      RefineBest o = red[threadIdx.x + off];
      RefineBest s = red[threadIdx.x];
      s.cc = fmaxf(s.cc, o.cc);
      if (ref_better(o.z, o.pix, o.k, s.z, s.pix, s.k)) {
        s.z = o.z;
        s.pix = o.pix;
        s.k = o.k;
      }
      red[threadIdx.x] = s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *slot = red[0];
}

// Pick the first batch whose z maximum strictly beats the running best
// (core_refine_template.py:725) and decode it; compose the offset with the pose and
// wrap the angles (core_refine_template.py:504-536).  out[11]: max_cc, max_z, phi,
// theta, psi (composed, wrapped), defocus offset, pixel-size offset, pos_y, pos_x
// (valid-crop coordinates), angle_idx, and the phi/theta/psi offsets in out[11..13].
__global__ void k_refine_select(const RefineBest* __restrict__ slots, int nbatch, int batch, int n_off, int D,
                                const float* __restrict__ off, const float* __restrict__ base,
                                const float* __restrict__ dfo, const float* __restrict__ pxo, int crop_w,
                                float* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float best = -1e9f;
  int bi = -1;
  for (int i = 0; i < nbatch; ++i)
    if (slots[i].z > best) {
      best = slots[i].z;
      bi = i;
    }
  float res[14] = {-1e9f, -1e9f, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  float phio = 0.f, theo = 0.f, psio = 0.f;
  if (bi >= 0) {
    const RefineBest s = slots[bi];
    const int s0 = bi * batch;
    const int nb = min(batch, n_off - s0);
    const int f = s.k;
    const int px = f / (D * nb), d = (f / nb) % D, a = f % nb;
    res[0] = s.cc;
    res[1] = s.z;
    phio = off[3 * (s0 + a)];
    theo = off[3 * (s0 + a) + 1];
    psio = off[3 * (s0 + a) + 2];
    res[5] = dfo[d];
    res[6] = pxo[px];
    res[7] = (float)(s.pix / crop_w);
    res[8] = (float)(s.pix % crop_w);
    res[9] = (float)(s0 + a);
  }
  // compose R(offset) @ R(pose) and convert back to ZYZ degrees
  float ro[9], rb[9], rc[9];
  const float oe[3] = {phio, theo, psio};
  rotmat_zyz(oe, ro);
  rotmat_zyz(base, rb);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) rc[3 * i + j] = ro[3 * i] * rb[j] + ro[3 * i + 1] * rb[3 + j] + ro[3 * i + 2] * rb[6 + j];
  const double r22 = fmin(1.0, fmax(-1.0, (double)rc[8]));
  double theta = acos(r22);
  double phi, psi;
  if (fabs(sin(theta)) < 1e-7) {
    phi = r22 > 0 ? atan2((double)rc[3], (double)rc[0]) : atan2(-(double)rc[3], -(double)rc[0]);
    psi = 0.0;
  } else {
    phi = atan2((double)rc[5], (double)rc[2]);
    psi = atan2((double)rc[7], -(double)rc[6]);
  }
  const double r2d = 57.29577951308232;
  float p = (float)(phi * r2d), th = (float)(theta * r2d), ps = (float)(psi * r2d);
  if (p < 0) p += 360.f;
  if (th < 0) th += 180.f;
  if (ps < 0) ps += 360.f;
  res[2] = p;
  res[3] = th;
  res[4] = ps;
  res[11] = phio;
  res[12] = theo;
  res[13] = psio;
  for (int i = 0; i < 14; ++i) out[i] = res[i];
}

}  // namespace tdtm

// CTF x B-envelope stack (C, D, h, h/2+1) from the optics parameters ctf_params =
// {voltage kV, Cs mm, amplitude contrast, phase shift deg, pixel size A, B-factor A^2}
// (utils/ctf_utils.py:19-116): defocus = offsets + (u + v)/2, astigmatism |u - v|/2
// at `defocus_angle` degrees, Cs / ((p + dp)/p)^4 per pixel size, relativistic
// wavelength from CODATA constants.
namespace tdtm {
static int launch_ctf_stack(int h, int D, int C, double defocus_u, double defocus_v, double defocus_angle,
                            const float* defocus_offsets, const float* pixel_offsets, const double* ctf_params,
                            const double* mag, float* out, hipStream_t st) {
  const double voltage = ctf_params[0], cs_mm = ctf_params[1], ac = ctf_params[2], phase = ctf_params[3];
  const double px = ctf_params[4], bfac = ctf_params[5];
  if (!(voltage > 0) || !(px > 0) || !(ac >= 0 && ac < 1)) return fail("invalid CTF parameters");
  const double v = voltage * 1e3;
  const double lam = 6.62607015e-34 /
                     std::sqrt(2 * 9.1093837015e-31 * 1.602176634e-19 * v *
                               (1 + 1.602176634e-19 * v / (2 * 9.1093837015e-31 * 299792458.0 * 299792458.0))) *
                     1e10;
  const long tot = (long)h * (h / 2 + 1) * D * C;
  CtfMag cm{1.0, 0.0, 0.0, 1.0, 0};
  if (mag) cm = CtfMag{mag[0], mag[1], mag[2], mag[3], 1};
  k_ctf_stack<<<(tot + 255) / 256, 256, 0, st>>>(h, D, C, 0.5 * (defocus_u + defocus_v),
                                                 0.5 * std::fabs(defocus_u - defocus_v), defocus_angle * M_PI / 180.0,
                                                 defocus_offsets, pixel_offsets, lam, cs_mm * 1e7, px,
                                                 std::atan(ac / std::sqrt(1 - ac * ac)), phase * M_PI / 180.0, bfac,
                                                 cm, out);
  TDTM_LAUNCHED();
  return 0;
}
}  // namespace tdtm

// --------------------------------------------------------------------------- plan
using namespace tdtm;

struct tdtm_match_plan {
  int H, W, h, B, D, C, K, G, Wr, wr;
  int Hn, Wn;  // image size of the projection normalisation (the FFT grid H x W, or the
               // unpadded image inside it: tdtm_match_plan_set_image_size)
  int Gs;  // tile-column stride of the intermediate (G rounded up to even: 128-byte lines)
  int num_cu = 256;  // compute units of the plan's device
  const void* occ_kern = nullptr;  // last column-pass kernel whose occupancy was queried
  int occ_per_cu = 1;              // its resident blocks per CU
  // launch-shape options (tdtm_match_plan_set_option); defaults chosen by measurement
  int opt_colpass_columns = 2;  // columns per column-pass block: 2 or 1
  int opt_colpass_pair = 0;     // two correlations per column-pass iteration
  int opt_colpass_chunks = 0;   // k-chunks per column group (0 = automatic)
  int opt_colpass_twiddles = 2; // four-step column transforms: 0 global table, 1 LDS copy, 2 registers
  int opt_colpass_store = 1;    // two-column blocks: 0 direct 8-byte stores, 1 whole rows via LDS
  int opt_colpass_persistent = 1;  // two-column blocks, one correlation per iteration: resident-sized grid
  int opt_rowpass_staging = 3;  // 0: synchronous de-interleaved, 1: async raw, 2: async raw + reformat,
                                // 3: synchronous de-interleaved with batched loads
  int opt_rowpass_twiddles = 1; // 0: twiddles from the global table, 1: LDS copy (staging 0 / 3 only)
  int opt_rowpass_rows = 1;     // rows of a tile row per row-pass block: 4, 2 or 1 (staging 0 / 3)
  int opt_phase_timing = 0;     // record HIP events between the phases of every step
  int opt_fft_engine = 0;       // column transforms: 0 FP32 VALU; MFMA (H = 4096) 1 fp16 split data,
                                // 2 bf16 split data, 3 fp16 rounded once per stage, 4 bf16 rounded once
  // anisotropic magnification of the slices and of the refinement CTF
  // (tdtm_match_plan_set_magnification; off by default)
  SliceMag mag{1.f, 0.f, 0.f, 1.f, 0};
  double mag_d[4] = {1.0, 0.0, 0.0, 1.0};  // the same matrix in double (CTF)
  // per-step events at the phase boundaries (project | rowfwd | colpass | rowpass), read and
  // released by tdtm_match_plan_phase_times
  std::vector<std::array<hipEvent_t, TDTM_NUM_PHASES + 1>> phase_events;
  float2* tw = nullptr;
  float2* Xt = nullptr;
  float* filt = nullptr;
  float2* S = nullptr;
  float2* Tb = nullptr;
  float* P = nullptr;
  float* part = nullptr;
  float2* ae = nullptr;
  float2* Rb = nullptr;
  float2* Gb = nullptr;
  float* ctfbuf = nullptr;
  RefineBest* rpart = nullptr;
  RefineBest* rslots = nullptr;
  int rslots_cap = 0;
  int64_t bytes = 0;
  bool image_set = false, filters_set = false;
  // graph replay of full-batch steps (tdtm_match_run with option "graph"): one captured
  // step reads its first orientation from d_o0 and advances it; the executable is kept
  // for the arguments it was captured with (graph_key) and re-captured when they change
  int opt_graph = 1;
  int* d_o0 = nullptr;
  const int* o0p_launch = nullptr;  // set while a step is being captured
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::array<const void*, 8> graph_key{};
  int graph_O = -1;
  // batched refinement: device state {offset start, slot, particle}, the current pose, and
  // one captured full-batch graph per (defocus count, pixel count, arguments)
  int* d_rstate = nullptr;
  float* d_rpose = nullptr;
  struct RefineGraph {
    int D, C;
    std::array<const void*, 5> key;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<RefineGraph> rgraphs;
};

template <class T>
static int dalloc(T** p, size_t count, int64_t& acc) {
  const size_t b = count * sizeof(T);
  hipError_t e = hipMalloc((void**)p, b);
  if (e != hipSuccess) return fail(std::string("hipMalloc(") + std::to_string(b) + "): " + hipGetErrorString(e));
  acc += (int64_t)b;
  return 0;
}

static void release_phase_events(tdtm_match_plan* p) {
  for (auto& ev : p->phase_events)
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  p->phase_events.clear();
}

static void release_graph(tdtm_match_plan* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
  p->graph_key = {};
  p->graph_O = -1;
}

static void release_refine_graphs(tdtm_match_plan* p) {
  for (auto& g : p->rgraphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  p->rgraphs.clear();
}

static void plan_free(tdtm_match_plan* p) {
  release_phase_events(p);
  release_graph(p);
  release_refine_graphs(p);
  if (p->d_rstate) (void)hipFree(p->d_rstate);
  if (p->d_rpose) (void)hipFree(p->d_rpose);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  if (p->d_o0) (void)hipFree(p->d_o0);
  void* bufs[] = {p->tw, p->Xt, p->filt, p->S, p->Tb, p->P, p->part, p->ae, p->Rb, p->Gb, p->ctfbuf, p->rpart,
                  p->rslots};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
}

// ---------------------------------------------------------- size dispatch helpers
template <int HS>
static void launch_proj(tdtm_match_plan* p, hipStream_t st) {
  constexpr int CP = (HS / V) >= 64 ? 4 : 8;
  constexpr int WR = HS / 2 + 1;
  const int nf = p->D * p->C, nb = p->K / nf;
  const int gx = (WR + CP - 1) / CP;
  // filter chunks: about 8 blocks per CU, each running several filters of one slice
  int nchunk = (int)std::min<long>(nf, std::max<long>(1, (8L * p->num_cu + (long)gx * nb - 1) / ((long)gx * nb)));
  const int fchunk = (nf + nchunk - 1) / nchunk;
  nchunk = (nf + fchunk - 1) / fchunk;
  k_proj_cols<HS, CP><<<dim3(gx, nb, nchunk), CP * (HS / V), 0, st>>>(p->S, p->filt, nb, nf, fchunk, p->tw, p->Tb);
  constexpr int TRW = (HS / 2) / V;
  constexpr int RPB0 = 256 / TRW;
  constexpr int RPB = RPB0 > HS ? HS : RPB0;
  dim3 gr(HS / RPB, p->K);
  k_proj_rows<HS, RPB><<<gr, RPB * TRW, 0, st>>>(p->Tb, p->tw, p->P, p->part);
}

static int dispatch_proj(tdtm_match_plan* p, hipStream_t st) {
  switch (p->h) {
    case 32: launch_proj<32>(p, st); break;
    case 64: launch_proj<64>(p, st); break;
    case 128: launch_proj<128>(p, st); break;
    case 256: launch_proj<256>(p, st); break;
    case 512: launch_proj<512>(p, st); break;
    default: return fail("unsupported template size");
  }
  return 0;
}

template <int WS, int NZJ>
static void launch_rowfwd_n(tdtm_match_plan* p, int K, hipStream_t st) {
  using C = RowFwdCfg<WS>;
  dim3 grid(p->h / C::RR, K);  // h is a power of two >= 32 >= RR
  k_rowfwd<WS, NZJ><<<grid, C::BLOCK, 0, st>>>(p->P, p->ae, p->h, p->G, p->tw, p->Rb);
}

template <int WS>
static void launch_rowfwd(tdtm_match_plan* p, int K, hipStream_t st) {
  constexpr int T = (WS / 2) / V;
  const int nz = (p->h / 2 + T - 1) / T;  // leading non-zero inputs per first-pass butterfly
  if (nz <= 1) launch_rowfwd_n<WS, 1>(p, K, st);
  else if (nz == 2) launch_rowfwd_n<WS, 2>(p, K, st);
  else launch_rowfwd_n<WS, 8>(p, K, st);
}

template <int HH, int NZJ>
static void launch_colpass_n(tdtm_match_plan* p, int K, int ystore, hipStream_t st) {
  constexpr bool FOUR = UseFour<HH>::value;
  const int cols = (FOUR && p->opt_colpass_columns == 1) ? 1 : CW;
  const int nt = (FOUR && p->opt_colpass_pair) ? 2 : 1;
  // one column per block: grid padded to groups of 16 (see k_colpass)
  const int npair8 = ((p->G + 1) / 2 + 7) / 8;  // groups of 8 tile-column pairs
  const int nblk_x = cols == 1 ? 32 * npair8 : 16 * npair8;
  // the four-step LDS twiddle copy when it fits next to the transforms
  constexpr int LFc = FOUR ? FS<(FOUR ? HH : 512)>::floats : LdsSize<HH>::floats;
  constexpr int TWF = FsTwLayout<(FOUR ? HH : 512)>::floats;
  const bool full = ystore == HH;
  const bool ltw = FOUR && p->opt_colpass_twiddles == 1;
  const bool rtw = FOUR && p->opt_colpass_twiddles == 2;
  using Kern = void (*)(const float2*, const float2*, int, int, int, int, int, int, const float2*, float2*, int);
  Kern kern = nullptr;
#define TDTM_COLPASS_T(CWB_, NT_, TWM_, FULL_)                                                       \
  do {                                                                                             \
    if constexpr (NT_ == 1 && FOUR) {                                                              \
      if (p->opt_colpass_store == 1) kern = k_colpass<HH, NZJ, CWB_, NT_, TWM_, FULL_, true>;       \
      else kern = k_colpass<HH, NZJ, CWB_, NT_, TWM_, FULL_, false>;                              \
    } else {                                                                                       \
      kern = k_colpass<HH, NZJ, CWB_, NT_, TWM_, FULL_, false>;                                    \
    }                                                                                              \
  } while (0)
#define TDTM_COLPASS(CWB_, NT_)                                                           \
  do {                                                                                    \
    constexpr int xs_ = NT_ == 1 ? 2 * CWB_ * HH : 0; /* transposed-store area */            \
    constexpr bool fits_ = FOUR && (CWB_ * NT_ * 2 * LFc + TWF + xs_) * 4 <= TDTM_LDS_BYTES;   \
    constexpr bool regs_ = FOUR && NT_ == 1;                                              \
    if (rtw && regs_) {                                                                   \
      if (full) TDTM_COLPASS_T(CWB_, NT_, (regs_ ? 2 : 0), true);                         \
      else TDTM_COLPASS_T(CWB_, NT_, (regs_ ? 2 : 0), false);                             \
    } else if ((ltw || rtw) && fits_) {                                                   \
      if (full) TDTM_COLPASS_T(CWB_, NT_, (fits_ ? 1 : 0), true);                         \
      else TDTM_COLPASS_T(CWB_, NT_, (fits_ ? 1 : 0), false);                             \
    } else {                                                                              \
      if (full) TDTM_COLPASS_T(CWB_, NT_, 0, true);                                       \
      else TDTM_COLPASS_T(CWB_, NT_, 0, false);                                           \
    }                                                                                     \
  } while (0)
  if constexpr (FOUR) {
    if (cols == 1 && nt == 2) TDTM_COLPASS(1, 2);
    else if (cols == 1) TDTM_COLPASS(1, 1);
    else if (nt == 2) TDTM_COLPASS(CW, 2);
    else TDTM_COLPASS(CW, 1);
  } else {
    TDTM_COLPASS(CW, 1);
  }
#undef TDTM_COLPASS
#undef TDTM_COLPASS_T
  const int threads = cols * (HH / V);
  // resident blocks per CU from the runtime's occupancy calculator (cached per kernel:
  // the launch shape of a plan's column pass does not change between steps)
  auto blocks_per_cu = [&]() {
    if (p->occ_kern != (const void*)kern) {
      int n = 1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, threads, 0) != hipSuccess || n < 1) n = 1;
      p->occ_kern = (const void*)kern;
      p->occ_per_cu = n;
    }
    return p->occ_per_cu;
  };
  if (FOUR && nt == 1 && p->opt_colpass_persistent) {
    // persistent schedule: as many blocks as are resident at once, in whole XCD groups of
    // the blocks that share lines (b, b + 8 with two columns per block; b .. b + 24 with
    // one: multiples of 16 / 32 blocks), each group taking an equal share of the
    // ceil(G/2) x K items
    const int per_cu = blocks_per_cu();
    const long gsize = cols == CW ? 2 : 4;  // blocks per group
    const long unit = 8 * gsize;
    long nblk = ((long)per_cu * p->num_cu) / unit * unit;
    if (nblk < unit) nblk = unit;
    const long total = (long)((p->G + 1) / 2) * K;
    const long npairs = nblk / gsize;
    const int persist = (int)((total + npairs - 1) / npairs);
    hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(threads), 0, st, (const float2*)p->Rb,
                       (const float2*)p->Xt, p->h, p->G, p->Gs, K, K, ystore, (const float2*)p->tw, p->Gb, persist);
    return;
  }
  // correlation chunks per column group: the split that minimises (rounds of resident
  // blocks) x (correlations per block + 2 for the block's set-up: image column, twiddle
  // table, first load), i.e. the last round of blocks is as full as possible without
  // many short blocks; resident blocks per CU from the runtime's occupancy calculator
  int nchunk = p->opt_colpass_chunks;
  if (nchunk <= 0) {
    const long slots = (long)blocks_per_cu() * p->num_cu;
    long best = -1;
    for (int n = 1; n <= 16 && n <= K; ++n) {
      int kc = (K + n - 1) / n;
      if (nt == 2 && (kc & 1) && kc < K) ++kc;
      const int nn = (K + kc - 1) / kc;
      const long rounds = ((long)nblk_x * nn + slots - 1) / slots;
      const long cost = rounds * (kc + 2);
      if (best < 0 || cost < best) {
        best = cost;
        nchunk = nn;
      }
    }
  }
  if (nchunk > K) nchunk = K;
  if (nchunk < 1) nchunk = 1;
  int kchunk = (K + nchunk - 1) / nchunk;
  if (nt == 2 && (kchunk & 1) && kchunk < K) ++kchunk;  // keep pairs inside a chunk
  nchunk = (K + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(kern, dim3(nblk_x, nchunk), dim3(threads), 0, st, (const float2*)p->Rb, (const float2*)p->Xt,
                     p->h, p->G, p->Gs, K, kchunk, ystore, (const float2*)p->tw, p->Gb, 0);
}

// Matrix-core column pass (H = 4096, whole columns stored): persistent grid of as many
// 1024-thread blocks as are resident, pairs (b, b + 8) over tile columns 2m / 2m + 1.
template <class ET, bool SPLIT>
static void launch_colpass_mfma(tdtm_match_plan* p, int K, hipStream_t st) {
  const void* kern = (const void*)k_colpass_mfma<ET, SPLIT>;
  if (p->occ_kern != kern) {
    int n = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_colpass_mfma<ET, SPLIT>, MF_BLOCK, 0) != hipSuccess ||
        n < 1)
      n = 1;
    p->occ_kern = kern;
    p->occ_per_cu = n;
  }
  long nblk = ((long)p->occ_per_cu * p->num_cu) / 16 * 16;
  if (nblk < 16) nblk = 16;
  const long total = (long)((p->G + 1) / 2) * K;
  const int persist = (int)((total + nblk / 2 - 1) / (nblk / 2));
  k_colpass_mfma<ET, SPLIT><<<dim3((unsigned)nblk), MF_BLOCK, 0, st>>>(p->Rb, p->Xt, p->h, p->G, p->Gs, K, p->tw,
                                                                      p->Gb, persist);
}

template <int HH>
static void launch_colpass(tdtm_match_plan* p, int K, int ystore, hipStream_t st) {
  if constexpr (HH == MF_N) {
    if (p->opt_fft_engine != 0 && ystore == HH && p->h <= 512) {
      switch (p->opt_fft_engine) {
        case 1: launch_colpass_mfma<_Float16, true>(p, K, st); break;
        case 2: launch_colpass_mfma<__bf16, true>(p, K, st); break;
        case 3: launch_colpass_mfma<_Float16, false>(p, K, st); break;
        default: launch_colpass_mfma<__bf16, false>(p, K, st); break;
      }
      return;
    }
  }
  constexpr int T = HH / V;
  const int nz = (p->h + T - 1) / T;
  if (nz <= 1) launch_colpass_n<HH, 1>(p, K, ystore, st);
  else if (nz == 2) launch_colpass_n<HH, 2>(p, K, ystore, st);
  else launch_colpass_n<HH, 8>(p, K, ystore, st);
}

template <int WS, int MODE, int TWM, int RB>
static void launch_rowpass_rb(tdtm_match_plan* p, int K, int nb, int O, int o0, float* mip, int* idx, float* shi,
                              float* slo, float* qhi, float* qlo, hipStream_t st) {
  k_rowpass<WS, MODE, TWM, RB><<<dim3((p->H / TR) * (TR / RB)), RB * (WS / 2) / V, 0, st>>>(
      p->Gb, p->H, p->G, K, nb, p->D, O, o0, p->tw, mip, idx, shi, slo, qhi, qlo, p->o0p_launch);
}

template <int WS, int MODE, int TWM>
static void launch_rowpass_mt(tdtm_match_plan* p, int K, int nb, int O, int o0, float* mip, int* idx, float* shi,
                              float* slo, float* qhi, float* qlo, hipStream_t st) {
  if constexpr (MODE == 0 || MODE == 3) {
    if (p->opt_rowpass_rows == 1) return launch_rowpass_rb<WS, MODE, TWM, 1>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
    if (p->opt_rowpass_rows == 2 || WS > 4096)
      return launch_rowpass_rb<WS, MODE, TWM, 2>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
  }
  if constexpr (WS <= 4096) launch_rowpass_rb<WS, MODE, TWM, TR>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
}

template <int WS>
static void launch_rowpass_pow2(tdtm_match_plan* p, int K, int nb, int O, int o0, float* mip, int* idx, float* shi,
                                float* slo, float* qhi, float* qlo, hipStream_t st);

template <int WS>
static void launch_rowpass(tdtm_match_plan* p, int K, int nb, int O, int o0, float* mip, int* idx, float* shi,
                           float* slo, float* qhi, float* qlo, hipStream_t st) {
  if constexpr (WS > 4096) {
    // 8192-sample rows: one or two rows per block, synchronous staging (the other shapes
    // do not fit the LDS); rowpass_rows = 4 runs as 2
    if (p->opt_rowpass_twiddles == 1)
      launch_rowpass_mt<WS, 3, 1>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
    else
      launch_rowpass_mt<WS, 3, 0>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
  } else {
    launch_rowpass_pow2<WS>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
  }
}

template <int WS>
static void launch_rowpass_pow2(tdtm_match_plan* p, int K, int nb, int O, int o0, float* mip, int* idx, float* shi,
                                float* slo, float* qhi, float* qlo, hipStream_t st) {
  const bool ltw = p->opt_rowpass_twiddles == 1;
  switch (p->opt_rowpass_staging) {
    case 1: launch_rowpass_mt<WS, 1, 0>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st); break;
    case 2: launch_rowpass_mt<WS, 2, 0>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st); break;
    case 3:
      if (ltw) launch_rowpass_mt<WS, 3, 1>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
      else launch_rowpass_mt<WS, 3, 0>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
      break;
    default:
      if (ltw) launch_rowpass_mt<WS, 0, 1>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
      else launch_rowpass_mt<WS, 0, 0>(p, K, nb, O, o0, mip, idx, shi, slo, qhi, qlo, st);
  }
}

template <int WS>
static void launch_rowpass_refine(tdtm_match_plan* p, int K, int kchunk, int crop_h, int crop_w, const float* mean,
                                  const float* sd, long plane, const int* partp, hipStream_t st) {
  const dim3 grid((crop_h + TR - 1) / TR, (K + kchunk - 1) / kchunk);
  k_rowpass_refine<WS><<<grid, TR * (WS / 2) / V, 0, st>>>(p->Gb, p->H, K, kchunk, crop_h, crop_w, mean, sd, plane,
                                                           partp, p->tw, p->rpart);
}

#define TDTM_SWITCH_POW2(val, FN, ...)             \
  switch (val) {                                   \
    case 64: FN<64>(__VA_ARGS__); break;           \
    case 128: FN<128>(__VA_ARGS__); break;         \
    case 256: FN<256>(__VA_ARGS__); break;         \
    case 512: FN<512>(__VA_ARGS__); break;         \
    case 1024: FN<1024>(__VA_ARGS__); break;       \
    case 2048: FN<2048>(__VA_ARGS__); break;       \
    case 4096: FN<4096>(__VA_ARGS__); break;       \
    default: return fail("unsupported image size"); \
  }

// row lengths of the search's row passes: up to 8192 (padded grids of wide micrographs)
#define TDTM_SWITCH_ROWS(val, FN, ...)             \
  switch (val) {                                   \
    case 64: FN<64>(__VA_ARGS__); break;           \
    case 128: FN<128>(__VA_ARGS__); break;         \
    case 256: FN<256>(__VA_ARGS__); break;         \
    case 512: FN<512>(__VA_ARGS__); break;         \
    case 1024: FN<1024>(__VA_ARGS__); break;       \
    case 2048: FN<2048>(__VA_ARGS__); break;       \
    case 4096: FN<4096>(__VA_ARGS__); break;       \
    case 8192: FN<8192>(__VA_ARGS__); break;       \
    default: return fail("unsupported image size"); \
  }

extern "C" {

int tdtm_abi_version(void) { return 1; }

int tdtm_ctf_filter_stack(int h, int num_defocus, int num_pixel, double defocus_u, double defocus_v,
                          double defocus_angle, const float* defocus_offsets, const float* pixel_offsets,
                          const double* ctf_params, float* out, void* stream) {
  if (!defocus_offsets || !pixel_offsets || !ctf_params || !out) return fail("null argument");
  if (h < 2 || (h & 1) || num_defocus < 1 || num_pixel < 1) return fail("bad sizes");
  return launch_ctf_stack(h, num_defocus, num_pixel, defocus_u, defocus_v, defocus_angle, defocus_offsets,
                          pixel_offsets, ctf_params, nullptr, out, static_cast<hipStream_t>(stream));
}

int tdtm_ctf_filter_stack_mag(int h, int num_defocus, int num_pixel, double defocus_u, double defocus_v,
                              double defocus_angle, const float* defocus_offsets, const float* pixel_offsets,
                              const double* ctf_params, const double* mag_matrix, float* out, void* stream) {
  if (!defocus_offsets || !pixel_offsets || !ctf_params || !out) return fail("null argument");
  if (h < 2 || (h & 1) || num_defocus < 1 || num_pixel < 1) return fail("bad sizes");
  return launch_ctf_stack(h, num_defocus, num_pixel, defocus_u, defocus_v, defocus_angle, defocus_offsets,
                          pixel_offsets, ctf_params, mag_matrix, out, static_cast<hipStream_t>(stream));
}

const char* tdtm_last_error(void) { return last_error().c_str(); }

int tdtm_match_plan_create(int H, int W, int h, int B, int D, int C, tdtm_match_plan_t* out) {
  if (!out) return fail("out_plan is NULL");
  *out = nullptr;
  if (!is_pow2(H) || !is_pow2(W) || H < 64 || W < 64 || H > 4096 || W > 8192)
    return fail("image size must be a power of two in [64, 4096] x [64, 8192], got " + std::to_string(H) + "x" +
                std::to_string(W));
  if (!is_pow2(h) || h < 32 || h > 512) return fail("template size must be a power of two in [32, 512]");
  if (h > H || h > W) return fail("template larger than image");
  if (B < 1 || D < 1 || C < 1) return fail("batch, defocus and pixel counts must be >= 1");
  auto* p = new tdtm_match_plan();
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      p->num_cu = ncu;
  }
  p->H = H;
  p->W = W;
  p->Hn = H;
  p->Wn = W;
  p->h = h;
  p->B = B;
  p->D = D;
  p->C = C;
  p->K = B * D * C;
  p->Wr = W / 2 + 1;
  p->wr = h / 2 + 1;
  p->G = (p->Wr + CW - 1) / CW;
  p->Gs = p->G + (p->G & 1);
  const size_t K = p->K;
  int rc = 0;
  rc |= dalloc(&p->tw, TW_N, p->bytes);
  rc |= rc ? 0 : dalloc(&p->Xt, (size_t)p->G * H * CW, p->bytes);
  rc |= rc ? 0 : dalloc(&p->filt, (size_t)C * D * h * p->wr, p->bytes);
  rc |= rc ? 0 : dalloc(&p->S, (size_t)B * h * p->wr, p->bytes);
  rc |= rc ? 0 : dalloc(&p->Tb, K * h * p->wr, p->bytes);
  rc |= rc ? 0 : dalloc(&p->P, K * h * h, p->bytes);
  rc |= rc ? 0 : dalloc(&p->part, K * h * 3, p->bytes);
  rc |= rc ? 0 : dalloc(&p->ae, K, p->bytes);
  rc |= rc ? 0 : dalloc(&p->Rb, K * p->G * h * CW, p->bytes);
  rc |= rc ? 0 : dalloc(&p->Gb, K * H * p->Gs * CW, p->bytes);
  rc |= rc ? 0 : dalloc(&p->ctfbuf, (size_t)C * D * h * p->wr, p->bytes);
  rc |= rc ? 0 : dalloc(&p->rpart, (size_t)H / TR * REFINE_CHUNKS_MAX, p->bytes);
  rc |= rc ? 0 : dalloc(&p->d_o0, 1, p->bytes);
  rc |= rc ? 0 : dalloc(&p->d_rstate, 3, p->bytes);
  rc |= rc ? 0 : dalloc(&p->d_rpose, 3, p->bytes);
  if (rc) {
    plan_free(p);
    delete p;
    return 1;
  }
  const std::vector<float2> tw = twiddle_table();
  hipError_t e = hipMemcpy(p->tw, tw.data(), TW_N * sizeof(float2), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    plan_free(p);
    delete p;
    return fail(std::string("twiddle upload: ") + hipGetErrorString(e));
  }
  *out = p;
  return 0;
}

int tdtm_match_plan_destroy(tdtm_match_plan_t p) {
  if (!p) return 0;
  plan_free(p);
  delete p;
  return 0;
}

int64_t tdtm_match_plan_bytes(tdtm_match_plan_t p) { return p ? p->bytes : -1; }

int tdtm_match_plan_set_image_size(tdtm_match_plan_t p, int image_h, int image_w) {
  if (!p) return fail("null argument");
  if (image_h < p->h || image_w < p->h || image_h > p->H || image_w > p->W)
    return fail("image size must lie between the template size and the plan's grid " + std::to_string(p->H) + "x" +
                std::to_string(p->W));
  if (image_h != p->Hn || image_w != p->Wn) {
    // the captured steps carry the image size (projection normalisation, refinement crop)
    release_graph(p);
    release_refine_graphs(p);
  }
  p->Hn = image_h;
  p->Wn = image_w;
  return 0;
}

int tdtm_match_plan_set_magnification(tdtm_match_plan_t p, const double* mag_matrix) {
  if (!p) return fail("null argument");
  SliceMag m{1.f, 0.f, 0.f, 1.f, 0};
  if (mag_matrix) {
    for (int i = 0; i < 4; ++i)
      if (!std::isfinite(mag_matrix[i])) return fail("magnification matrix entries must be finite");
    m = SliceMag{(float)mag_matrix[0], (float)mag_matrix[1], (float)mag_matrix[2], (float)mag_matrix[3], 1};
  }
  // the captured steps carry the slice kernel's arguments
  release_graph(p);
  release_refine_graphs(p);
  p->mag = m;
  p->mag_d[0] = mag_matrix ? mag_matrix[0] : 1.0;
  p->mag_d[1] = mag_matrix ? mag_matrix[1] : 0.0;
  p->mag_d[2] = mag_matrix ? mag_matrix[2] : 0.0;
  p->mag_d[3] = mag_matrix ? mag_matrix[3] : 1.0;
  return 0;
}

int tdtm_match_plan_set_option(tdtm_match_plan_t p, const char* key, int value) {
  if (!p || !key) return fail("null argument");
  const std::string k(key);
  if (k == "colpass_columns") {
    if (value != 1 && value != 2) return fail("colpass_columns must be 1 or 2");
    p->opt_colpass_columns = value;
  } else if (k == "colpass_pair") {
    p->opt_colpass_pair = value != 0;
  } else if (k == "colpass_store") {
    if (value < 0 || value > 1) return fail("colpass_store must be 0 or 1");
    p->opt_colpass_store = value;
  } else if (k == "colpass_twiddles") {
    if (value < 0 || value > 2) return fail("colpass_twiddles must be 0, 1 or 2");
    p->opt_colpass_twiddles = value;
  } else if (k == "colpass_persistent") {
    p->opt_colpass_persistent = value != 0;
  } else if (k == "colpass_chunks") {
    if (value < 0) return fail("colpass_chunks must be >= 0");
    p->opt_colpass_chunks = value;
  } else if (k == "rowpass_staging") {
    if (value < 0 || value > 3) return fail("rowpass_staging must be 0, 1, 2 or 3");
    if ((value == 1 || value == 2) && p->opt_rowpass_twiddles)
      return fail("rowpass_staging 1 / 2 leave no LDS for rowpass_twiddles = 1");
    if ((value == 1 || value == 2) && p->opt_rowpass_rows != 4)
      return fail("rowpass_staging 1 / 2 need rowpass_rows = 4");
    // the asynchronous copy loop is written for whole waves: a row-pass block of W / 4
    // threads is below one wave for W < 256 (and rows above 4096 always stage synchronously)
    if ((value == 1 || value == 2) && (p->W < 256 || p->W > 4096))
      return fail("rowpass_staging 1 / 2 need 256 <= W <= 4096 (a row-pass block of at least one wave)");
    p->opt_rowpass_staging = value;
  } else if (k == "rowpass_rows") {
    if (value != 1 && value != 2 && value != 4) return fail("rowpass_rows must be 1, 2 or 4");
    if (value != 4 && (p->opt_rowpass_staging == 1 || p->opt_rowpass_staging == 2))
      return fail("rowpass_rows < 4 needs rowpass_staging 0 or 3");
    p->opt_rowpass_rows = value;
  } else if (k == "rowpass_twiddles") {
    if (value < 0 || value > 1) return fail("rowpass_twiddles must be 0 or 1");
    if (value && (p->opt_rowpass_staging == 1 || p->opt_rowpass_staging == 2))
      return fail("rowpass_twiddles = 1 needs rowpass_staging 0 or 3");
    p->opt_rowpass_twiddles = value;
  } else if (k == "fft_engine") {
    if (value < 0 || value > 4)
      return fail("fft_engine must be 0 (FP32 VALU) or 1..4 (MFMA: fp16 / bf16 with split data, fp16 / bf16 "
                  "rounded once per stage)");
    if (value != 0 && p->H != MF_N) return fail("fft_engine 1..4 need a 4096-row search grid");
    p->opt_fft_engine = value;
    release_graph(p);
  } else if (k == "graph") {
    p->opt_graph = value != 0;
    if (!p->opt_graph) {
      release_graph(p);
      release_refine_graphs(p);
    }
  } else if (k == "phase_timing") {
    p->opt_phase_timing = value != 0;
    if (!p->opt_phase_timing) release_phase_events(p);
  } else {
    return fail("unknown option '" + k + "'");
  }
  // captured steps keep the launch shapes they were recorded with: record again
  if (k != "phase_timing") {
    release_graph(p);
    release_refine_graphs(p);
  }
  return 0;
}

int tdtm_match_set_image(tdtm_match_plan_t p, const void* image_dft, void* stream) {
  if (!p || !image_dft) return fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  const long total = (long)p->G * p->H * CW;
  const float scale = (float)(1.0 / ((double)p->H * (double)p->W));
  k_tile_image<<<(total + 255) / 256, 256, 0, st>>>((const float2*)image_dft, p->H, p->Wr, p->G, scale, p->Xt);
  TDTM_LAUNCHED();
  p->image_set = true;
  return 0;
}

int tdtm_match_set_filters(tdtm_match_plan_t p, const float* ctf, const float* wf, void* stream) {
  if (!p || !ctf || !wf) return fail("null argument");
  const long plane = (long)p->h * p->wr;
  const long total = plane * p->C * p->D;
  k_filters<<<(total + 255) / 256, 256, 0, (hipStream_t)stream>>>(ctf, wf, plane, total, p->filt);
  TDTM_LAUNCHED();
  p->filters_set = true;
  return 0;
}

int tdtm_stats_init(int64_t n, float* mip, int32_t* idx, float* shi, float* slo, float* qhi, float* qlo,
                    void* stream) {
  if (n <= 0) return fail("n must be positive");
  if (!mip || !idx || !shi || !slo || !qhi || !qlo) return fail("null argument");
  k_stats_init<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, mip, idx, shi, slo, qhi, qlo);
  TDTM_LAUNCHED();
  return 0;
}

int tdtm_stats_merge(int64_t n, float* mip, int32_t* idx, float* shi, float* slo, float* qhi, float* qlo,
                     const float* mip_b, const int32_t* idx_b, const float* shi_b, const float* slo_b,
                     const float* qhi_b, const float* qlo_b, void* stream) {
  if (n <= 0) return fail("n must be positive");
  if (!mip || !idx || !shi || !slo || !qhi || !qlo || !mip_b || !idx_b || !shi_b || !slo_b || !qhi_b || !qlo_b)
    return fail("null argument");
  k_stats_merge<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, mip, idx, shi, slo, qhi, qlo, mip_b, idx_b,
                                                                 shi_b, slo_b, qhi_b, qlo_b);
  TDTM_LAUNCHED();
  return 0;
}

}  // extern "C"

// One orientation batch [start, start + count) enqueued on `st`.  With o0p set (graph
// capture) the kernels take the first orientation from device memory instead of `start`.
static int enqueue_step(tdtm_match_plan* p, const void* template_dft, const float* euler, int O, int start, int count,
                        float* mip, int32_t* idx, float* shi, float* slo, float* qhi, float* qlo, hipStream_t st,
                        const int* o0p) {
  const int K = count * p->D * p->C;
  hipEvent_t* ev = nullptr;
  if (p->opt_phase_timing && o0p == nullptr) {
    std::array<hipEvent_t, TDTM_NUM_PHASES + 1> e{};
    for (auto& x : e) TDTM_HIP(hipEventCreate(&x));
    p->phase_events.push_back(e);
    ev = p->phase_events.back().data();
    TDTM_HIP(hipEventRecord(ev[0], st));
  }
  const long ns = (long)count * p->h * p->wr;
  k_slice<<<(ns + 255) / 256, 256, 0, st>>>((const float2*)template_dft, p->h, euler, start, count, nullptr, p->S,
                                             o0p, p->mag);
  TDTM_LAUNCHED();
  // the projection kernels take their batch size from p->K; run them on exactly K
  const int Ksave = p->K;
  p->K = K;
  int rc = dispatch_proj(p, st);
  p->K = Ksave;
  if (rc) return rc;
  TDTM_LAUNCHED();
  k_proj_norm<<<K, 256, 0, st>>>(p->part, p->h, p->Hn, p->Wn, p->ae);
  TDTM_LAUNCHED();
  if (ev) TDTM_HIP(hipEventRecord(ev[1], st));
  TDTM_SWITCH_ROWS(p->W, launch_rowfwd, p, K, st);
  TDTM_LAUNCHED();
  if (ev) TDTM_HIP(hipEventRecord(ev[2], st));
  TDTM_SWITCH_POW2(p->H, launch_colpass, p, K, p->H, st);
  TDTM_LAUNCHED();
  if (ev) TDTM_HIP(hipEventRecord(ev[3], st));
  p->o0p_launch = o0p;
  TDTM_SWITCH_ROWS(p->W, launch_rowpass, p, K, count, O, start, mip, idx, shi, slo, qhi, qlo, st);
  p->o0p_launch = nullptr;
  TDTM_LAUNCHED();
  if (ev) TDTM_HIP(hipEventRecord(ev[4], st));
  return 0;
}

static int step_checks(tdtm_match_plan* p, const void* template_dft, const float* euler, int O, float* mip,
                       int32_t* idx, float* shi, float* slo, float* qhi, float* qlo) {
  if (!p || !template_dft || !euler || !mip || !idx || !shi || !slo || !qhi || !qlo) return fail("null argument");
  if (!p->image_set || !p->filters_set) return fail("tdtm_match_set_image / set_filters not called");
  if ((long)p->C * p->D * O > 2147483647L) return fail("search space exceeds int32 global index");
  if ((long)p->B * p->D * p->C >= 65535) return fail("orientations x defocus x pixel sizes per step must be < 65535");
  return 0;
}

// Capture one full-batch step (first orientation from d_o0, advanced by B at its end) for
// these arguments, unless the kept executable already is for them.
static int ensure_graph(tdtm_match_plan* p, const void* template_dft, const float* euler, int O, float* mip,
                        int32_t* idx, float* shi, float* slo, float* qhi, float* qlo) {
  const std::array<const void*, 8> key{template_dft, euler, mip, idx, shi, slo, qhi, qlo};
  if (p->gexec && key == p->graph_key && O == p->graph_O) return 0;
  release_graph(p);
  if (!p->cap_stream) TDTM_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
  TDTM_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(p, template_dft, euler, O, 0, p->B, mip, idx, shi, slo, qhi, qlo, p->cap_stream, p->d_o0);
  if (!rc) {
    k_advance_start<<<1, 1, 0, p->cap_stream>>>(p->d_o0, p->B);
    if (hipGetLastError() != hipSuccess) rc = fail("kernel launch during capture failed");
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  p->graph = g;
  TDTM_HIP(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
  p->graph_key = key;
  p->graph_O = O;
  return 0;
}

extern "C" {

int tdtm_match_step(tdtm_match_plan_t p, const void* template_dft, const float* euler, int O, int start,
                    int count, float* mip, int32_t* idx, float* shi, float* slo, float* qhi, float* qlo,
                    void* stream) {
  if (step_checks(p, template_dft, euler, O, mip, idx, shi, slo, qhi, qlo)) return 1;
  if (count < 1 || count > p->B) return fail("count must be in [1, max_orientation_batch]");
  if (start < 0 || start + count > O) return fail("orientation range out of bounds");
  return enqueue_step(p, template_dft, euler, O, start, count, mip, idx, shi, slo, qhi, qlo, (hipStream_t)stream,
                      nullptr);
}

int tdtm_match_run(tdtm_match_plan_t p, const void* template_dft, const float* euler, int O, int start, int end,
                   float* mip, int32_t* idx, float* shi, float* slo, float* qhi, float* qlo, void* stream) {
  if (step_checks(p, template_dft, euler, O, mip, idx, shi, slo, qhi, qlo)) return 1;
  if (start < 0 || end > O || start > end) return fail("orientation range out of bounds");
  hipStream_t st = (hipStream_t)stream;
  const int nfull = (end - start) / p->B;
  int o = start;
  if (p->opt_graph && !p->opt_phase_timing && nfull > 0) {
    if (ensure_graph(p, template_dft, euler, O, mip, idx, shi, slo, qhi, qlo)) return 1;
    k_set_start<<<1, 1, 0, st>>>(p->d_o0, start);
    TDTM_LAUNCHED();
    for (int i = 0; i < nfull; ++i) TDTM_HIP(hipGraphLaunch(p->gexec, st));
    o = start + nfull * p->B;
  }
  for (; o < end; o += p->B) {
    const int count = std::min(p->B, end - o);
    if (enqueue_step(p, template_dft, euler, O, o, count, mip, idx, shi, slo, qhi, qlo, st, nullptr)) return 1;
  }
  return 0;
}

int tdtm_match_plan_phase_times(tdtm_match_plan_t p, double* out_ms, int* out_steps) {
  if (!p || !out_ms || !out_steps) return fail("null argument");
  for (int i = 0; i < TDTM_NUM_PHASES; ++i) out_ms[i] = 0.0;
  *out_steps = (int)p->phase_events.size();
  int rc = 0;
  for (auto& e : p->phase_events) {
    if (rc) break;
    hipError_t err = hipEventSynchronize(e[TDTM_NUM_PHASES]);
    if (err != hipSuccess) {
      rc = fail(std::string("hipEventSynchronize: ") + hipGetErrorString(err));
      break;
    }
    for (int i = 0; i < TDTM_NUM_PHASES; ++i) {
      float ms = 0.f;
      err = hipEventElapsedTime(&ms, e[i], e[i + 1]);
      if (err != hipSuccess) {
        rc = fail(std::string("hipEventElapsedTime: ") + hipGetErrorString(err));
        break;
      }
      out_ms[i] += ms;
    }
  }
  release_phase_events(p);
  return rc;
}

int tdtm_match_finalize(int64_t n, int64_t total, const float* mip, const int32_t* idx, const float* shi,
                        const float* slo, const float* qhi, const float* qlo, const float* euler, int O,
                        const float* dvals, int D, float* zs, float* mean, float* sd, float* phi, float* theta,
                        float* psi, float* df, void* stream) {
  if (n <= 0 || total <= 0 || O < 1 || D < 1) return fail("bad sizes");
  k_finalize<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, (double)total, mip, idx, shi, slo, qhi, qlo,
                                                               euler, O, dvals, D, zs, mean, sd, phi, theta, psi,
                                                               df);
  TDTM_LAUNCHED();
  return 0;
}

}  // extern "C"

// Launches of one refinement batch of `count` offsets from `start` (K = count D C
// correlations, the plan's D and C): slice (offsets composed with `pose`), projection,
// padded forward FFT, column pass over the valid rows, row pass with the z epilogue against
// mean / sd, and the reduction into result slot `bi`.  With `state` (graph capture) the
// first offset, the slot and the particle index come from state[0..2] (mean / sd are then
// the bases of per-particle maps `plane` floats apart).
static int refine_batch(tdtm_match_plan* p, const void* template_dft, const float* offsets, int start, int count,
                        const float* pose, const float* mean, const float* sd, long plane, int bi, const int* state,
                        hipStream_t st) {
  const int crop_h = p->Hn - p->h + 1, crop_w = p->Wn - p->h + 1;
  const int ystore = ((crop_h + TR - 1) / TR) * TR;
  const int nparts = (crop_h + TR - 1) / TR;
  const int K = count * p->D * p->C;
  const long ns = (long)count * p->h * p->wr;
  k_slice<<<(ns + 255) / 256, 256, 0, st>>>((const float2*)template_dft, p->h, offsets, start, count, pose, p->S,
                                             state, p->mag);
  TDTM_LAUNCHED();
  const int Ksave = p->K;
  p->K = K;
  int rc = dispatch_proj(p, st);
  p->K = Ksave;
  if (rc) return rc;
  TDTM_LAUNCHED();
  k_proj_norm<<<K, 256, 0, st>>>(p->part, p->h, p->Hn, p->Wn, p->ae);
  TDTM_LAUNCHED();
  TDTM_SWITCH_POW2(p->W, launch_rowfwd, p, K, st);
  TDTM_LAUNCHED();
  TDTM_SWITCH_POW2(p->H, launch_colpass, p, K, ystore, st);
  TDTM_LAUNCHED();
  // correlation chunks: about 8 blocks per CU over the tile rows of the crop
  int nchunk = std::max(1, std::min({(8 * p->num_cu + nparts - 1) / nparts, K, REFINE_CHUNKS_MAX}));
  const int kchunk = (K + nchunk - 1) / nchunk;
  nchunk = (K + kchunk - 1) / kchunk;
  TDTM_SWITCH_POW2(p->W, launch_rowpass_refine, p, K, kchunk, crop_h, crop_w, mean, sd, plane,
                   state ? state + 2 : nullptr, st);
  TDTM_LAUNCHED();
  k_refine_reduce<<<1, 256, 0, st>>>(p->rpart, nparts * nchunk, p->rslots, bi, state ? state + 1 : nullptr);
  TDTM_LAUNCHED();
  return 0;
}

static int ensure_rslots(tdtm_match_plan* p, int nbatch) {
  if (nbatch <= p->rslots_cap) return 0;
  release_refine_graphs(p);  // captured reductions point at the old slots
  if (p->rslots) TDTM_HIP(hipFree(p->rslots));
  p->rslots = nullptr;
  p->rslots_cap = 0;
  if (dalloc(&p->rslots, (size_t)nbatch, p->bytes)) return 1;
  p->rslots_cap = nbatch;
  return 0;
}

// Captured full batch (B offsets) for the plan's current D and C, or the kept one.
static int refine_graph(tdtm_match_plan* p, const void* template_dft, const float* offsets, const float* mean,
                        const float* sd, long plane, hipGraphExec_t* out) {
  const std::array<const void*, 5> key{template_dft, offsets, mean, sd, (const void*)p->rslots};
  for (auto& g : p->rgraphs)
    if (g.D == p->D && g.C == p->C && g.key == key) {
      *out = g.exec;
      return 0;
    }
  if (p->rgraphs.size() >= 16) release_refine_graphs(p);
  if (!p->cap_stream) TDTM_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
  TDTM_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = refine_batch(p, template_dft, offsets, 0, p->B, p->d_rpose, mean, sd, plane, 0, p->d_rstate,
                        p->cap_stream);
  if (!rc) {
    k_refine_advance<<<1, 1, 0, p->cap_stream>>>(p->d_rstate, p->B);
    if (hipGetLastError() != hipSuccess) rc = fail("kernel launch during capture failed");
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  p->rgraphs.push_back({p->D, p->C, key, g, ex});
  *out = ex;
  return 0;
}

// CTF stack x projective filter of one particle into the plan's filters (D, C of the plan).
static int refine_filters(tdtm_match_plan* p, double du, double dv, double da, const float* dfo, const float* pxo,
                          const double* ctf_params, const float* projective_filter, hipStream_t st) {
  if (launch_ctf_stack(p->h, p->D, p->C, du, dv, da, dfo, pxo, ctf_params, p->mag.on ? p->mag_d : nullptr,
                       p->ctfbuf, st))
    return 1;
  const long plane = (long)p->h * p->wr;
  const long tot = plane * p->D * p->C;
  k_filters<<<(tot + 255) / 256, 256, 0, st>>>(p->ctfbuf, projective_filter, plane, tot, p->filt);
  TDTM_LAUNCHED();
  return 0;
}

extern "C" {

int tdtm_refine_particle(tdtm_match_plan_t p, const void* template_dft, const void* particle_dft,
                         const float* euler_pose, const float* offsets, int n_off, const float* defocus_offsets,
                         const float* pixel_offsets, double defocus_u, double defocus_v, double defocus_angle,
                         const double* ctf_params, const float* projective_filter, const float* corr_mean,
                         const float* corr_std, float* out, void* stream) {
  if (!p || !template_dft || !particle_dft || !euler_pose || !offsets || !defocus_offsets || !pixel_offsets ||
      !ctf_params || !projective_filter || !corr_mean || !corr_std || !out)
    return fail("null argument");
  if (n_off < 1) return fail("no orientation offsets");
  if ((long)p->B * p->D * p->C >= 65535) return fail("batch x defocus x pixel sizes must be < 65535");
  hipStream_t st = (hipStream_t)stream;
  if (p->W > 4096) return fail("refinement crops are at most 4096 wide");
  const int crop_w = p->Wn - p->h + 1;
  const int nbatch = (n_off + p->B - 1) / p->B;
  if (ensure_rslots(p, nbatch)) return 1;
  // image tiles (match plan semantics: scale 1/(H W))
  if (tdtm_match_set_image(p, particle_dft, stream)) return 1;
  if (refine_filters(p, defocus_u, defocus_v, defocus_angle, defocus_offsets, pixel_offsets, ctf_params,
                     projective_filter, st))
    return 1;
  for (int bi = 0; bi < nbatch; ++bi) {
    const int start = bi * p->B;
    if (refine_batch(p, template_dft, offsets, start, std::min(p->B, n_off - start), euler_pose, corr_mean, corr_std,
                     0, bi, nullptr, st))
      return 1;
  }
  k_refine_select<<<1, 64, 0, st>>>(p->rslots, nbatch, p->B, n_off, p->D, offsets, euler_pose, defocus_offsets,
                                     pixel_offsets, crop_w, out);
  TDTM_LAUNCHED();
  return 0;
}

int tdtm_refine_batch(tdtm_match_plan_t p, const void* template_dft, const void* particle_dft, int n_particles,
                      int64_t particle_stride, const float* euler_poses, const float* offsets, const int* off_start,
                      const int* off_count, const float* defocus_offsets, const int* d_start, const int* d_count,
                      const float* pixel_offsets, const int* c_start, const int* c_count, const double* defocus_u,
                      const double* defocus_v, const double* defocus_angle, const double* ctf_params,
                      const float* projective_filters, const float* corr_mean, const float* corr_std, float* out,
                      void* stream) {
  if (!p || !template_dft || !particle_dft || !euler_poses || !offsets || !off_start || !off_count ||
      !defocus_offsets || !d_start || !d_count || !pixel_offsets || !c_start || !c_count || !defocus_u ||
      !defocus_v || !defocus_angle || !ctf_params || !projective_filters || !corr_mean || !corr_std || !out)
    return fail("null argument");
  if (n_particles < 1) return fail("no particles");
  if (p->W > 4096) return fail("refinement crops are at most 4096 wide");
  if (particle_stride < (int64_t)p->H * p->Wr) return fail("particle stride below the spectrum size");
  const int Dmax = p->D, Cmax = p->C;
  int maxbatch = 0;
  for (int i = 0; i < n_particles; ++i) {
    if (off_count[i] < 1 || off_start[i] < 0) return fail("particle " + std::to_string(i) + ": no orientation offsets");
    if (d_count[i] < 1 || d_count[i] > Dmax || d_start[i] < 0)
      return fail("particle " + std::to_string(i) + ": defocus count outside [1, plan num_defocus]");
    if (c_count[i] < 1 || c_count[i] > Cmax || c_start[i] < 0)
      return fail("particle " + std::to_string(i) + ": pixel-size count outside [1, plan num_pixel]");
    maxbatch = std::max(maxbatch, (off_count[i] + p->B - 1) / p->B);
  }
  if ((long)p->B * Dmax * Cmax >= 65535) return fail("batch x defocus x pixel sizes must be < 65535");
  hipStream_t st = (hipStream_t)stream;
  if (ensure_rslots(p, maxbatch)) return 1;
  const int crop_h = p->Hn - p->h + 1, crop_w = p->Wn - p->h + 1;
  const long plane = (long)crop_h * crop_w;
  const long fplane = (long)p->h * p->wr;
  const float2* X = static_cast<const float2*>(particle_dft);
  int rc = 0;
  for (int i = 0; i < n_particles && !rc; ++i) {
    p->D = d_count[i];
    p->C = c_count[i];
    const int n_off = off_count[i];
    const float* offs = offsets + 3L * off_start[i];
    const float* dfo = defocus_offsets + d_start[i];
    const float* pxo = pixel_offsets + c_start[i];
    const int nbatch = (n_off + p->B - 1) / p->B, nfull = n_off / p->B;
    rc = tdtm_match_set_image(p, X + (long)i * particle_stride, stream);
    if (!rc)
      rc = refine_filters(p, defocus_u[i], defocus_v[i], defocus_angle[i], dfo, pxo, ctf_params,
                          projective_filters + (long)i * fplane, st);
    if (rc) break;
    // device state: first offset, slot 0, particle i; pose copied to the plan's slot
    k_refine_state<<<1, 1, 0, st>>>(p->d_rstate, off_start[i], 0, i, euler_poses + 3L * i, p->d_rpose);
    if (hipGetLastError() != hipSuccess) {
      rc = fail("kernel launch failed");
      break;
    }
    int done = 0;
    if (p->opt_graph && nfull > 0) {
      hipGraphExec_t ex = nullptr;
      rc = refine_graph(p, template_dft, offsets, corr_mean, corr_std, plane, &ex);
      for (int b = 0; b < nfull && !rc; ++b) {
        const hipError_t e = hipGraphLaunch(ex, st);
        if (e != hipSuccess) rc = fail(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
      }
      done = nfull;
    }
    for (int bi = done; bi < nbatch && !rc; ++bi) {
      const int start = bi * p->B;
      rc = refine_batch(p, template_dft, offs, start, std::min(p->B, n_off - start), p->d_rpose,
                        corr_mean + (long)i * plane, corr_std + (long)i * plane, 0, bi, nullptr, st);
    }
    if (rc) break;
    k_refine_select<<<1, 64, 0, st>>>(p->rslots, nbatch, p->B, n_off, p->D, offs, p->d_rpose, dfo, pxo, crop_w,
                                       out + 14L * i);
    if (hipGetLastError() != hipSuccess) rc = fail("kernel launch failed");
  }
  p->D = Dmax;
  p->C = Cmax;
  return rc;
}

}  // extern "C"
