// Orientation-proposal network on gfx950: training and inference kernels and C ABI.
//
// A small CNN scores coarse orientation bins (plus one "empty tile" class) for a 64 x 64
// image tile (a 256 x 256 micrograph tile averaged 4 x 4):
//   conv 3x3 (1 -> 16) + ReLU, avg-pool 2   -> 32 x 32 x 16
//   conv 3x3 (16 -> 32) + ReLU, avg-pool 2  -> 16 x 16 x 32
//   conv 3x3 (32 -> 64) + ReLU, avg-pool 2  ->  8 x  8 x 64
//   conv 3x3 (64 -> 64) + ReLU, avg-pool 2  ->  4 x  4 x 64 = 1024
//   fc 1024 -> 256 + ReLU, fc 256 -> L = bins + 1
// Activations are NHWC float32.  Every matrix product -- the convolutions as im2col x
// weights, the fully connected layers, and all backward products -- runs on
// v_mfma_f32_16x16x32_bf16: operands rounded to bf16 (weights are kept as float32 master
// copies for the optimiser and re-rounded after every update), float32 accumulation.
// Loss: softmax cross-entropy against a target distribution.  Optimiser: Adam.
// Reductions (bias gradients, col2im) run in a fixed order: results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "tdtm.h"

namespace tdtm {
namespace {

typedef __bf16 pn_b8 __attribute__((ext_vector_type(8)));
typedef float pn_f4 __attribute__((ext_vector_type(4)));

constexpr int PN_IN = 64;         // input tile edge
constexpr int PN_LAYERS = 4;      // convolution layers
constexpr int PN_CH[PN_LAYERS + 1] = {1, 16, 32, 64, 64};
constexpr int PN_FEAT = 4 * 4 * 64;
constexpr int PN_HID = 256;

// C[m][n] (beta = 0: =, 1: +=) sum_k A(m, k) B(k, n) with A(m, k) = A[m lda + k] (TA:
// A[k lda + m]) and B(k, n) = B[k ldb + n] (TB: B[n ldb + k]), bf16 operands, float32
// result.  One wave per 16 x 16 tile of C, four waves per block along n; lane l = j + 16 g
// holds row j of the A tile and column j of the B tile for k = k0 + 8 g .. + 7 and rows
// 4 g .. 4 g + 3 of column j of the result.
// Split K (gridDim.z > 1): block z sums k in [z kchunk, (z + 1) kchunk) into its own
// (M x N) slice C + z M N (ldc = N, beta 0); k_sum_splits adds the slices in order.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                   const __bf16* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, int beta, int kchunk) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, j = l & 15, g = l >> 4;
  const int m0 = blockIdx.y * 16, n0 = (blockIdx.x * 4 + wave) * 16;
  if (n0 >= N) return;  // whole wave leaves together
  const int row = m0 + j, col = n0 + j;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  C += (long)blockIdx.z * M * N;
  pn_f4 acc = {0.f, 0.f, 0.f, 0.f};
  const __bf16 zero = (__bf16)0.f;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    pn_b8 a, b;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = k0 + 8 * g + s;
      const bool ia = row < M && k < ke, ib = col < N && k < ke;
      a[s] = ia ? (TA ? A[(long)k * lda + row] : A[(long)row * lda + k]) : zero;
      b[s] = ib ? (TB ? B[(long)col * ldb + k] : B[(long)k * ldb + col]) : zero;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + 4 * g + r;
    if (m < M && col < N) {
      float* c = C + (long)m * ldc + col;
      *c = beta ? *c + acc[r] : acc[r];
    }
  }
}

// 3x3 patches (zero padding 1) of an NHWC image as rows of a bf16 matrix:
// cols[(n, y, x)][(ky, kx, c)] = X[n][y + ky - 1][x + kx - 1][c].
__global__ void k_im2col(const float* __restrict__ X, int Nb, int H, int W, int Cc, __bf16* __restrict__ cols) {
  const long K = 9L * Cc;
  const long total = (long)Nb * H * W * K;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int kk = (int)(e % K);
  const long px = e / K;
  const int c = kk % Cc, tap = kk / Cc, kx = tap % 3, ky = tap / 3;
  const int x = (int)(px % W), y = (int)((px / W) % H), n = (int)(px / ((long)W * H));
  const int yy = y + ky - 1, xx = x + kx - 1;
  float v = 0.f;
  if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = X[(((long)n * H + yy) * W + xx) * Cc + c];
  cols[e] = (__bf16)v;
}

// Gradient of the image from the gradient of its patch matrix (the transpose of k_im2col):
// dX[n][y][x][c] = sum over the 9 taps of dcols at the patch rows that read (y, x).
__global__ void k_col2im(const float* __restrict__ dcols, int Nb, int H, int W, int Cc, float* __restrict__ dX) {
  const long total = (long)Nb * H * W * Cc;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int c = (int)(e % Cc);
  const long px = e / Cc;
  const int x = (int)(px % W), y = (int)((px / W) % H), n = (int)(px / ((long)W * H));
  const long K = 9L * Cc;
  float s = 0.f;
  for (int ky = 0; ky < 3; ++ky)
    for (int kx = 0; kx < 3; ++kx) {
      const int py = y - ky + 1, pxx = x - kx + 1;  // patch centre that reads (y, x) with tap (ky, kx)
      if (py >= 0 && py < H && pxx >= 0 && pxx < W)
        s += dcols[(((long)n * H + py) * W + pxx) * K + (ky * 3 + kx) * Cc + c];
    }
  dX[e] = s;
}

__global__ void k_bias_relu(float* __restrict__ Y, const float* __restrict__ bias, long rows, int Cc) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= rows * Cc) return;
  Y[e] = fmaxf(Y[e] + bias[e % Cc], 0.f);
}

__global__ void k_bias(float* __restrict__ Y, const float* __restrict__ bias, long rows, int Cc) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= rows * Cc) return;
  Y[e] += bias[e % Cc];
}

// ReLU gradient through the stored post-activation output (y > 0), in place on dY.
__global__ void k_relu_grad(float* __restrict__ dY, const float* __restrict__ Y, long n) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (!(Y[e] > 0.f)) dY[e] = 0.f;
}

__global__ void k_avgpool(const float* __restrict__ X, int Nb, int H, int W, int Cc, float* __restrict__ Y) {
  const int Ho = H / 2, Wo = W / 2;
  const long total = (long)Nb * Ho * Wo * Cc;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int c = (int)(e % Cc);
  const long px = e / Cc;
  const int x = (int)(px % Wo), y = (int)((px / Wo) % Ho), n = (int)(px / ((long)Wo * Ho));
  const float* b = X + (((long)n * H + 2 * y) * W + 2 * x) * Cc + c;
  Y[e] = 0.25f * (b[0] + b[Cc] + b[(long)W * Cc] + b[(long)W * Cc + Cc]);
}

__global__ void k_avgpool_grad(const float* __restrict__ dY, int Nb, int H, int W, int Cc, float* __restrict__ dX) {
  const long total = (long)Nb * H * W * Cc;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int c = (int)(e % Cc);
  const long px = e / Cc;
  const int x = (int)(px % W), y = (int)((px / W) % H), n = (int)(px / ((long)W * H));
  dX[e] = 0.25f * dY[(((long)n * (H / 2) + y / 2) * (W / 2) + x / 2) * Cc + c];
}

__global__ void k_zero(float* __restrict__ x, long n) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) x[e] = 0.f;
}

__global__ void k_to_bf16(const float* __restrict__ x, long n, __bf16* __restrict__ y) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) y[e] = (__bf16)x[e];
}

// out[m ldc + n] (beta = 0: =, 1: +=) sum over z of the split-K slices part[z][m][n], in z order.
__global__ void k_sum_splits(const float* __restrict__ part, int splits, int M, int N, float* __restrict__ C,
                             int ldc, int beta) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)M * N) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(long)z * M * N + e];
  float* c = C + (e / N) * ldc + e % N;
  *c = beta ? *c + s : s;
}

// Column sums of a (rows x Cc) matrix (bias gradients), Cc dividing 256: block b takes rows
// [b rchunk, (b + 1) rchunk); thread t adds column t % Cc over rows t / Cc, + 256 / Cc, ...
// in order (consecutive threads read consecutive addresses), then the 256 / Cc phases are
// added in a fixed tree; part[b][c] holds the block's sums.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ G, long rows, int Cc, long rchunk,
                                                     float* __restrict__ part) {
  __shared__ float red[256];
  const int t = threadIdx.x, P = 256 / Cc, c = t % Cc, ph = t / Cc;
  const long r0 = blockIdx.x * rchunk, r1 = min(rows, r0 + rchunk);
  float s = 0.f;
  for (long r = r0 + ph; r < r1; r += P) s += G[r * Cc + c];
  red[t] = s;
  __syncthreads();
  for (int w = P / 2; w > 0; w >>= 1) {
    if (ph < w) red[t] += red[t + w * Cc];
    __syncthreads();
  }
  if (ph == 0) part[(long)blockIdx.x * Cc + c] = red[t];
}

// out[c] = sum over b of part[b][c], in b order.
__global__ void k_colsum_final(const float* __restrict__ part, int nb, int Cc, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cc) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(long)b * Cc + c];
  out[c] = s;
}

// Column sums for any Cc: one block per column, thread t sums rows t, t + 256, ... in
// order, then a fixed tree over the threads (used for the classifier's L columns).
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ G, long rows, int Cc,
                                                float* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (long r = threadIdx.x; r < rows; r += 256) s += G[r * Cc + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// Softmax cross-entropy against target distributions T (rows sum to 1): loss[n] and
// dlogits = (softmax - T) / Nb.  One block per sample.
__global__ __launch_bounds__(256) void k_softmax_xent(const float* __restrict__ Z, const float* __restrict__ T, int L,
                                                      int Nb, float* __restrict__ dZ, float* __restrict__ loss) {
  __shared__ float red[256];
  const int n = blockIdx.x;
  const float* z = Z + (long)n * L;
  float m = -INFINITY;
  for (int i = threadIdx.x; i < L; i += 256) m = fmaxf(m, z[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  m = red[0];
  __syncthreads();
  float s = 0.f;
  for (int i = threadIdx.x; i < L; i += 256) s += expf(z[i] - m);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const float lse = m + logf(red[0]);
  __syncthreads();
  float lo = 0.f;
  for (int i = threadIdx.x; i < L; i += 256) {
    const float t = T[(long)n * L + i];
    const float p = expf(z[i] - lse);
    dZ[(long)n * L + i] = (p - t) / (float)Nb;
    if (t > 0.f) lo -= t * (z[i] - lse);
  }
  red[threadIdx.x] = lo;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[n] = red[0];
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ gr, long n, float lr, float b1, float b2, float eps, float c1,
                       float c2, __bf16* __restrict__ wb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float g = gr[e];
  const float mm = b1 * m[e] + (1.f - b1) * g;
  const float vv = b2 * v[e] + (1.f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  const float ww = w[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[e] = ww;
  wb[e] = (__bf16)ww;
}

unsigned nblk(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace tdtm

using namespace tdtm;

// Parameter layout (one float32 vector): for each conv layer l: W_l (Cout, 9 Cin) as
// [cout][(ky, kx, cin)], b_l (Cout); then W_fc1 (256, 1024) with inputs in NHWC order
// (y, x, c) of the 4 x 4 x 64 features, b_fc1 (256), W_fc2 (L, 256), b_fc2 (L).
struct tdtm_proposal_net {
  int L = 0, maxN = 0;
  long nparam = 0;
  long woff[PN_LAYERS + 2] = {}, boff[PN_LAYERS + 2] = {};
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  __bf16* wb = nullptr;
  // activations (maxN samples): a[0] input, conv outputs c[l] (post-ReLU), pooled p[l]
  float* act[2 * PN_LAYERS + 1] = {};
  float *h1 = nullptr, *z = nullptr, *dz = nullptr, *loss = nullptr;
  __bf16* cols = nullptr;  // largest patch matrix, bf16
  __bf16* ab = nullptr;    // bf16 copy of a GEMM's activation / gradient operand
  float* dbuf[3] = {};     // gradient scratch (largest activation)
  float* dcols = nullptr;
  float* ws = nullptr;    // split-K slices of the weight-gradient products
  float* csum = nullptr;  // per-block column sums (bias gradients)
  long cols_max = 0, act_max = 0;
  int steps = 0;
};

namespace {

int pn_alloc(void** p, size_t bytes) {
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

void pn_free(tdtm_proposal_net* n) {
  void* bufs[] = {n->w, n->g, n->m, n->v, n->wb, n->h1, n->z, n->dz, n->loss, n->cols, n->ab, n->dcols,
                  n->dbuf[0], n->dbuf[1], n->dbuf[2], n->ws, n->csum};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (float* a : n->act)
    if (a) (void)hipFree(a);
}

int edge(int l) { return PN_IN >> l; }  // spatial edge of the input of conv layer l

// Split K for products with few output tiles and a long K (the weight gradients, K = batch
// x pixels): aim at PN_SPLIT_TARGET blocks, at least PN_SPLIT_MIN_K per split.
constexpr long PN_SPLIT_TARGET = 512, PN_SPLIT_MIN_K = 1024, PN_SPLIT_MAX = 128;
constexpr long PN_COLSUM_BLOCKS = 128, PN_COLSUM_ROWS = 1024;

// Upper bound on the splits of a product: non-decreasing in K.
long split_bound(int M, int N, int K) {
  const long base = (long)((N + 63) / 64) * ((M + 15) / 16);
  const long s = std::min((PN_SPLIT_TARGET + base - 1) / base, std::max(1L, (long)K / PN_SPLIT_MIN_K));
  return std::max(1L, std::min(s, PN_SPLIT_MAX));
}

// (splits, k per split) of a product, splits <= split_bound; splits = 1 runs the whole K
// in one block column.
void gemm_splits(int M, int N, int K, int* splits, int* kchunk) {
  const long s = split_bound(M, N, K);
  const long kc = ((K + s - 1) / s + 31) / 32 * 32;
  *kchunk = (int)kc;
  *splits = (int)((K + kc - 1) / kc);
}

// Floats of split-K slices any product of this shape with K' <= K needs.
long splits_floats(int M, int N, int K) {
  const long s = split_bound(M, N, K);
  return s > 1 ? s * M * N : 0;
}

// C = / += A B; with ws (room for splits_floats(M, N, K) floats) the product is split
// over K and the slices summed in order.
template <bool TA, bool TB>
int gemm(int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb, float* C, int ldc, int beta,
         hipStream_t st, float* ws = nullptr) {
  int splits = 1, kchunk = K;
  if (ws) gemm_splits(M, N, K, &splits, &kchunk);
  if (splits == 1) kchunk = K;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 15) / 16), (unsigned)splits);
  if (splits > 1) {
    k_gemm_bf16<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, ws, N, 0, kchunk);
    TDTM_LAUNCHED();
    k_sum_splits<<<nblk((long)M * N), 256, 0, st>>>(ws, splits, M, N, C, ldc, beta);
  } else {
    k_gemm_bf16<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta, kchunk);
  }
  TDTM_LAUNCHED();
  return 0;
}

long colsum_blocks(long rows) { return std::min(PN_COLSUM_BLOCKS, (rows + PN_COLSUM_ROWS - 1) / PN_COLSUM_ROWS); }

// out[c] = sum over rows of G[r][c] (fixed order: deterministic).
int colsum(tdtm_proposal_net* n, const float* G, long rows, int Cc, float* out, hipStream_t st) {
  if (Cc <= 256 && 256 % Cc == 0 && rows > PN_COLSUM_ROWS) {
    const long nb = colsum_blocks(rows), rchunk = (rows + nb - 1) / nb;
    k_colsum_part<<<(unsigned)nb, 256, 0, st>>>(G, rows, Cc, rchunk, n->csum);
    TDTM_LAUNCHED();
    k_colsum_final<<<nblk(Cc), 256, 0, st>>>(n->csum, (int)nb, Cc, out);
  } else {
    k_colsum<<<(unsigned)Cc, 256, 0, st>>>(G, rows, Cc, out);
  }
  TDTM_LAUNCHED();
  return 0;
}

int to_bf16(const float* x, long n, __bf16* y, hipStream_t st) {
  k_to_bf16<<<nblk(n), 256, 0, st>>>(x, n, y);
  TDTM_LAUNCHED();
  return 0;
}

int pn_forward(tdtm_proposal_net* n, int Nb, hipStream_t st) {
  for (int l = 0; l < PN_LAYERS; ++l) {
    const int E = edge(l), Ci = PN_CH[l], Co = PN_CH[l + 1];
    const long rows = (long)Nb * E * E;
    const float* X = n->act[2 * l];
    float* Y = n->act[2 * l + 1];
    k_im2col<<<nblk(rows * 9 * Ci), 256, 0, st>>>(X, Nb, E, E, Ci, n->cols);
    TDTM_LAUNCHED();
    // Y (rows x Co) = cols (rows x 9 Ci) W^T, W (Co x 9 Ci)
    if (gemm<false, true>((int)rows, Co, 9 * Ci, n->cols, 9 * Ci, n->wb + n->woff[l], 9 * Ci, Y, Co, 0, st)) return 1;
    k_bias_relu<<<nblk(rows * Co), 256, 0, st>>>(Y, n->w + n->boff[l], rows, Co);
    TDTM_LAUNCHED();
    k_avgpool<<<nblk(rows * Co / 4), 256, 0, st>>>(Y, Nb, E, E, Co, n->act[2 * l + 2]);
    TDTM_LAUNCHED();
  }
  const float* F = n->act[2 * PN_LAYERS];  // (Nb, 1024)
  if (to_bf16(F, (long)Nb * PN_FEAT, n->ab, st)) return 1;
  if (gemm<false, true>(Nb, PN_HID, PN_FEAT, n->ab, PN_FEAT, n->wb + n->woff[PN_LAYERS], PN_FEAT, n->h1, PN_HID, 0,
                        st))
    return 1;
  k_bias_relu<<<nblk((long)Nb * PN_HID), 256, 0, st>>>(n->h1, n->w + n->boff[PN_LAYERS], Nb, PN_HID);
  TDTM_LAUNCHED();
  if (to_bf16(n->h1, (long)Nb * PN_HID, n->ab, st)) return 1;
  if (gemm<false, true>(Nb, n->L, PN_HID, n->ab, PN_HID, n->wb + n->woff[PN_LAYERS + 1], PN_HID, n->z, n->L, 0, st))
    return 1;
  k_bias<<<nblk((long)Nb * n->L), 256, 0, st>>>(n->z, n->w + n->boff[PN_LAYERS + 1], Nb, n->L);
  TDTM_LAUNCHED();
  return 0;
}

// Gradients of every parameter into n->g (overwritten) from dz.
int pn_backward(tdtm_proposal_net* n, int Nb, hipStream_t st) {
  // fc2: dW2 (L x 256) = dz^T h1, db2 = colsum dz, dh1 = dz W2 (Nb x 256)
  float* dh1 = n->dbuf[0];
  if (to_bf16(n->dz, (long)Nb * n->L, n->ab, st)) return 1;
  __bf16* hb = reinterpret_cast<__bf16*>(n->dcols);  // bf16 copy of h1 (dcols is free here)
  if (to_bf16(n->h1, (long)Nb * PN_HID, hb, st)) return 1;
  if (gemm<true, false>(n->L, PN_HID, Nb, n->ab, n->L, hb, PN_HID, n->g + n->woff[PN_LAYERS + 1], PN_HID, 0, st,
                        n->ws))
    return 1;
  if (colsum(n, n->dz, Nb, n->L, n->g + n->boff[PN_LAYERS + 1], st)) return 1;
  if (gemm<false, false>(Nb, PN_HID, n->L, n->ab, n->L, n->wb + n->woff[PN_LAYERS + 1], PN_HID, dh1, PN_HID, 0, st))
    return 1;
  k_relu_grad<<<nblk((long)Nb * PN_HID), 256, 0, st>>>(dh1, n->h1, (long)Nb * PN_HID);
  TDTM_LAUNCHED();
  // fc1: dW1 (256 x 1024) = dh1^T F, db1, dF = dh1 W1 (Nb x 1024)
  float* dF = n->dbuf[1];
  if (to_bf16(dh1, (long)Nb * PN_HID, n->ab, st)) return 1;
  if (to_bf16(n->act[2 * PN_LAYERS], (long)Nb * PN_FEAT, hb, st)) return 1;
  if (gemm<true, false>(PN_HID, PN_FEAT, Nb, n->ab, PN_HID, hb, PN_FEAT, n->g + n->woff[PN_LAYERS], PN_FEAT, 0, st,
                        n->ws))
    return 1;
  if (colsum(n, dh1, Nb, PN_HID, n->g + n->boff[PN_LAYERS], st)) return 1;
  if (gemm<false, false>(Nb, PN_FEAT, PN_HID, n->ab, PN_HID, n->wb + n->woff[PN_LAYERS], PN_FEAT, dF, PN_FEAT, 0, st))
    return 1;
  // conv layers, last to first; dP = gradient of the pooled output of layer l
  float* dP = dF;
  for (int l = PN_LAYERS - 1; l >= 0; --l) {
    const int E = edge(l), Ci = PN_CH[l], Co = PN_CH[l + 1];
    const long rows = (long)Nb * E * E;
    float* dY = (dP == n->dbuf[2]) ? n->dbuf[0] : n->dbuf[2];
    k_avgpool_grad<<<nblk(rows * Co), 256, 0, st>>>(dP, Nb, E, E, Co, dY);
    TDTM_LAUNCHED();
    k_relu_grad<<<nblk(rows * Co), 256, 0, st>>>(dY, n->act[2 * l + 1], rows * Co);
    TDTM_LAUNCHED();
    if (colsum(n, dY, rows, Co, n->g + n->boff[l], st)) return 1;
    // dW (Co x 9 Ci) = dY^T cols; cols rebuilt from the layer input
    k_im2col<<<nblk(rows * 9 * Ci), 256, 0, st>>>(n->act[2 * l], Nb, E, E, Ci, n->cols);
    TDTM_LAUNCHED();
    if (to_bf16(dY, rows * Co, n->ab, st)) return 1;
    if (gemm<true, false>(Co, 9 * Ci, (int)rows, n->ab, Co, n->cols, 9 * Ci, n->g + n->woff[l], 9 * Ci, 0, st,
                          n->ws))
      return 1;
    if (l == 0) break;  // no gradient of the network input
    // dcols (rows x 9 Ci) = dY W; dX = col2im(dcols), the gradient of pooled output l - 1
    if (gemm<false, false>((int)rows, 9 * Ci, Co, n->ab, Co, n->wb + n->woff[l], 9 * Ci, n->dcols, 9 * Ci, 0, st))
      return 1;
    float* dX = (dY == n->dbuf[0]) ? n->dbuf[1] : n->dbuf[0];
    k_col2im<<<nblk(rows * Ci), 256, 0, st>>>(n->dcols, Nb, E, E, Ci, dX);
    TDTM_LAUNCHED();
    dP = dX;
  }
  return 0;
}

}  // namespace

extern "C" {

int tdtm_proposal_create(int num_classes, int max_batch, tdtm_proposal_net_t* out) {
  if (!out) return fail("out is NULL");
  *out = nullptr;
  if (num_classes < 2 || max_batch < 1 || max_batch > 4096) return fail("num_classes >= 2, 1 <= max_batch <= 4096");
  auto* n = new tdtm_proposal_net();
  n->L = num_classes;
  n->maxN = max_batch;
  long off = 0;
  for (int l = 0; l < PN_LAYERS; ++l) {
    n->woff[l] = off;
    off += (long)PN_CH[l + 1] * 9 * PN_CH[l];
    n->boff[l] = off;
    off += PN_CH[l + 1];
  }
  n->woff[PN_LAYERS] = off;
  off += (long)PN_HID * PN_FEAT;
  n->boff[PN_LAYERS] = off;
  off += PN_HID;
  n->woff[PN_LAYERS + 1] = off;
  off += (long)n->L * PN_HID;
  n->boff[PN_LAYERS + 1] = off;
  off += n->L;
  n->nparam = off;
  const long N = max_batch;
  long amax = N * PN_IN * PN_IN * PN_CH[1], cmax = 0;
  for (int l = 0; l < PN_LAYERS; ++l) {
    const long rows = N * edge(l) * edge(l);
    cmax = std::max(cmax, rows * 9 * PN_CH[l]);
    amax = std::max(amax, rows * PN_CH[l + 1]);
  }
  amax = std::max(amax, N * (long)std::max(n->L, PN_FEAT));
  n->cols_max = cmax;
  n->act_max = amax;
  int rc = 0;
  rc |= pn_alloc((void**)&n->w, off * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->g, off * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->m, off * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->v, off * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->wb, off * 2);
  for (int i = 0; i <= 2 * PN_LAYERS && !rc; ++i) {
    const int l = i / 2;
    long sz;
    if (i % 2 == 0) sz = N * edge(l) * edge(l) * PN_CH[l];          // layer input (pooled output)
    else sz = N * edge(l) * edge(l) * PN_CH[l + 1];                  // conv output
    rc |= pn_alloc((void**)&n->act[i], sz * 4);
  }
  rc |= rc ? 0 : pn_alloc((void**)&n->h1, N * PN_HID * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->z, N * n->L * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->dz, N * n->L * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->loss, N * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->cols, cmax * 2);
  rc |= rc ? 0 : pn_alloc((void**)&n->ab, std::max(cmax, amax) * 2);
  rc |= rc ? 0 : pn_alloc((void**)&n->dcols, std::max(cmax, amax) * 4);
  for (int i = 0; i < 3 && !rc; ++i) rc |= pn_alloc((void**)&n->dbuf[i], amax * 4);
  // split-K slices: the largest over the weight-gradient products at the largest batch
  // (splits grow with K, so smaller batches need no more)
  long wsmax = std::max(splits_floats(n->L, PN_HID, max_batch), splits_floats(PN_HID, PN_FEAT, max_batch));
  for (int l = 0; l < PN_LAYERS; ++l)
    wsmax = std::max(wsmax, splits_floats(PN_CH[l + 1], 9 * PN_CH[l], (int)(N * edge(l) * edge(l))));
  rc |= rc ? 0 : pn_alloc((void**)&n->ws, std::max(1L, wsmax) * 4);
  rc |= rc ? 0 : pn_alloc((void**)&n->csum, PN_COLSUM_BLOCKS * 256 * 4);
  if (rc) {
    pn_free(n);
    delete n;
    return 1;
  }
  *out = n;
  return 0;
}

int tdtm_proposal_destroy(tdtm_proposal_net_t n) {
  if (!n) return 0;
  pn_free(n);
  delete n;
  return 0;
}

int64_t tdtm_proposal_num_params(tdtm_proposal_net_t n) { return n ? n->nparam : -1; }

int tdtm_proposal_set_params(tdtm_proposal_net_t n, const float* params, void* stream) {
  if (!n || !params) return fail("null argument");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  TDTM_HIP(hipMemcpyAsync(n->w, params, n->nparam * 4, hipMemcpyDeviceToDevice, st));
  k_zero<<<nblk(n->nparam), 256, 0, st>>>(n->m, n->nparam);
  k_zero<<<nblk(n->nparam), 256, 0, st>>>(n->v, n->nparam);
  TDTM_LAUNCHED();
  n->steps = 0;
  return to_bf16(n->w, n->nparam, n->wb, st);
}

int tdtm_proposal_get_params(tdtm_proposal_net_t n, float* params, float* grads, void* stream) {
  if (!n) return fail("null argument");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (params) TDTM_HIP(hipMemcpyAsync(params, n->w, n->nparam * 4, hipMemcpyDeviceToDevice, st));
  if (grads) TDTM_HIP(hipMemcpyAsync(grads, n->g, n->nparam * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int tdtm_proposal_forward(tdtm_proposal_net_t n, const float* tiles, int batch, float* logits, void* stream) {
  if (!n || !tiles || !logits) return fail("null argument");
  if (batch < 1 || batch > n->maxN) return fail("batch must be in [1, max_batch]");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  TDTM_HIP(hipMemcpyAsync(n->act[0], tiles, (size_t)batch * PN_IN * PN_IN * 4, hipMemcpyDeviceToDevice, st));
  if (pn_forward(n, batch, st)) return 1;
  TDTM_HIP(hipMemcpyAsync(logits, n->z, (size_t)batch * n->L * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int tdtm_proposal_backward(tdtm_proposal_net_t n, const float* targets, int batch, float* loss, void* stream) {
  if (!n || !targets || !loss) return fail("null argument");
  if (batch < 1 || batch > n->maxN) return fail("batch must be in [1, max_batch]");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  k_softmax_xent<<<batch, 256, 0, st>>>(n->z, targets, n->L, batch, n->dz, n->loss);
  TDTM_LAUNCHED();
  if (pn_backward(n, batch, st)) return 1;
  TDTM_HIP(hipMemcpyAsync(loss, n->loss, (size_t)batch * 4, hipMemcpyDeviceToDevice, st));
  return 0;
}

int tdtm_proposal_adam_step(tdtm_proposal_net_t n, float lr, float beta1, float beta2, float eps, void* stream) {
  if (!n) return fail("null argument");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  n->steps += 1;
  const float c1 = 1.f - std::pow(beta1, (float)n->steps), c2 = 1.f - std::pow(beta2, (float)n->steps);
  k_adam<<<nblk(n->nparam), 256, 0, st>>>(n->w, n->m, n->v, n->g, n->nparam, lr, beta1, beta2, eps, c1, c2, n->wb);
  TDTM_LAUNCHED();
  return 0;
}

}  // extern "C"
