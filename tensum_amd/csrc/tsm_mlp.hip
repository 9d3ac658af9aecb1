// tsm_mlp.hip -- MLP surrogate of the Monte Carlo quantities of interest:
// BF16 weights and activations on MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation, FP32 master weights, Adam.  No counterpart in the reference.
//
// Layers: z_l = x_l W_l^T + b_l (W_l: n_out x n_in, row-major), x_{l+1} =
// GELU(z_l) (exact, erf form) for hidden layers, y = z_last.  Activations are
// row-major [batch x features] in bf16; pre-activations z_l are kept in FP32
// for the backward pass.  Backward (loss = mean over batch and outputs of
// (y - t)^2):
//   dy = 2 (y - t) / (B n_out)            (fp32 -> bf16)
//   dW_l = dz_l^T x_l,  db_l = column sums of dz_l   (fp32)
//   dx_l = dz_l W_l,    dz_{l-1} = dx_l * GELU'(z_{l-1})   (-> bf16)
// Every product is one launch of the tiled GEMM below with the operands'
// transposes absorbed by the LDS staging.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tsm.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

__host__ __device__ inline u16 f2bf(float f) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_GELU_BWD = 2, EPI_BIAS_F32 = 3 };

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS rows padded by 16 bytes

// C[M x N] = A[M x K] * B[K x N].  AK: A stored M x K with k contiguous (else
// K x M with m contiguous); BK_: B stored N x K with k contiguous (else K x N
// with n contiguous).  256 threads = 4 waves, each wave one 32 x 32 MFMA tile
// of the 64 x 64 block tile.
template <bool AK, bool BKM, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const u16* __restrict__ A,
                                              int lda, const u16* __restrict__ B, int ldb,
                                              float* __restrict__ Cf, u16* __restrict__ Cb, int ldc,
                                              const float* __restrict__ bias,
                                              const float* __restrict__ Z, float* __restrict__ Zout) {
  __shared__ __attribute__((aligned(16))) u16 As[BM * LDK];
  __shared__ __attribute__((aligned(16))) u16 Bs[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;

  for (int k0 = 0; k0 < K; k0 += BK) {
    // ---- stage A (64 x 32) and B (64 x 32) in LDS, k contiguous per row ----
    if (AK) {
      const int row = tid >> 2, ks = (tid & 3) * 8, gm = m0 + row, gk = k0 + ks;
      u16 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (gm < M && gk + j < K) ? A[(size_t)gm * lda + gk + j] : (u16)0;
#pragma unroll
      for (int j = 0; j < 8; ++j) As[row * LDK + ks + j] = v[j];
    } else {
      const int kk = tid >> 3, ms = (tid & 7) * 8, gk = k0 + kk;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gm = m0 + ms + j;
        As[(ms + j) * LDK + kk] = (gm < M && gk < K) ? A[(size_t)gk * lda + gm] : (u16)0;
      }
    }
    if (BKM) {
      const int row = tid >> 2, ks = (tid & 3) * 8, gn = n0 + row, gk = k0 + ks;
      u16 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (gn < N && gk + j < K) ? B[(size_t)gn * ldb + gk + j] : (u16)0;
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[row * LDK + ks + j] = v[j];
    } else {
      const int kk = tid >> 3, ns = (tid & 7) * 8, gk = k0 + kk;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gn = n0 + ns + j;
        Bs[(ns + j) * LDK + kk] = (gn < N && gk < K) ? B[(size_t)gk * ldb + gn] : (u16)0;
      }
    }
    __syncthreads();
    // ---- two MFMA 32x32x16 per k tile; lanes 0-31 take k 0-7, 32-63 k 8-15
    //      of each 16-slice (the same map for A and B, so the k order is moot)
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int ko = kk + (lane >> 5) * 8;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wm + (lane & 31)) * LDK + ko]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wn + (lane & 31)) * LDK + ko]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // ---- epilogue: acc[i] is C[row][col] with col = lane % 32 and
  //      row = 8 (i / 4) + 4 (lane / 32) + i % 4 within the wave tile
  const int gn = n0 + wn + (lane & 31);
  if (gn >= N) return;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int gm = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    if (gm >= M) continue;
    const size_t o = (size_t)gm * ldc + gn;
    float c = acc[i];
    if (EPI == EPI_F32) {
      Cf[o] = c;
    } else if (EPI == EPI_BIAS_F32) {
      Cf[o] = c + bias[gn];
    } else if (EPI == EPI_BIAS_GELU) {
      c += bias[gn];
      if (Zout) Zout[o] = c;
      Cb[o] = f2bf(gelu(c));
    } else {  // EPI_GELU_BWD
      Cb[o] = f2bf(c * gelu_d(Z[o]));
    }
  }
}

// Column sums of a [rows x cols] bf16 or fp32 matrix, fp32, fixed order.
__global__ void k_colsum_bf16(int rows, int cols, const u16* X, float* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += bf2f(X[(size_t)r * cols + c]);
  out[c] = s;
}

// dy = 2 (y - t) / (B n), loss partials per block (sum of squares)
__global__ __launch_bounds__(256) void k_mse(int n, const float* y, const float* t, float scale,
                                             u16* dy, float* partial) {
  __shared__ float red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < n) {
    const float d = y[i] - t[i];
    e2 = d * d;
    dy[i] = f2bf(2.0f * d * scale);
  }
  for (int o = 32; o > 0; o >>= 1) e2 += __shfl_down(e2, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Adam on fp32 master weights; refreshes the bf16 copy.
__global__ void k_adam(int n, float* w, const float* g, float* m, float* v, u16* wb, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(int n, const float* x, u16* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// rows of a [rows x cols] fp32 matrix picked by idx -> bf16 (inputs) or fp32 (targets)
__global__ void k_gather_rows(int rows, int cols, const int* idx, const float* src, u16* db,
                              float* df) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int r = i / cols, c = i % cols;
  const float x = src[(size_t)idx[r] * cols + c];
  if (db) db[i] = f2bf(x);
  if (df) df[i] = x;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct tsm_mlp {
  int L = 0, maxb = 0;
  std::vector<int> w;                 // widths, L + 1
  std::vector<size_t> woff, boff;     // offsets of W_l, b_l in the flat parameter vector
  size_t np = 0;
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;  // fp32 params, grads, Adam
  u16* Pb = nullptr;                  // bf16 copy of the parameters
  std::vector<u16*> X;                // activations x_0..x_{L-1} (bf16), maxb x w_l
  std::vector<float*> Z;              // pre-activations z_0..z_{L-1} (fp32)
  u16 *dA = nullptr, *dB = nullptr;   // backward ping-pong, maxb x max width (bf16)
  float *Y = nullptr, *T = nullptr, *part = nullptr;
  float *Xin = nullptr, *Tin = nullptr;  // staged training set
  int* idx = nullptr;
  size_t nin_cap = 0;
  long adam_t = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<void*> allocs;
  template <class T>
  hipError_t al(T** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(T) + 16);
    if (e == hipSuccess) allocs.push_back((void*)*p);
    return e;
  }
  ~tsm_mlp() {
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {
thread_local std::string g_mlp_err;
int mfail(int code, const std::string& m) {
  g_mlp_err = m;
  return code;
}
#define MLP_TRY(x)                                                     \
  do {                                                                 \
    hipError_t _e = (x);                                               \
    if (_e != hipSuccess) return mfail(TSM_E_HIP, hipGetErrorString(_e)); \
  } while (0)

template <bool AK, bool BKM, int EPI>
void gemm(hipStream_t st, int M, int N, int K, const u16* A, int lda, const u16* B, int ldb,
          float* Cf, u16* Cb, int ldc, const float* bias, const float* Z, float* Zout) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN);
  hipLaunchKernelGGL((k_gemm<AK, BKM, EPI>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, Cf,
                     Cb, ldc, bias, Z, Zout);
}

// forward of nb rows already in m->X[0]; output fp32 in m->Y (nb x w[L])
void forward(tsm_mlp* m, int nb, bool keep_z) {
  for (int l = 0; l < m->L; ++l) {
    const int K = m->w[l], N = m->w[l + 1];
    const u16* W = m->Pb + m->woff[l];
    const float* b = m->P + m->boff[l];
    if (l + 1 < m->L)
      gemm<true, true, EPI_BIAS_GELU>(m->st, nb, N, K, m->X[l], K, W, K, nullptr, m->X[l + 1], N, b,
                                      nullptr, keep_z ? m->Z[l] : nullptr);
    else
      gemm<true, true, EPI_BIAS_F32>(m->st, nb, N, K, m->X[l], K, W, K, m->Y, nullptr, N, b,
                                     nullptr, nullptr);
  }
}

// loss and gradient (into m->G) of nb rows in X[0] with targets in m->T
void backward(tsm_mlp* m, int nb) {
  const int L = m->L, no = m->w[L];
  const int n = nb * no;
  hipLaunchKernelGGL(k_mse, dim3((n + 255) / 256), dim3(256), 0, m->st, n, m->Y, m->T,
                     1.0f / (float)n, m->dA, m->part);
  u16 *dz = m->dA, *dx = m->dB;
  for (int l = L - 1; l >= 0; --l) {
    const int K = m->w[l], N = m->w[l + 1];
    // dW_l [N x K] = dz^T [N x nb] * x_l [nb x K]
    gemm<false, false, EPI_F32>(m->st, N, K, nb, dz, N, m->X[l], K, m->G + m->woff[l], nullptr, K,
                                nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_colsum_bf16, dim3((N + 63) / 64), dim3(64), 0, m->st, nb, N, dz,
                       m->G + m->boff[l]);
    if (l > 0) {
      // dx_l [nb x K] = dz [nb x N] * W_l [N x K]; times GELU'(z_{l-1})
      gemm<true, false, EPI_GELU_BWD>(m->st, nb, K, N, dz, N, m->Pb + m->woff[l], K, nullptr, dx, K,
                                      nullptr, m->Z[l - 1], nullptr);
      std::swap(dz, dx);
    }
  }
}

int stage_rows(tsm_mlp* m, int rows, const float* x, const float* t) {
  const int ni = m->w[0], no = m->w[m->L];
  float* tmp = nullptr;
  MLP_TRY(hipMalloc(&tmp, (size_t)rows * ni * sizeof(float)));
  hipError_t e = hipMemcpyAsync(tmp, x, (size_t)rows * ni * sizeof(float), hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((rows * ni + 255) / 256), dim3(256), 0, m->st, rows * ni, tmp,
                       m->X[0]);
    e = hipGetLastError();
  }
  if (e == hipSuccess && t)
    e = hipMemcpyAsync(m->T, t, (size_t)rows * no * sizeof(float), hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  (void)hipFree(tmp);
  MLP_TRY(e);
  return TSM_OK;
}

}  // namespace

extern "C" const char* tsm_mlp_last_error(void) { return g_mlp_err.c_str(); }

extern "C" int tsm_mlp_create(int n_layers, const int* widths, int max_batch,
                              unsigned long long seed, tsm_mlp** out) {
  if (!out || n_layers < 1 || n_layers > 16 || !widths || max_batch < 1)
    return mfail(TSM_E_ARG, "bad argument");
  for (int l = 0; l <= n_layers; ++l)
    if (widths[l] < 1 || widths[l] > 4096) return mfail(TSM_E_ARG, "layer width out of range");
  tsm_mlp* m = new tsm_mlp();
  m->L = n_layers;
  m->maxb = max_batch;
  m->w.assign(widths, widths + n_layers + 1);
  int wmax = 0;
  for (int l = 0; l < n_layers; ++l) {
    m->woff.push_back(m->np);
    m->np += (size_t)m->w[l + 1] * m->w[l];
    m->boff.push_back(m->np);
    m->np += (size_t)m->w[l + 1];
  }
  for (int x : m->w) wmax = std::max(wmax, x);
  // He-style uniform init, variance 1/fan_in (splitmix64 stream)
  std::vector<float> h(m->np, 0.0f);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t q = z;
    q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ull;
    q = (q ^ (q >> 27)) * 0x94d049bb133111ebull;
    return (double)((q ^ (q >> 31)) >> 11) * 0x1p-53;
  };
  for (int l = 0; l < n_layers; ++l) {
    const double a = std::sqrt(3.0 / m->w[l]);
    for (size_t i = 0; i < (size_t)m->w[l + 1] * m->w[l]; ++i)
      h[m->woff[l] + i] = (float)(a * (2.0 * next() - 1.0));
  }
  hipError_t e = hipStreamCreate(&m->st);
  if (e == hipSuccess) e = m->al(&m->P, m->np);
  if (e == hipSuccess) e = m->al(&m->G, m->np);
  if (e == hipSuccess) e = m->al(&m->Mo, m->np);
  if (e == hipSuccess) e = m->al(&m->Vo, m->np);
  if (e == hipSuccess) e = m->al(&m->Pb, m->np);
  m->X.assign(n_layers, nullptr);
  m->Z.assign(n_layers, nullptr);
  for (int l = 0; l < n_layers && e == hipSuccess; ++l) {
    e = m->al(&m->X[l], (size_t)max_batch * m->w[l]);
    if (e == hipSuccess) e = m->al(&m->Z[l], (size_t)max_batch * m->w[l + 1]);
  }
  if (e == hipSuccess) e = m->al(&m->dA, (size_t)max_batch * wmax);
  if (e == hipSuccess) e = m->al(&m->dB, (size_t)max_batch * wmax);
  if (e == hipSuccess) e = m->al(&m->Y, (size_t)max_batch * m->w[n_layers]);
  if (e == hipSuccess) e = m->al(&m->T, (size_t)max_batch * m->w[n_layers]);
  if (e == hipSuccess) e = m->al(&m->part, (size_t)max_batch * m->w[n_layers] / 256 + 2);
  if (e == hipSuccess) e = hipMemcpy(m->P, h.data(), m->np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->Mo, 0, m->np * sizeof(float));
  if (e == hipSuccess) e = hipMemset(m->Vo, 0, m->np * sizeof(float));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->st,
                       (int)m->np, m->P, m->Pb);
    e = hipStreamSynchronize(m->st);
  }
  if (e != hipSuccess) {
    delete m;
    return mfail(TSM_E_HIP, hipGetErrorString(e));
  }
  *out = m;
  return TSM_OK;
}

extern "C" void tsm_mlp_destroy(tsm_mlp* m) { delete m; }

extern "C" long tsm_mlp_n_params(const tsm_mlp* m) { return m ? (long)m->np : -1; }

extern "C" int tsm_mlp_get_params(tsm_mlp* m, float* p) {
  if (!m || !p) return mfail(TSM_E_ARG, "bad argument");
  MLP_TRY(hipStreamSynchronize(m->st));
  MLP_TRY(hipMemcpy(p, m->P, m->np * sizeof(float), hipMemcpyDeviceToHost));
  return TSM_OK;
}

extern "C" int tsm_mlp_set_params(tsm_mlp* m, const float* p) {
  if (!m || !p) return mfail(TSM_E_ARG, "bad argument");
  MLP_TRY(hipStreamSynchronize(m->st));
  MLP_TRY(hipMemcpy(m->P, p, m->np * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->st,
                     (int)m->np, m->P, m->Pb);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(m->st));
  return TSM_OK;
}

extern "C" int tsm_mlp_forward(tsm_mlp* m, long n, const float* x, float* y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return mfail(TSM_E_ARG, "bad argument");
  const int ni = m->w[0], no = m->w[m->L];
  for (long r0 = 0; r0 < n; r0 += m->maxb) {
    const int nb = (int)std::min<long>(m->maxb, n - r0);
    int rc = stage_rows(m, nb, x + (size_t)r0 * ni, nullptr);
    if (rc) return rc;
    forward(m, nb, false);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpyAsync(y + (size_t)r0 * no, m->Y, (size_t)nb * no * sizeof(float),
                           hipMemcpyDeviceToHost, m->st));
    MLP_TRY(hipStreamSynchronize(m->st));
  }
  return TSM_OK;
}

extern "C" int tsm_mlp_gradient(tsm_mlp* m, int n, const float* x, const float* t, double* loss,
                                float* grad) {
  if (!m || n < 1 || n > m->maxb || !x || !t) return mfail(TSM_E_ARG, "bad argument");
  int rc = stage_rows(m, n, x, t);
  if (rc) return rc;
  forward(m, n, true);
  backward(m, n);
  MLP_TRY(hipGetLastError());
  const int nparts = (n * m->w[m->L] + 255) / 256;
  std::vector<float> part(nparts);
  MLP_TRY(hipMemcpyAsync(part.data(), m->part, nparts * sizeof(float), hipMemcpyDeviceToHost, m->st));
  if (grad) MLP_TRY(hipMemcpyAsync(grad, m->G, m->np * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MLP_TRY(hipStreamSynchronize(m->st));
  double s = 0.0;
  for (float p : part) s += p;
  if (loss) *loss = s / ((double)n * m->w[m->L]);
  return TSM_OK;
}

extern "C" int tsm_mlp_train(tsm_mlp* m, long n, const float* x, const float* t, int epochs,
                             int batch, float lr, unsigned long long seed, double* epoch_loss) {
  if (!m || n < 1 || !x || !t || epochs < 0 || batch < 1 || batch > m->maxb)
    return mfail(TSM_E_ARG, "bad argument");
  const int ni = m->w[0], no = m->w[m->L];
  if ((size_t)n > m->nin_cap) {
    MLP_TRY(m->al(&m->Xin, (size_t)n * ni));
    MLP_TRY(m->al(&m->Tin, (size_t)n * no));
    MLP_TRY(m->al(&m->idx, (size_t)n));
    m->nin_cap = (size_t)n;
  }
  MLP_TRY(hipMemcpyAsync(m->Xin, x, (size_t)n * ni * sizeof(float), hipMemcpyHostToDevice, m->st));
  MLP_TRY(hipMemcpyAsync(m->Tin, t, (size_t)n * no * sizeof(float), hipMemcpyHostToDevice, m->st));
  std::vector<int> perm(n);
  for (long i = 0; i < n; ++i) perm[i] = (int)i;
  uint64_t z = seed ^ 0x5851f42d4c957f2dull;
  auto next = [&]() {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t q = z;
    q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ull;
    q = (q ^ (q >> 27)) * 0x94d049bb133111ebull;
    return q ^ (q >> 31);
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int nmb = (int)((n + batch - 1) / batch);
  std::vector<float> part;
  for (int ep = 0; ep < epochs; ++ep) {
    for (long i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
    MLP_TRY(hipMemcpyAsync(m->idx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, m->st));
    MLP_TRY(hipStreamSynchronize(m->st));  // perm is reused by the next epoch
    double esum = 0.0;
    for (int b = 0; b < nmb; ++b) {
      const int r0 = b * batch, nb = (int)std::min<long>(batch, n - r0);
      hipLaunchKernelGGL(k_gather_rows, dim3((nb * ni + 255) / 256), dim3(256), 0, m->st, nb, ni,
                         m->idx + r0, m->Xin, m->X[0], nullptr);
      hipLaunchKernelGGL(k_gather_rows, dim3((nb * no + 255) / 256), dim3(256), 0, m->st, nb, no,
                         m->idx + r0, m->Tin, nullptr, m->T);
      forward(m, nb, true);
      backward(m, nb);
      ++m->adam_t;
      const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->st,
                         (int)m->np, m->P, m->G, m->Mo, m->Vo, m->Pb, lr, b1, b2, eps, c1, c2);
      MLP_TRY(hipGetLastError());
      if (epoch_loss) {
        const int nparts = (nb * no + 255) / 256;
        part.resize(nparts);
        MLP_TRY(hipMemcpyAsync(part.data(), m->part, nparts * sizeof(float), hipMemcpyDeviceToHost,
                               m->st));
        MLP_TRY(hipStreamSynchronize(m->st));
        double s = 0.0;
        for (float p : part) s += p;
        esum += s;
      }
    }
    if (epoch_loss) epoch_loss[ep] = esum / ((double)n * no);
  }
  MLP_TRY(hipStreamSynchronize(m->st));
  return TSM_OK;
}

// Plain GEMM through the same kernel, for tests of the MFMA tiling:
// C[M x N] (fp32) = A * B with A, B given in bf16 bit patterns and the layout
// flags of the kernel (a_kmajor: A is M x K row-major, else K x M; b_kmajor: B
// is N x K row-major, else K x N).
extern "C" int tsm_mlp_gemm_test(int M, int N, int K, const unsigned short* A, int a_kmajor,
                                 const unsigned short* B, int b_kmajor, float* C) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return mfail(TSM_E_ARG, "bad argument");
  u16 *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  MLP_TRY(hipMalloc(&dA, (size_t)M * K * 2));
  hipError_t e = hipMalloc(&dB, (size_t)N * K * 2);
  if (e == hipSuccess) e = hipMalloc(&dC, (size_t)M * N * 4);
  if (e == hipSuccess) e = hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, (size_t)N * K * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const int lda = a_kmajor ? K : M, ldb = b_kmajor ? K : N;
    if (a_kmajor && b_kmajor) gemm<true, true, EPI_F32>(0, M, N, K, dA, lda, dB, ldb, dC, nullptr, N, nullptr, nullptr, nullptr);
    else if (a_kmajor) gemm<true, false, EPI_F32>(0, M, N, K, dA, lda, dB, ldb, dC, nullptr, N, nullptr, nullptr, nullptr);
    else if (b_kmajor) gemm<false, true, EPI_F32>(0, M, N, K, dA, lda, dB, ldb, dC, nullptr, N, nullptr, nullptr, nullptr);
    else gemm<false, false, EPI_F32>(0, M, N, K, dA, lda, dB, ldb, dC, nullptr, N, nullptr, nullptr, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
  MLP_TRY(e);
  return TSM_OK;
}
