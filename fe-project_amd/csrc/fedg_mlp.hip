// Per-column MLP that predicts the initial iterate of the HEVI Newton solve
// (no counterpart in the reference, whose Newton loop starts from q^n and
// stops after one iteration).  Dense layers n_in -> H -> H -> H -> n_out with
// ReLU, BF16 operands on the matrix cores (v_mfma_f32_16x16x16_bf16) with
// FP32 accumulation, FP32 master weights and Adam.
//
// All products are "NT" GEMMs, C[m][n] = sum_k A[m][k] B[n][k] with A, B
// row-major BF16 and K padded to a multiple of 16:
//   forward     Z  = X  W^T        (A = X [B x K],    B = W  [N x K])
//   input grad  dX = dZ W          (A = dZ [B x N],   B = W^T [K x N])
//   weight grad dW = dZ^T X        (A = dZ^T [N x B], B = X^T [K x B])
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "fedg_internal.h"
#include "fedg_mlp.h"

namespace fedg {
namespace {

typedef short short4_ __attribute__((ext_vector_type(4)));
typedef float float4_ __attribute__((ext_vector_type(4)));

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffffu) ? 0x40 : 0);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// C (fp32, ldc) = A (M x K bf16, lda) * B (N x K bf16, ldb)^T; K % 16 == 0.
// 256 threads, 64x64 block tile, each wavefront a 32x32 tile of 2x2 MFMA
// blocks.  Out-of-range rows and columns load zeros and are not stored.
__global__ void __launch_bounds__(256) gemm_nt_bf16_kernel(const uint16_t* __restrict__ A, int lda,
                                                           const uint16_t* __restrict__ B, int ldb,
                                                           float* __restrict__ C, int ldc, int M, int N, int K,
                                                           int accumulate) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (w >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (w & 1) * 32;
  float4_ acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (float4_){0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, kq = (lane >> 4) * 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    short4_ fa[2], fb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = m0 + a * 16 + r;
      if (row < M) {
        const uint16_t* p = A + (size_t)row * lda + k0 + kq;
        fa[a] = *reinterpret_cast<const short4_*>(p);
      } else {
        fa[a] = (short4_){0, 0, 0, 0};
      }
      const int col = n0 + a * 16 + r;
      if (col < N) {
        const uint16_t* p = B + (size_t)col * ldb + k0 + kq;
        fb[a] = *reinterpret_cast<const short4_*>(p);
      } else {
        fb[a] = (short4_){0, 0, 0, 0};
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
  // D layout: lane holds rows 4*(lane/16) + i of column lane % 16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + b * 16 + r;
      if (col >= N) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = m0 + a * 16 + kq + i;
        if (row < M) {
          float* c = C + (size_t)row * ldc + col;
          *c = accumulate ? *c + acc[a][b][i] : acc[a][b][i];
        }
      }
    }
}

// fp32 [R x C] (ld) -> bf16 [R x Cp] with zero padding of columns C..Cp-1
__global__ void to_bf16_kernel(const float* __restrict__ x, int ldx, uint16_t* __restrict__ y, int R, int C, int Cp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)R * Cp) return;
  const int r = t / Cp, c = t % Cp;
  y[t] = c < C ? f2bf(x[(size_t)r * ldx + c]) : (uint16_t)0;
}

// fp32 [R x C] (ld) -> bf16 transposed [C x Rp] with zero padding of columns R..Rp-1
__global__ void to_bf16_t_kernel(const float* __restrict__ x, int ldx, uint16_t* __restrict__ y, int R, int C, int Rp) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // x: columns of x, y: rows of x
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    tile[j][tx] = (r < R && c < C) ? x[(size_t)r * ldx + c] : 0.f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;  // output row c, column r
    if (c < C && r < Rp) y[(size_t)c * Rp + r] = f2bf(tile[tx][j]);
  }
}

// H = relu(Z + b) (relu when act != 0); Z is [B x N]
__global__ void bias_act_kernel(float* __restrict__ Z, const float* __restrict__ bias, int Bn, int N, int act) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)Bn * N) return;
  float v = Z[t] + bias[t % N];
  if (act) v = v > 0.f ? v : 0.f;
  Z[t] = v;
}

// dZ = dH * (H > 0)
__global__ void relu_back_kernel(float* __restrict__ dH, const float* __restrict__ H, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(H[t] > 0.f)) dH[t] = 0.f;
}

// column sums of dZ [B x N] -> db[N]: 64 columns x 16 row groups per
// 1024-thread block, the groups' partial sums added in group order
__global__ void __launch_bounds__(1024) colsum_kernel(const float* __restrict__ dZ, float* __restrict__ db, int Bn,
                                                      int N) {
  __shared__ float part[16][64];
  const int tx = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + tx;
  float s = 0.f;
  if (c < N) {
    int b = g;
    for (; b + 48 < Bn; b += 64) {
      const float a0 = dZ[(size_t)b * N + c], a1 = dZ[(size_t)(b + 16) * N + c];
      const float a2 = dZ[(size_t)(b + 32) * N + c], a3 = dZ[(size_t)(b + 48) * N + c];
      s += a0;
      s += a1;
      s += a2;
      s += a3;
    }
    for (; b < Bn; b += 16) s += dZ[(size_t)b * N + c];
  }
  part[g][tx] = s;
  __syncthreads();
  if (g == 0 && c < N) {
    float t = part[0][tx];
    for (int k = 1; k < 16; ++k) t += part[k][tx];
    db[c] = t;
  }
}

// mean squared error: dY = 2 (Y - T) / (B N), per-block partial sums of (Y-T)^2
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY, long n,
                           float* __restrict__ partial) {
  __shared__ float s[256];
  float acc = 0.f;
  const float sc = 2.f / (float)n;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const float d = Y[t] - T[t];
    dY[t] = sc * d;
    acc += d * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
inline int pad16(int k) { return (k + 15) / 16 * 16; }

}  // namespace

struct MlpImpl {
  int nl;                 // layers
  int dims[8];            // dims[0] = n_in, ..., dims[nl] = n_out
  int kp[8];              // padded input width of each layer
  float* params;          // all W (N x K, unpadded) then b, fp32 master copy
  float* grads;
  float* m;
  float* v;
  long nparam;
  long woff[8], boff[8];
  uint16_t* Wb[8];        // bf16 [N x Kp]
  uint16_t* WbT[8];       // bf16 [K x Np16] (transposed, for the input gradient)
  int cap;                // batch capacity of the workspaces
  float* act[8];          // activations: act[0] = input (B x n_in), act[l] = H_l (B x dims[l])
  float* dact[8];         // gradients w.r.t. act
  uint16_t* xb;           // bf16 staging (largest B x Kp)
  uint16_t* tb1;          // bf16 staging for transposes
  uint16_t* tb2;
  float* partial;
  long step;
  hipStream_t stream;
};

static int mlp_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return FEDG_ERR_HIP;
  }
  return FEDG_OK;
}
#define MLP_TRY(x)                   \
  do {                               \
    int rc_ = (x);                   \
    if (rc_ != FEDG_OK) return rc_;  \
  } while (0)

static int gemm_nt(MlpImpl* p, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, int M,
                   int N, int K, int acc) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(gemm_nt_bf16_kernel, grid, dim3(256), 0, p->stream, A, lda, B, ldb, C, ldc, M, N, K, acc);
  return mlp_hip(hipGetLastError(), "gemm_nt_bf16_kernel");
}

int mlp_create(int n_in, int n_hidden, int n_hidden_layers, int n_out, unsigned seed, MlpImpl** out) {
  if (n_in < 1 || n_hidden < 1 || n_hidden_layers < 1 || n_hidden_layers > 6 || n_out < 1) {
    set_error("mlp: bad layer sizes");
    return FEDG_ERR_ARG;
  }
  MlpImpl* p = new MlpImpl();
  memset(p, 0, sizeof(*p));
  p->nl = n_hidden_layers + 1;
  p->dims[0] = n_in;
  for (int l = 1; l <= n_hidden_layers; ++l) p->dims[l] = n_hidden;
  p->dims[p->nl] = n_out;
  long off = 0;
  for (int l = 0; l < p->nl; ++l) {
    p->kp[l] = pad16(p->dims[l]);
    p->woff[l] = off;
    off += (long)p->dims[l + 1] * p->dims[l];
  }
  for (int l = 0; l < p->nl; ++l) {
    p->boff[l] = off;
    off += p->dims[l + 1];
  }
  p->nparam = off;
  // He-uniform initialisation, zero biases
  std::vector<float> h((size_t)off, 0.f);
  std::mt19937 rng(seed);
  for (int l = 0; l < p->nl; ++l) {
    const float lim = std::sqrt(6.f / (float)p->dims[l]) * (l == p->nl - 1 ? 0.1f : 1.f);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (long t = 0; t < (long)p->dims[l + 1] * p->dims[l]; ++t) h[p->woff[l] + t] = U(rng);
  }
  int rc = mlp_hip(hipStreamCreate(&p->stream), "hipStreamCreate");
  auto al = [&](void** q, size_t bytes) {
    if (rc == FEDG_OK) rc = mlp_hip(hipMalloc(q, bytes), "hipMalloc");
    if (rc == FEDG_OK) rc = mlp_hip(hipMemset(*q, 0, bytes), "hipMemset");
  };
  al((void**)&p->params, sizeof(float) * off);
  al((void**)&p->grads, sizeof(float) * off);
  al((void**)&p->m, sizeof(float) * off);
  al((void**)&p->v, sizeof(float) * off);
  al((void**)&p->partial, sizeof(float) * 256);
  for (int l = 0; l < p->nl; ++l) {
    al((void**)&p->Wb[l], sizeof(uint16_t) * (size_t)p->dims[l + 1] * p->kp[l]);
    al((void**)&p->WbT[l], sizeof(uint16_t) * (size_t)p->dims[l] * pad16(p->dims[l + 1]));
  }
  if (rc == FEDG_OK) rc = mlp_hip(hipMemcpy(p->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice), "upload");
  if (rc != FEDG_OK) {
    mlp_destroy(p);
    return rc;
  }
  *out = p;
  return FEDG_OK;
}

void mlp_destroy(MlpImpl* p) {
  if (!p) return;
  void* bufs[] = {p->params, p->grads, p->m, p->v, p->partial, p->xb, p->tb1, p->tb2};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int l = 0; l < 8; ++l) {
    if (p->Wb[l]) (void)hipFree(p->Wb[l]);
    if (p->WbT[l]) (void)hipFree(p->WbT[l]);
    if (p->act[l]) (void)hipFree(p->act[l]);
    if (p->dact[l]) (void)hipFree(p->dact[l]);
  }
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

static int mlp_reserve(MlpImpl* p, int B) {
  if (B <= p->cap) return FEDG_OK;
  const int cap = pad16(B);
  for (int l = 0; l <= p->nl; ++l) {
    if (p->act[l]) (void)hipFree(p->act[l]);
    if (p->dact[l]) (void)hipFree(p->dact[l]);
    p->act[l] = p->dact[l] = nullptr;
  }
  if (p->xb) (void)hipFree(p->xb);
  if (p->tb1) (void)hipFree(p->tb1);
  if (p->tb2) (void)hipFree(p->tb2);
  p->xb = p->tb1 = p->tb2 = nullptr;
  size_t wmax = 0;
  for (int l = 0; l <= p->nl; ++l) wmax = std::max(wmax, (size_t)pad16(p->dims[l]));
  int rc = FEDG_OK;
  for (int l = 0; l <= p->nl && rc == FEDG_OK; ++l) {
    rc = mlp_hip(hipMalloc((void**)&p->act[l], sizeof(float) * (size_t)cap * p->dims[l]), "hipMalloc");
    if (rc == FEDG_OK) rc = mlp_hip(hipMalloc((void**)&p->dact[l], sizeof(float) * (size_t)cap * p->dims[l]), "hipMalloc");
  }
  if (rc == FEDG_OK) rc = mlp_hip(hipMalloc((void**)&p->xb, sizeof(uint16_t) * (size_t)cap * wmax), "hipMalloc");
  if (rc == FEDG_OK) rc = mlp_hip(hipMalloc((void**)&p->tb1, sizeof(uint16_t) * (size_t)cap * wmax), "hipMalloc");
  if (rc == FEDG_OK) rc = mlp_hip(hipMalloc((void**)&p->tb2, sizeof(uint16_t) * (size_t)cap * wmax), "hipMalloc");
  if (rc == FEDG_OK) p->cap = cap;
  return rc;
}

// BF16 copies of the weights (and their transposes) from the fp32 masters
static int mlp_sync_weights(MlpImpl* p) {
  for (int l = 0; l < p->nl; ++l) {
    const int N = p->dims[l + 1], K = p->dims[l];
    const float* W = p->params + p->woff[l];
    hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk((long)N * p->kp[l], 256)), dim3(256), 0, p->stream, W, K, p->Wb[l],
                       N, K, p->kp[l]);
    MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_kernel"));
    dim3 g((K + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(to_bf16_t_kernel, g, dim3(256), 0, p->stream, W, K, p->WbT[l], N, K, pad16(N));
    MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_t_kernel"));
  }
  return FEDG_OK;
}

// forward of B rows of X (device fp32 [B x n_in]); output in act[nl]
static int mlp_fwd(MlpImpl* p, const float* X, int B) {
  MLP_TRY(mlp_reserve(p, B));
  MLP_TRY(mlp_hip(hipMemcpyAsync(p->act[0], X, sizeof(float) * (size_t)B * p->dims[0], hipMemcpyDeviceToDevice,
                                 p->stream), "copy"));
  for (int l = 0; l < p->nl; ++l) {
    const int K = p->dims[l], N = p->dims[l + 1], Kp = p->kp[l];
    hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk((long)B * Kp, 256)), dim3(256), 0, p->stream, p->act[l], K, p->xb,
                       B, K, Kp);
    MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_kernel"));
    MLP_TRY(gemm_nt(p, p->xb, Kp, p->Wb[l], Kp, p->act[l + 1], N, B, N, Kp, 0));
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((long)B * N, 256)), dim3(256), 0, p->stream, p->act[l + 1],
                       p->params + p->boff[l], B, N, l < p->nl - 1 ? 1 : 0);
    MLP_TRY(mlp_hip(hipGetLastError(), "bias_act_kernel"));
  }
  return FEDG_OK;
}

// backward from dact[nl] (set by the loss); gradients into p->grads
static int mlp_bwd(MlpImpl* p, int B) {
  for (int l = p->nl - 1; l >= 0; --l) {
    const int K = p->dims[l], N = p->dims[l + 1];
    float* dZ = p->dact[l + 1];
    if (l < p->nl - 1) {
      hipLaunchKernelGGL(relu_back_kernel, dim3(nblk((long)B * N, 256)), dim3(256), 0, p->stream, dZ, p->act[l + 1],
                         (long)B * N);
      MLP_TRY(mlp_hip(hipGetLastError(), "relu_back_kernel"));
    }
    hipLaunchKernelGGL(colsum_kernel, dim3(nblk(N, 64)), dim3(1024), 0, p->stream, dZ, p->grads + p->boff[l], B, N);
    MLP_TRY(mlp_hip(hipGetLastError(), "colsum_kernel"));
    // dW = dZ^T X : A = dZ^T [N x Bp], B = X^T [K x Bp]
    const int Bp = pad16(B);
    dim3 gz((N + 31) / 32, (B + 31) / 32), gx((K + 31) / 32, (B + 31) / 32);
    hipLaunchKernelGGL(to_bf16_t_kernel, gz, dim3(256), 0, p->stream, dZ, N, p->tb1, B, N, Bp);
    MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_t_kernel"));
    hipLaunchKernelGGL(to_bf16_t_kernel, gx, dim3(256), 0, p->stream, p->act[l], K, p->tb2, B, K, Bp);
    MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_t_kernel"));
    MLP_TRY(gemm_nt(p, p->tb1, Bp, p->tb2, Bp, p->grads + p->woff[l], K, N, K, Bp, 0));
    if (l > 0) {
      // dX = dZ W : A = dZ [B x Np], B = W^T [K x Np]
      const int Np = pad16(N);
      hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk((long)B * Np, 256)), dim3(256), 0, p->stream, dZ, N, p->tb1, B, N,
                         Np);
      MLP_TRY(mlp_hip(hipGetLastError(), "to_bf16_kernel"));
      MLP_TRY(gemm_nt(p, p->tb1, Np, p->WbT[l], Np, p->dact[l], K, B, K, Np, 0));
    }
  }
  return FEDG_OK;
}

int mlp_forward(MlpImpl* p, const float* X, float* Y, int B) {
  MLP_TRY(mlp_sync_weights(p));
  MLP_TRY(mlp_fwd(p, X, B));
  MLP_TRY(mlp_hip(hipMemcpyAsync(Y, p->act[p->nl], sizeof(float) * (size_t)B * p->dims[p->nl],
                                 hipMemcpyDeviceToDevice, p->stream), "copy"));
  return mlp_hip(hipStreamSynchronize(p->stream), "sync");
}

int mlp_loss_grad(MlpImpl* p, const float* X, const float* T, int B, float* loss) {
  MLP_TRY(mlp_sync_weights(p));
  MLP_TRY(mlp_fwd(p, X, B));
  const long n = (long)B * p->dims[p->nl];
  hipLaunchKernelGGL(mse_kernel, dim3(256), dim3(256), 0, p->stream, p->act[p->nl], T, p->dact[p->nl], n, p->partial);
  MLP_TRY(mlp_hip(hipGetLastError(), "mse_kernel"));
  MLP_TRY(mlp_bwd(p, B));
  float part[256];
  MLP_TRY(mlp_hip(hipMemcpyAsync(part, p->partial, sizeof(part), hipMemcpyDeviceToHost, p->stream), "copy"));
  MLP_TRY(mlp_hip(hipStreamSynchronize(p->stream), "sync"));
  double s = 0.0;
  for (int b = 0; b < 256; ++b) s += part[b];
  if (loss) *loss = (float)(s / (double)n);
  return FEDG_OK;
}

int mlp_adam(MlpImpl* p, float lr, float b1, float b2, float eps) {
  p->step += 1;
  const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(p->nparam, 256)), dim3(256), 0, p->stream, p->params, p->grads, p->m,
                     p->v, p->nparam, lr, b1, b2, eps, c1, c2);
  MLP_TRY(mlp_hip(hipGetLastError(), "adam_kernel"));
  return FEDG_OK;
}

long mlp_param_count(const MlpImpl* p) { return p->nparam; }
int mlp_get(MlpImpl* p, float* params_host, float* grads_host) {
  MLP_TRY(mlp_hip(hipStreamSynchronize(p->stream), "sync"));
  if (params_host) MLP_TRY(mlp_hip(hipMemcpy(params_host, p->params, sizeof(float) * p->nparam, hipMemcpyDeviceToHost), "copy"));
  if (grads_host) MLP_TRY(mlp_hip(hipMemcpy(grads_host, p->grads, sizeof(float) * p->nparam, hipMemcpyDeviceToHost), "copy"));
  return FEDG_OK;
}
int mlp_set(MlpImpl* p, const float* params_host) {
  MLP_TRY(mlp_hip(hipStreamSynchronize(p->stream), "sync"));
  MLP_TRY(mlp_hip(hipMemcpy(p->params, params_host, sizeof(float) * p->nparam, hipMemcpyHostToDevice), "copy"));
  MLP_TRY(mlp_hip(hipMemset(p->m, 0, sizeof(float) * p->nparam), "memset"));
  MLP_TRY(mlp_hip(hipMemset(p->v, 0, sizeof(float) * p->nparam), "memset"));
  p->step = 0;
  return FEDG_OK;
}
int mlp_dims(const MlpImpl* p, int* n_in, int* n_out) {
  if (n_in) *n_in = p->dims[0];
  if (n_out) *n_out = p->dims[p->nl];
  return FEDG_OK;
}
hipStream_t mlp_stream(const MlpImpl* p) { return p->stream; }

}  // namespace fedg
