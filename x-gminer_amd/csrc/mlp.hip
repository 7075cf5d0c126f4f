// Cost model of the learned scheduler: a 16 -> 64 -> 64 -> 5 MLP (ReLU),
// BF16 operands on the matrix cores (v_mfma_f32_16x16x16_bf16), FP32
// accumulation, FP32 master weights with Adam.  Training (forward,
// backward, Adam) and batched inference run entirely on the GPU through
// the kernels below; the host only orders batches.  The model only ranks
// schedules -- it never touches a count.
//
// Parameter vector (FP32, kMlpParams floats): W1[16x64] b1[64] W2[64x64]
// b2[64] W3[64x5] b3[5], row-major (input index major).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/xgm.h"
#include "host_util.h"

namespace xgm {
namespace {

constexpr int IN = 16, H = 64, OUT = 5;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * OUT, NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == XGM_MLP_PARAMS, "parameter layout");

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ bf16 f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

// C[M x N] (+)= A[M x K] * B[K x N]; A(r,k) = A[r*sar + k*sac], B(k,c) =
// B[k*sbr + c*sbc], C(r,c) = C[r*ldc + c].  One wave per 16 x 16 output tile;
// the K loop feeds 16 x 16 x 16 BF16 MFMAs.  Out-of-range operands read 0.
// Lane l holds A[l%16][4*(l/16)+i], B[4*(l/16)+i][l%16] and produces
// D[4*(l/16)+i][l%16] (i = 0..3).
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const bf16 *__restrict__ A, int64_t sar,
                                                   int64_t sac, const bf16 *__restrict__ B, int64_t sbr,
                                                   int64_t sbc, float *__restrict__ C, int64_t ldc, int accumulate) {
  const int lane = threadIdx.x & 63;
  const int tiles_n = (N + 15) / 16;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  if (tm * 16 >= M) return;
  const int r = tm * 16 + (lane & 15);   // A row of this lane
  const int c = tn * 16 + (lane & 15);   // B column of this lane
  const int kq = 4 * (lane >> 4);
  v4f32 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4i16 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + kq + i;
      a[i] = (r < M && k < K) ? (short)A[r * sar + k * sac] : (short)0;
      b[i] = (c < N && k < K) ? (short)B[k * sbr + c * sbc] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rr = tm * 16 + kq + i;
    if (rr < M && c < N) {
      float *p = C + rr * ldc + c;
      *p = accumulate ? *p + acc[i] : acc[i];
    }
  }
}

// Z += bias (per column), then act = relu(Z) (or identity); writes FP32 Z
// (pre-activation, kept for the backward pass) and the BF16 activation.
__global__ void k_bias_act(int B, int N, float *__restrict__ Z, const float *__restrict__ bias, int relu,
                           bf16 *__restrict__ act_bf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * N) return;
  const float z = Z[i] + bias[i % N];
  Z[i] = z;
  if (act_bf) act_bf[i] = f2bf(relu ? fmaxf(z, 0.f) : z);
}

// dY = 2 * mask * (Y - T) * inv_count; loss += sum mask * (Y - T)^2 * inv_count
__global__ void k_loss_grad(int B, const float *__restrict__ Y, const float *__restrict__ T,
                            const float *__restrict__ Mk, float inv_count, float *__restrict__ dY,
                            bf16 *__restrict__ dYb, float *__restrict__ loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (int64_t)B * OUT) {
    const float d = (Y[i] - T[i]) * Mk[i];
    const float g = 2.f * d * inv_count;
    dY[i] = g;
    dYb[i] = f2bf(g);
    l = d * d * inv_count;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// dZ = dA * (Zpre > 0) (FP32 for bias grads, BF16 for the next GEMM)
__global__ void k_relu_back(int64_t n, const float *__restrict__ dA, const float *__restrict__ Zpre,
                            float *__restrict__ dZ, bf16 *__restrict__ dZb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = Zpre[i] > 0.f ? dA[i] : 0.f;
  dZ[i] = g;
  dZb[i] = f2bf(g);
}

// db[c] = sum_r dZ[r, c] (one thread per column, rows strided by blocks)
__global__ void k_colsum(int B, int N, const float *__restrict__ dZ, float *__restrict__ db) {
  const int c = threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int r = blockIdx.x; r < B; r += gridDim.x) s += dZ[(int64_t)r * N + c];
  atomicAdd(&db[c], s);
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_to_bf16(int n, const float *__restrict__ x, bf16 *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Gather a batch: rows idx[0..B) of X (n x 16 FP32) -> BF16, T and mask.
__global__ void k_gather(int B, const int32_t *__restrict__ idx, const float *__restrict__ X,
                         const float *__restrict__ T, const float *__restrict__ Mk, bf16 *__restrict__ Xb,
                         float *__restrict__ Tb, float *__restrict__ Mb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * IN) return;
  const int r = i / IN, f = i - r * IN;
  const int64_t src = idx ? idx[r] : r;
  Xb[i] = f2bf(X[src * IN + f]);
  if (T && f < OUT) {
    Tb[r * OUT + f] = T[src * OUT + f];
    Mb[r * OUT + f] = Mk[src * OUT + f];
  }
}

__global__ void k_zero_f32(int64_t n, float *p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace xgm

using namespace xgm;

struct xgm_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 master, grads, Adam moments
  bf16 *w1 = nullptr, *w2 = nullptr, *w3 = nullptr;              // BF16 weight copies
  int64_t adam_t = 0;
  // workspace for batches of up to cap rows
  int64_t cap = 0;
  bf16 *xb = nullptr, *a1b = nullptr, *a2b = nullptr, *dyb = nullptr, *dzb = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *tb = nullptr, *mb = nullptr, *dy = nullptr, *da = nullptr,
        *dz = nullptr, *loss = nullptr;
  int32_t *idx = nullptr;

  int fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return XGM_EHIP;
  }
  void free_ws() {
    for (void *q : {(void *)xb, (void *)a1b, (void *)a2b, (void *)dyb, (void *)dzb, (void *)z1, (void *)z2,
                    (void *)y, (void *)tb, (void *)mb, (void *)dy, (void *)da, (void *)dz, (void *)idx})
      (void)hipFree(q);
    xb = a1b = a2b = dyb = dzb = nullptr;
    z1 = z2 = y = tb = mb = dy = da = dz = nullptr;
    idx = nullptr;
    cap = 0;
  }
  ~xgm_mlp() {
    (void)hipSetDevice(device);
    free_ws();
    for (void *q : {(void *)p, (void *)g, (void *)m, (void *)v, (void *)w1, (void *)w2, (void *)w3, (void *)loss})
      (void)hipFree(q);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

#define MLP_TRY(mdl, call)                                    \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return (mdl)->fail(_e, #call);      \
  } while (0)

int ensure_ws(xgm_mlp *md, int64_t rows) {
  if (md->cap >= rows) return XGM_OK;
  md->free_ws();
  const int64_t r = std::max<int64_t>(rows, 1024);
  MLP_TRY(md, hipMalloc(&md->xb, r * IN * sizeof(bf16)));
  MLP_TRY(md, hipMalloc(&md->a1b, r * H * sizeof(bf16)));
  MLP_TRY(md, hipMalloc(&md->a2b, r * H * sizeof(bf16)));
  MLP_TRY(md, hipMalloc(&md->dyb, r * OUT * sizeof(bf16)));
  MLP_TRY(md, hipMalloc(&md->dzb, r * H * sizeof(bf16)));
  MLP_TRY(md, hipMalloc(&md->z1, r * H * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->z2, r * H * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->y, r * OUT * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->tb, r * OUT * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->mb, r * OUT * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->dy, r * OUT * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->da, r * H * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->dz, r * H * sizeof(float)));
  MLP_TRY(md, hipMalloc(&md->idx, r * sizeof(int32_t)));
  md->cap = r;
  return XGM_OK;
}

void gemm(hipStream_t s, int M, int N, int K, const bf16 *A, int64_t sar, int64_t sac, const bf16 *B, int64_t sbr,
          int64_t sbc, float *C, int64_t ldc, int acc) {
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  k_gemm_bf16<<<(tiles + 3) / 4, 256, 0, s>>>(M, N, K, A, sar, sac, B, sbr, sbc, C, ldc, acc);
}

int refresh_bf16(xgm_mlp *md) {
  k_to_bf16<<<blocks_for(IN * H), 256, 0, md->stream>>>(IN * H, md->p + OFF_W1, md->w1);
  k_to_bf16<<<blocks_for(H * H), 256, 0, md->stream>>>(H * H, md->p + OFF_W2, md->w2);
  k_to_bf16<<<blocks_for(H * OUT), 256, 0, md->stream>>>(H * OUT, md->p + OFF_W3, md->w3);
  MLP_TRY(md, hipGetLastError());
  return XGM_OK;
}

// Forward over rows already gathered into md->xb: fills z1, a1b, z2, a2b, y.
void forward(xgm_mlp *md, int B) {
  hipStream_t s = md->stream;
  gemm(s, B, H, IN, md->xb, IN, 1, md->w1, H, 1, md->z1, H, 0);
  k_bias_act<<<blocks_for((int64_t)B * H), 256, 0, s>>>(B, H, md->z1, md->p + OFF_B1, 1, md->a1b);
  gemm(s, B, H, H, md->a1b, H, 1, md->w2, H, 1, md->z2, H, 0);
  k_bias_act<<<blocks_for((int64_t)B * H), 256, 0, s>>>(B, H, md->z2, md->p + OFF_B2, 1, md->a2b);
  gemm(s, B, OUT, H, md->a2b, H, 1, md->w3, OUT, 1, md->y, OUT, 0);
  k_bias_act<<<blocks_for((int64_t)B * OUT), 256, 0, s>>>(B, OUT, md->y, md->p + OFF_B3, 0, nullptr);
}

// Backward after forward + loss grad (md->dy / md->dyb): writes md->g.
void backward(xgm_mlp *md, int B) {
  hipStream_t s = md->stream;
  float *g = md->g;
  k_zero_f32<<<blocks_for(NPARAM), 256, 0, s>>>(NPARAM, g);
  // layer 3: dW3 = A2^T dY, db3 = colsum dY, dA2 = dY W3^T
  gemm(s, H, OUT, B, md->a2b, 1, H, md->dyb, OUT, 1, g + OFF_W3, OUT, 0);
  k_colsum<<<64, 64, 0, s>>>(B, OUT, md->dy, g + OFF_B3);
  gemm(s, B, H, OUT, md->dyb, OUT, 1, md->w3, 1, OUT, md->da, H, 0);
  k_relu_back<<<blocks_for((int64_t)B * H), 256, 0, s>>>((int64_t)B * H, md->da, md->z2, md->dz, md->dzb);
  // layer 2
  gemm(s, H, H, B, md->a1b, 1, H, md->dzb, H, 1, g + OFF_W2, H, 0);
  k_colsum<<<64, 64, 0, s>>>(B, H, md->dz, g + OFF_B2);
  gemm(s, B, H, H, md->dzb, H, 1, md->w2, 1, H, md->da, H, 0);
  k_relu_back<<<blocks_for((int64_t)B * H), 256, 0, s>>>((int64_t)B * H, md->da, md->z1, md->dz, md->dzb);
  // layer 1
  gemm(s, IN, H, B, md->xb, 1, IN, md->dzb, H, 1, g + OFF_W1, H, 0);
  k_colsum<<<64, 64, 0, s>>>(B, H, md->dz, g + OFF_B1);
}

int upload(xgm_mlp *md, void **dst, const void *src, size_t bytes) {
  MLP_TRY(md, hipMalloc(dst, std::max<size_t>(bytes, 4)));
  if (bytes) MLP_TRY(md, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, md->stream));
  return XGM_OK;
}

}  // namespace

namespace xgm {

// Inference on device arrays (X: n x 16 FP32, Y: n x 5 FP32), used by the
// scheduler without host round trips.  Synchronous.
int mlp_predict_device(xgm_mlp *md, const float *dX, int64_t n, float *dY) {
  if (!md || n < 0) return XGM_EINVAL;
  if (!n) return XGM_OK;
  MLP_TRY(md, hipSetDevice(md->device));
  const int64_t chunk = 1 << 16;
  int rc = ensure_ws(md, std::min(n, chunk));
  if (rc) return rc;
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int B = (int)std::min(chunk, n - r0);
    k_gather<<<blocks_for((int64_t)B * IN), 256, 0, md->stream>>>(B, nullptr, dX + r0 * IN, nullptr, nullptr, md->xb,
                                                                   nullptr, nullptr);
    forward(md, B);
    MLP_TRY(md, hipMemcpyAsync(dY + r0 * OUT, md->y, (size_t)B * OUT * sizeof(float), hipMemcpyDeviceToDevice,
                               md->stream));
  }
  MLP_TRY(md, hipStreamSynchronize(md->stream));
  return XGM_OK;
}

}  // namespace xgm

extern "C" {

int xgm_mlp_create(int device, uint64_t seed, xgm_mlp **out) {
  if (!out) return XGM_EINVAL;
  *out = nullptr;
  std::unique_ptr<xgm_mlp> md(new (std::nothrow) xgm_mlp);
  if (!md) return XGM_ENOMEM;
  md->device = device;
  MLP_TRY(md.get(), hipSetDevice(device));
  MLP_TRY(md.get(), hipStreamCreateWithFlags(&md->stream, hipStreamNonBlocking));
  for (float **q : {&md->p, &md->g, &md->m, &md->v}) MLP_TRY(md.get(), hipMalloc(q, NPARAM * sizeof(float)));
  MLP_TRY(md.get(), hipMalloc(&md->loss, sizeof(float)));
  MLP_TRY(md.get(), hipMalloc(&md->w1, IN * H * sizeof(bf16)));
  MLP_TRY(md.get(), hipMalloc(&md->w2, H * H * sizeof(bf16)));
  MLP_TRY(md.get(), hipMalloc(&md->w3, H * OUT * sizeof(bf16)));
  // He-uniform initial weights from a counter-based stream, zero biases
  std::vector<float> p(NPARAM, 0.f);
  auto fill = [&](int off, int fan_in, int count) {
    const float lim = std::sqrt(6.f / (float)fan_in);
    for (int i = 0; i < count; ++i) {
      const double u = (double)(stream_at(mix64(seed), (uint64_t)(off + i)) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = (float)((2.0 * u - 1.0) * lim);
    }
  };
  fill(OFF_W1, IN, IN * H);
  fill(OFF_W2, H, H * H);
  fill(OFF_W3, H, H * OUT);
  int rc = xgm_mlp_set_params(md.get(), p.data());
  if (rc) return rc;
  *out = md.release();
  return XGM_OK;
}

void xgm_mlp_free(xgm_mlp *md) { delete md; }

const char *xgm_mlp_last_error(const xgm_mlp *md) { return md ? md->err.c_str() : "null model"; }

int xgm_mlp_set_params(xgm_mlp *md, const float *params) {
  if (!md || !params) return XGM_EINVAL;
  MLP_TRY(md, hipSetDevice(md->device));
  MLP_TRY(md, hipMemcpyAsync(md->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, md->stream));
  k_zero_f32<<<blocks_for(NPARAM), 256, 0, md->stream>>>(NPARAM, md->m);
  k_zero_f32<<<blocks_for(NPARAM), 256, 0, md->stream>>>(NPARAM, md->v);
  md->adam_t = 0;
  int rc = refresh_bf16(md);
  if (rc) return rc;
  MLP_TRY(md, hipStreamSynchronize(md->stream));
  return XGM_OK;
}

int xgm_mlp_get_params(const xgm_mlp *md, float *params) {
  if (!md || !params) return XGM_EINVAL;
  xgm_mlp *mm = const_cast<xgm_mlp *>(md);
  MLP_TRY(mm, hipSetDevice(md->device));
  MLP_TRY(mm, hipMemcpyAsync(params, md->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, md->stream));
  MLP_TRY(mm, hipStreamSynchronize(md->stream));
  return XGM_OK;
}

int xgm_mlp_predict(xgm_mlp *md, const float *X, int64_t n, float *Y) {
  if (!md || (n > 0 && (!X || !Y)) || n < 0) return XGM_EINVAL;
  MLP_TRY(md, hipSetDevice(md->device));
  const int64_t chunk = 1 << 16;
  int rc = ensure_ws(md, std::min(n, chunk));
  if (rc) return rc;
  float *dx = nullptr;
  MLP_TRY(md, hipMalloc(&dx, std::min(n, chunk) * IN * sizeof(float) + 4));
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int B = (int)std::min(chunk, n - r0);
    hipError_t e = hipMemcpyAsync(dx, X + r0 * IN, (size_t)B * IN * sizeof(float), hipMemcpyHostToDevice, md->stream);
    if (e != hipSuccess) { (void)hipFree(dx); return md->fail(e, "H2D features"); }
    k_gather<<<blocks_for((int64_t)B * IN), 256, 0, md->stream>>>(B, nullptr, dx, nullptr, nullptr, md->xb, nullptr,
                                                                   nullptr);
    forward(md, B);
    e = hipMemcpyAsync(Y + r0 * OUT, md->y, (size_t)B * OUT * sizeof(float), hipMemcpyDeviceToHost, md->stream);
    if (e != hipSuccess) { (void)hipFree(dx); return md->fail(e, "D2H predictions"); }
  }
  hipError_t e = hipStreamSynchronize(md->stream);
  (void)hipFree(dx);
  if (e != hipSuccess) return md->fail(e, "predict sync");
  return XGM_OK;
}

int xgm_mlp_grads(xgm_mlp *md, const float *X, const float *T, const float *Mk, int64_t n, float *grads,
                  float *loss) {
  if (!md || !X || !T || !Mk || n <= 0 || n > (1 << 20)) return XGM_EINVAL;
  MLP_TRY(md, hipSetDevice(md->device));
  int rc = ensure_ws(md, n);
  if (rc) return rc;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr;
  if ((rc = upload(md, (void **)&dx, X, n * IN * sizeof(float))) ||
      (rc = upload(md, (void **)&dt, T, n * OUT * sizeof(float))) ||
      (rc = upload(md, (void **)&dm, Mk, n * OUT * sizeof(float)))) {
    (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dm);
    return rc;
  }
  double msum = 0;
  for (int64_t i = 0; i < n * OUT; ++i) msum += Mk[i];
  const int B = (int)n;
  k_gather<<<blocks_for((int64_t)B * IN), 256, 0, md->stream>>>(B, nullptr, dx, dt, dm, md->xb, md->tb, md->mb);
  forward(md, B);
  k_zero_f32<<<1, 64, 0, md->stream>>>(1, md->loss);
  k_loss_grad<<<blocks_for((int64_t)B * OUT), 256, 0, md->stream>>>(B, md->y, md->tb, md->mb,
                                                                     (float)(1.0 / std::max(1.0, msum)), md->dy,
                                                                     md->dyb, md->loss);
  backward(md, B);
  hipError_t e = hipMemcpyAsync(grads, md->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost, md->stream);
  if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, md->loss, sizeof(float), hipMemcpyDeviceToHost, md->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(md->stream);
  (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dm);
  if (e != hipSuccess) return md->fail(e, "grads");
  return XGM_OK;
}

int xgm_mlp_train(xgm_mlp *md, const float *X, const float *T, const float *Mk, int64_t n, int epochs, int batch,
                  float lr, uint64_t seed, float *final_loss) {
  if (!md || !X || !T || !Mk || n <= 0 || epochs <= 0 || batch <= 0 || batch > (1 << 20) || !(lr > 0))
    return XGM_EINVAL;
  MLP_TRY(md, hipSetDevice(md->device));
  int rc = ensure_ws(md, batch);
  if (rc) return rc;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr;
  int32_t *didx = nullptr;
  auto cleanup = [&](int r) {
    (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dm); (void)hipFree(didx);
    return r;
  };
  if ((rc = upload(md, (void **)&dx, X, n * IN * sizeof(float))) ||
      (rc = upload(md, (void **)&dt, T, n * OUT * sizeof(float))) ||
      (rc = upload(md, (void **)&dm, Mk, n * OUT * sizeof(float))))
    return cleanup(rc);
  hipError_t e = hipMalloc(&didx, n * sizeof(int32_t));
  if (e != hipSuccess) return cleanup(md->fail(e, "hipMalloc(idx)"));
  // mean mask weight per row, to normalise each batch's loss
  double msum = 0;
  for (int64_t i = 0; i < n * OUT; ++i) msum += Mk[i];
  const double mrow = std::max(1e-9, msum / (double)n);
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float last = 0.f;
  for (int ep = 0; ep < epochs; ++ep) {
    // Fisher-Yates with the counter-based stream: deterministic order
    for (int64_t i = n - 1; i > 0; --i) {
      const int64_t j = (int64_t)bounded_draw(stream_at(mix64(seed + 0x51ULL * (uint64_t)(ep + 1)), (uint64_t)i),
                                              (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    e = hipMemcpyAsync(didx, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, md->stream);
    if (e != hipSuccess) return cleanup(md->fail(e, "H2D perm"));
    k_zero_f32<<<1, 64, 0, md->stream>>>(1, md->loss);
    for (int64_t r0 = 0; r0 < n; r0 += batch) {
      const int B = (int)std::min<int64_t>(batch, n - r0);
      k_gather<<<blocks_for((int64_t)B * IN), 256, 0, md->stream>>>(B, didx + r0, dx, dt, dm, md->xb, md->tb, md->mb);
      forward(md, B);
      const float inv = (float)(1.0 / (mrow * (double)B));
      k_loss_grad<<<blocks_for((int64_t)B * OUT), 256, 0, md->stream>>>(B, md->y, md->tb, md->mb, inv, md->dy,
                                                                         md->dyb, md->loss);
      backward(md, B);
      ++md->adam_t;
      const float c1 = 1.f - std::pow(b1, (float)md->adam_t), c2 = 1.f - std::pow(b2, (float)md->adam_t);
      k_adam<<<blocks_for(NPARAM), 256, 0, md->stream>>>(NPARAM, md->p, md->g, md->m, md->v, lr, b1, b2, eps, c1, c2);
      if ((rc = refresh_bf16(md))) return cleanup(rc);
    }
    e = hipMemcpyAsync(&last, md->loss, sizeof(float), hipMemcpyDeviceToHost, md->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(md->stream);
    if (e != hipSuccess) return cleanup(md->fail(e, "train epoch"));
    last /= (float)((n + batch - 1) / batch);  // mean batch loss of the epoch
  }
  if (final_loss) *final_loss = last;
  return cleanup(XGM_OK);
}

// C[M x N] = A[M x K] * B[K x N] on the matrix cores (BF16 operands
// rounded from the FP32 inputs, FP32 accumulation), for checking the MFMA
// operand layout from the host.
int xgm_mfma_gemm_check(int device, int M, int N, int K, const float *A, const float *B, float *C) {
  if (M <= 0 || N <= 0 || K <= 0 || M > 4096 || N > 4096 || K > 4096 || !A || !B || !C) return XGM_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return XGM_EHIP;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  bf16 *hA = nullptr, *hB = nullptr;
  int rc = XGM_OK;
  auto ok = [&](hipError_t e) { if (e != hipSuccess) rc = XGM_EHIP; return e == hipSuccess; };
  if (ok(hipMalloc(&dA, (size_t)M * K * 4)) && ok(hipMalloc(&dB, (size_t)K * N * 4)) &&
      ok(hipMalloc(&dC, (size_t)M * N * 4)) && ok(hipMalloc(&hA, (size_t)M * K * 2)) &&
      ok(hipMalloc(&hB, (size_t)K * N * 2)) && ok(hipMemcpy(dA, A, (size_t)M * K * 4, hipMemcpyHostToDevice)) &&
      ok(hipMemcpy(dB, B, (size_t)K * N * 4, hipMemcpyHostToDevice))) {
    k_to_bf16<<<blocks_for((int64_t)M * K), 256>>>(M * K, dA, hA);
    k_to_bf16<<<blocks_for((int64_t)K * N), 256>>>(K * N, dB, hB);
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    k_gemm_bf16<<<(tiles + 3) / 4, 256>>>(M, N, K, hA, K, 1, hB, N, 1, dC, N, 0);
    if (ok(hipGetLastError()) && ok(hipDeviceSynchronize()))
      ok(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
  }
  (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(hA); (void)hipFree(hB);
  return rc;
}

}  // extern "C"
