// clr_mlp.hip -- the learned surrogate of the per-shell bias normalisation:
// a 4 -> 64 -> 64 -> 2 perceptron (tanh hidden layers, linear output) with
// its own training kernels (FP32 forward + backward, Adam) and a BF16 MFMA
// inference kernel (FP32 accumulate).
//
// Inputs (raw units) x = (D(z), b(z), sigma_G(z), r_smooth), standardised by
// the network's input scaling; outputs are standardised targets, mapped back
// by the output scaling: y0 = ln norm(z) (the reference's count/sum of
// compute_density_normalization, density.c:1272-1355), y1 = ln <B>, the
// expected count per cell of the un-normalised sampler in units of n dx^3.
//
// Parameter vector (FP32, CLR_MLP_NPARAM = 4610):
//   W1[64][4], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]   (W[out][in])
#include "clr_internal.h"
#include <math.h>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr int kIn = CLR_MLP_IN, kHid = CLR_MLP_HID, kOut = CLR_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kOut * kHid, kNP = oB3 + kOut;
static_assert(kNP == CLR_MLP_NPARAM, "parameter count");
constexpr int kS = 64;           // samples per training block (one per thread)
constexpr int kPad = kHid + 1;   // LDS row stride of per-sample activations

struct Norm {
  float xm[kIn], xs[kIn];   // x_n = (x - xm) / xs
  float ym[kOut], ys[kOut]; // y = ym + ys * y_n
};

// Forward + backward of the mean squared error (1/2n) sum |out - t_n|^2 over
// this block's kS samples; writes the block's gradient (kNP floats) and loss.
__global__ __launch_bounds__(kS) void k_mlp_grad(const float *__restrict__ P, Norm nm, const float *__restrict__ x,
                                                  const float *__restrict__ y, int n, float inv_n,
                                                  float *__restrict__ gpart, double *__restrict__ lpart) {
  __shared__ float w[kNP];
  __shared__ float h1[kS * kPad], h2[kS * kPad], g2[kS * kPad];
  __shared__ float xs[kS * kIn], g3[kS * kOut];
  const int t = threadIdx.x;
  const int s = blockIdx.x * kS + t;
  const bool valid = s < n;
  for (int i = t; i < kNP; i += kS) w[i] = P[i];
  float xn[kIn], tn[kOut];
#pragma unroll
  for (int i = 0; i < kIn; i++) {
    xn[i] = valid ? (x[(size_t)s * kIn + i] - nm.xm[i]) / nm.xs[i] : 0.f;
    xs[t * kIn + i] = xn[i];
  }
#pragma unroll
  for (int k = 0; k < kOut; k++) tn[k] = valid ? (y[(size_t)s * kOut + k] - nm.ym[k]) / nm.ys[k] : 0.f;
  __syncthreads();
  // layer 1
  for (int o = 0; o < kHid; o++) {
    float a = w[oB1 + o];
#pragma unroll
    for (int i = 0; i < kIn; i++) a += w[oW1 + o * kIn + i] * xn[i];
    h1[t * kPad + o] = tanhf(a);
  }
  // layer 2
  for (int o = 0; o < kHid; o++) {
    float a = w[oB2 + o];
    for (int i = 0; i < kHid; i++) a += w[oW2 + o * kHid + i] * h1[t * kPad + i];
    h2[t * kPad + o] = tanhf(a);
  }
  // output, loss, output gradient
  double loss = 0;
  float d3[kOut];
#pragma unroll
  for (int k = 0; k < kOut; k++) {
    float a = w[oB3 + k];
    for (int i = 0; i < kHid; i++) a += w[oW3 + k * kHid + i] * h2[t * kPad + i];
    const float e = valid ? a - tn[k] : 0.f;
    loss += 0.5 * (double)e * (double)e;
    d3[k] = e * inv_n;
    g3[t * kOut + k] = d3[k];
  }
  // back through layer 3 and tanh: g2 = (W3^T d3) (1 - h2^2)
  for (int i = 0; i < kHid; i++) {
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < kOut; k++) a += w[oW3 + k * kHid + i] * d3[k];
    const float hv = h2[t * kPad + i];
    g2[t * kPad + i] = a * (1.f - hv * hv);
  }
  // back through layer 2: g1 = (W2^T g2) (1 - h1^2), written over the
  // thread's own row of h2 (until the barrier every thread reads only its
  // own rows; h2 is recomputed for the W3 gradient below)
  for (int i = 0; i < kHid; i++) {
    float a = 0.f;
    for (int o = 0; o < kHid; o++) a += w[oW2 + o * kHid + i] * g2[t * kPad + o];
    const float hv = h1[t * kPad + i];
    h2[t * kPad + i] = a * (1.f - hv * hv);
  }
  __syncthreads();
  const float *g1 = h2;
  // block sums over the samples, one parameter column per thread (t = input
  // index), fixed order
  float *out = gpart + (size_t)blockIdx.x * kNP;
  // W2[o][t] = sum_s g2[s][o] h1[s][t]; b2[t] = sum_s g2[s][t]
  for (int o = 0; o < kHid; o++) {
    float a = 0.f;
    for (int q = 0; q < kS; q++) a += g2[q * kPad + o] * h1[q * kPad + t];
    out[oW2 + o * kHid + t] = a;
  }
  {
    float a = 0.f, b = 0.f;
    for (int q = 0; q < kS; q++) { a += g2[q * kPad + t]; b += g1[q * kPad + t]; }
    out[oB2 + t] = a;
    out[oB1 + t] = b;
  }
  // W3[k][t] = sum_s g3[s][k] h2orig[s][t]: h2 was overwritten, recompute it
  // from h1 (cheap) for this thread's column t
  {
    float a3[kOut] = {0.f, 0.f};
    for (int q = 0; q < kS; q++) {
      float a = w[oB2 + t];
      for (int i = 0; i < kHid; i++) a += w[oW2 + t * kHid + i] * h1[q * kPad + i];
      const float hv = tanhf(a);
#pragma unroll
      for (int k = 0; k < kOut; k++) a3[k] += g3[q * kOut + k] * hv;
    }
#pragma unroll
    for (int k = 0; k < kOut; k++) out[oW3 + k * kHid + t] = a3[k];
  }
  if (t < kOut) {
    float a = 0.f;
    for (int q = 0; q < kS; q++) a += g3[q * kOut + t];
    out[oB3 + t] = a;
  }
  // W1[t][i] = sum_s g1[s][t] x[s][i]
  {
    float a1[kIn] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < kS; q++) {
      const float gv = g1[q * kPad + t];
#pragma unroll
      for (int i = 0; i < kIn; i++) a1[i] += gv * xs[q * kIn + i];
    }
#pragma unroll
    for (int i = 0; i < kIn; i++) out[oW1 + t * kIn + i] = a1[i];
  }
  // loss: wavefront sum, one value per block
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (t == 0) lpart[blockIdx.x] = loss * (double)inv_n;
}

// Sum of the block gradients in block order (deterministic), then one Adam
// step (Kingma & Ba 2015) with bias correction; grad_out (optional) gets the
// summed gradient.
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ gpart, int nblocks, float lr, float b1, float b2, float eps,
                           float c1, float c2, int apply, float *__restrict__ grad_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  float g = 0.f;
  for (int b = 0; b < nblocks; b++) g += gpart[(size_t)b * kNP + p];
  if (grad_out) grad_out[p] = g;
  if (!apply) return;
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_sum_loss(const double *__restrict__ lpart, int nblocks, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0;
  for (int b = 0; b < nblocks; b++) a += lpart[b];
  *out = a;
}

// FP32 forward, one thread per sample (the reference the BF16 path is checked
// against, and the host-visible FP32 prediction).
__global__ void k_mlp_fwd_f32(const float *__restrict__ P, Norm nm, const float *__restrict__ x, int n,
                              float *__restrict__ y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xn[kIn], h1[kHid];
#pragma unroll
  for (int i = 0; i < kIn; i++) xn[i] = (x[(size_t)s * kIn + i] - nm.xm[i]) / nm.xs[i];
  for (int o = 0; o < kHid; o++) {
    float a = P[oB1 + o];
#pragma unroll
    for (int i = 0; i < kIn; i++) a += P[oW1 + o * kIn + i] * xn[i];
    h1[o] = tanhf(a);
  }
  float o3[kOut];
#pragma unroll
  for (int k = 0; k < kOut; k++) o3[k] = P[oB3 + k];
  for (int o = 0; o < kHid; o++) {
    float a = P[oB2 + o];
    for (int i = 0; i < kHid; i++) a += P[oW2 + o * kHid + i] * h1[i];
    const float hv = tanhf(a);
#pragma unroll
    for (int k = 0; k < kOut; k++) o3[k] += P[oW3 + k * kHid + o] * hv;
  }
#pragma unroll
  for (int k = 0; k < kOut; k++) y[(size_t)s * kOut + k] = nm.ym[k] + nm.ys[k] * o3[k];
}

// ---------------------------------------------------------------------------
// BF16 MFMA inference.  v_mfma_f32_16x16x16_bf16: per wavefront a 16x16 tile
// D = A B + C with A 16x16 (rows = samples, K), B 16x16 (K, cols), lane l
// holding A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16] (4 bf16 each) and
// D[4(l/16)..+3][l%16] (4 floats).  One wavefront handles 16 samples:
//   layer 1: K = 4 inputs padded to 16, 4 column tiles of 16 hidden units;
//   layer 2: K = 64 (4 k-tiles) x 4 column tiles;
//   layer 3: K = 64 x 1 column tile (2 of 16 columns used).
// Activations go through LDS between layers (the D layout is the transpose
// of what the next A operand needs); bias and tanh in FP32.
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (unsigned short)(u >> 16);
}

// B-operand packs: for tile (kt, ct) and lane l, 4 bf16 = B[16kt + 4(l/16) + i][16ct + l%16].
//   W1B: [4 ct][64 lanes] (K padded to 16), W2B: [4 kt][4 ct][64], W3B: [4 kt][64]
constexpr int kW1B = 4 * 64, kW2B = 16 * 64, kW3B = 4 * 64;
__global__ void k_mlp_pack_bf16(const float *__restrict__ P, v4s *__restrict__ pk) {
  const int l = threadIdx.x;   // 64 threads, one block
  const int col = l % 16, k0 = 4 * (l / 16);
  for (int ct = 0; ct < 4; ct++) {
    v4s v;
    for (int i = 0; i < 4; i++) {
      const int k = k0 + i, o = 16 * ct + col;
      v[i] = (short)(k < kIn ? f2bf(P[oW1 + o * kIn + k]) : 0);
    }
    pk[ct * 64 + l] = v;
  }
  for (int kt = 0; kt < 4; kt++)
    for (int ct = 0; ct < 4; ct++) {
      v4s v;
      for (int i = 0; i < 4; i++) {
        const int k = 16 * kt + k0 + i, o = 16 * ct + col;
        v[i] = (short)f2bf(P[oW2 + o * kHid + k]);
      }
      pk[kW1B + (kt * 4 + ct) * 64 + l] = v;
    }
  for (int kt = 0; kt < 4; kt++) {
    v4s v;
    for (int i = 0; i < 4; i++) {
      const int k = 16 * kt + k0 + i;
      v[i] = (short)(col < kOut ? f2bf(P[oW3 + col * kHid + k]) : 0);
    }
    pk[kW1B + kW2B + kt * 64 + l] = v;
  }
}

// One wavefront per block, so that the block barriers below are wavefront
// barriers and every lane of a block leaves together.
__global__ __launch_bounds__(64) void k_mlp_infer_bf16(const v4s *__restrict__ pk, const float *__restrict__ P,
                                                       Norm nm, const float *__restrict__ x, int n,
                                                       float *__restrict__ y) {
  __shared__ float A[16 * kPad];
  const int l = threadIdx.x;
  const int row0 = blockIdx.x * 16;
  const int r = l % 16, k0 = 4 * (l / 16);
  // layer 1: A = standardised inputs (K padded to 16)
  v4s a1;
  for (int i = 0; i < 4; i++) {
    const int k = k0 + i, s = row0 + r;
    const float v = (k < kIn && s < n) ? (x[(size_t)s * kIn + k] - nm.xm[k]) / nm.xs[k] : 0.f;
    a1[i] = (short)f2bf(v);
  }
  for (int ct = 0; ct < 4; ct++) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, pk[ct * 64 + l], acc, 0, 0, 0);
    const int o = 16 * ct + r;
    for (int i = 0; i < 4; i++) A[(k0 + i) * kPad + o] = tanhf(acc[i] + P[oB1 + o]);
  }
  __syncthreads();
  // layer 2
  v4s a2[4];
  for (int kt = 0; kt < 4; kt++)
    for (int i = 0; i < 4; i++) a2[kt][i] = (short)f2bf(A[r * kPad + 16 * kt + k0 + i]);
  __syncthreads();
  for (int ct = 0; ct < 4; ct++) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; kt++)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[kt], pk[kW1B + (kt * 4 + ct) * 64 + l], acc, 0, 0, 0);
    const int o = 16 * ct + r;
    for (int i = 0; i < 4; i++) A[(k0 + i) * kPad + o] = tanhf(acc[i] + P[oB2 + o]);
  }
  __syncthreads();
  // layer 3
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int kt = 0; kt < 4; kt++) {
    v4s a3;
    for (int i = 0; i < 4; i++) a3[i] = (short)f2bf(A[r * kPad + 16 * kt + k0 + i]);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, pk[kW1B + kW2B + kt * 64 + l], acc, 0, 0, 0);
  }
  if (r < kOut)
    for (int i = 0; i < 4; i++) {
      const int s = row0 + k0 + i;
      if (s < n) y[(size_t)s * kOut + r] = nm.ym[r] + nm.ys[r] * (acc[i] + P[oB3 + r]);
    }
}

uint64_t splitmix64(uint64_t &z) {
  uint64_t r = (z += 0x9E3779B97F4A7C15ull);
  r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
  r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
  return r ^ (r >> 31);
}

}  // namespace

struct clr_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *P = nullptr, *m = nullptr, *v = nullptr;   // parameters and Adam moments
  v4s *pk = nullptr;                                // BF16 B-operand packs
  int packed = 0;
  int steps_done = 0;
  Norm nm{};
  std::string err;
};

static int mlp_fail(clr_mlp *n, int code, const char *msg, hipError_t e = hipSuccess) {
  if (n) n->err = e == hipSuccess ? std::string(msg) : std::string(msg) + ": " + hipGetErrorString(e);
  return code;
}
#define MLP_HIP(net, call)                                                          \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return mlp_fail((net), CLR_ERR_HIP, #call, e_);           \
  } while (0)

// A batch on the device: x (n x 4) and y (n x 2) floats.
struct DevBatch {
  float *x = nullptr, *y = nullptr;
  float *gpart = nullptr;
  double *lpart = nullptr, *loss = nullptr;
  int nb = 0;
  ~DevBatch() {
    for (void *p : {(void *)x, (void *)y, (void *)gpart, (void *)lpart, (void *)loss})
      if (p) (void)hipFree(p);
  }
};

static int upload_batch(clr_mlp *net, const float *x, const float *y, int n, DevBatch &b) {
  b.nb = (n + kS - 1) / kS;
  MLP_HIP(net, hipMalloc((void **)&b.x, (size_t)n * kIn * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&b.y, (size_t)n * kOut * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&b.gpart, (size_t)b.nb * kNP * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&b.lpart, (size_t)b.nb * sizeof(double)));
  MLP_HIP(net, hipMalloc((void **)&b.loss, sizeof(double)));
  MLP_HIP(net, hipMemcpyAsync(b.x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, net->stream));
  MLP_HIP(net, hipMemcpyAsync(b.y, y, (size_t)n * kOut * sizeof(float), hipMemcpyHostToDevice, net->stream));
  return CLR_OK;
}

extern "C" {

int clr_mlp_create(clr_mlp **out, int device, uint32_t seed) {
  if (!out) return CLR_ERR_ARG;
  clr_mlp *net = new clr_mlp();
  *out = net;
  net->device = device;
  MLP_HIP(net, hipSetDevice(device));
  MLP_HIP(net, hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
  MLP_HIP(net, hipMalloc((void **)&net->P, kNP * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&net->m, kNP * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&net->v, kNP * sizeof(float)));
  MLP_HIP(net, hipMalloc((void **)&net->pk, (kW1B + kW2B + kW3B) * sizeof(v4s)));
  // Glorot-uniform weights, zero biases
  std::vector<float> p(kNP, 0.f);
  uint64_t z = 0x636F6C6F7265ull ^ ((uint64_t)seed << 20);
  auto fill = [&](int off, int nout, int nin) {
    const double a = sqrt(6.0 / (nin + nout));
    for (int i = 0; i < nout * nin; i++) p[off + i] = (float)(a * (2.0 * (splitmix64(z) >> 11) * 0x1.0p-53 - 1.0));
  };
  fill(oW1, kHid, kIn);
  fill(oW2, kHid, kHid);
  fill(oW3, kOut, kHid);
  MLP_HIP(net, hipMemcpy(net->P, p.data(), kNP * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(net, hipMemset(net->m, 0, kNP * sizeof(float)));
  MLP_HIP(net, hipMemset(net->v, 0, kNP * sizeof(float)));
  for (int i = 0; i < kIn; i++) { net->nm.xm[i] = 0.f; net->nm.xs[i] = 1.f; }
  for (int k = 0; k < kOut; k++) { net->nm.ym[k] = 0.f; net->nm.ys[k] = 1.f; }
  return CLR_OK;
}

void clr_mlp_destroy(clr_mlp *net) {
  if (!net) return;
  (void)hipSetDevice(net->device);
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  for (void *q : {(void *)net->P, (void *)net->m, (void *)net->v, (void *)net->pk})
    if (q) (void)hipFree(q);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

const char *clr_mlp_last_error(const clr_mlp *net) { return net ? net->err.c_str() : "null network"; }

int clr_mlp_set_scaling(clr_mlp *net, const float *x_mean, const float *x_scale, const float *y_mean,
                        const float *y_scale) {
  if (!net || !x_mean || !x_scale || !y_mean || !y_scale) return CLR_ERR_ARG;
  for (int i = 0; i < kIn; i++) {
    if (!(x_scale[i] > 0.f)) return mlp_fail(net, CLR_ERR_ARG, "input scales must be positive");
    net->nm.xm[i] = x_mean[i];
    net->nm.xs[i] = x_scale[i];
  }
  for (int k = 0; k < kOut; k++) {
    if (!(y_scale[k] > 0.f)) return mlp_fail(net, CLR_ERR_ARG, "output scales must be positive");
    net->nm.ym[k] = y_mean[k];
    net->nm.ys[k] = y_scale[k];
  }
  return CLR_OK;
}

int clr_mlp_set_params(clr_mlp *net, const float *params) {
  if (!net || !params) return CLR_ERR_ARG;
  MLP_HIP(net, hipSetDevice(net->device));
  MLP_HIP(net, hipMemcpy(net->P, params, kNP * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(net, hipMemset(net->m, 0, kNP * sizeof(float)));
  MLP_HIP(net, hipMemset(net->v, 0, kNP * sizeof(float)));
  net->steps_done = 0;
  net->packed = 0;
  return CLR_OK;
}

int clr_mlp_get_params(clr_mlp *net, float *params) {
  if (!net || !params) return CLR_ERR_ARG;
  MLP_HIP(net, hipSetDevice(net->device));
  MLP_HIP(net, hipStreamSynchronize(net->stream));
  MLP_HIP(net, hipMemcpy(params, net->P, kNP * sizeof(float), hipMemcpyDeviceToHost));
  return CLR_OK;
}

int clr_mlp_grad(clr_mlp *net, const float *x, const float *y, int n, float *grad, double *loss) {
  if (!net || !x || !y || n < 1) return CLR_ERR_ARG;
  MLP_HIP(net, hipSetDevice(net->device));
  DevBatch b;
  int rc = upload_batch(net, x, y, n, b);
  if (rc) return rc;
  float *dg = nullptr;
  MLP_HIP(net, hipMalloc((void **)&dg, kNP * sizeof(float)));
  hipLaunchKernelGGL(k_mlp_grad, dim3(b.nb), dim3(kS), 0, net->stream, net->P, net->nm, b.x, b.y, n, 1.f / n,
                     b.gpart, b.lpart);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + 255) / 256), dim3(256), 0, net->stream, net->P, net->m, net->v,
                     b.gpart, b.nb, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0, dg);
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, net->stream, b.lpart, b.nb, b.loss);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, dg, kNP * sizeof(float), hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, b.loss, sizeof(double), hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  (void)hipFree(dg);
  if (e != hipSuccess) return mlp_fail(net, CLR_ERR_HIP, "gradient", e);
  return CLR_OK;
}

int clr_mlp_train(clr_mlp *net, const float *x, const float *y, int n, int steps, float lr, float lr_final,
                  double *loss_out) {
  if (!net || !x || !y || n < 1 || steps < 0 || !(lr > 0.f) || !(lr_final > 0.f)) return CLR_ERR_ARG;
  MLP_HIP(net, hipSetDevice(net->device));
  DevBatch b;
  int rc = upload_batch(net, x, y, n, b);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; it++) {
    // learning rate decays geometrically from lr to lr_final
    const float lrt = lr * powf(lr_final / lr, steps > 1 ? (float)it / (float)(steps - 1) : 0.f);
    const int t = ++net->steps_done;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_mlp_grad, dim3(b.nb), dim3(kS), 0, net->stream, net->P, net->nm, b.x, b.y, n, 1.f / n,
                       b.gpart, b.lpart);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + 255) / 256), dim3(256), 0, net->stream, net->P, net->m, net->v,
                       b.gpart, b.nb, lrt, b1, b2, eps, c1, c2, 1, (float *)nullptr);
    if (loss_out) {
      hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, net->stream, b.lpart, b.nb, b.loss);
      MLP_HIP(net, hipMemcpyAsync(&loss_out[it], b.loss, sizeof(double), hipMemcpyDeviceToHost, net->stream));
    }
  }
  MLP_HIP(net, hipGetLastError());
  MLP_HIP(net, hipStreamSynchronize(net->stream));
  net->packed = 0;
  return CLR_OK;
}

int clr_mlp_predict(clr_mlp *net, const float *x, int n, float *y, int precision) {
  if (!net || !x || !y || n < 0) return CLR_ERR_ARG;
  if (precision != CLR_MLP_FP32 && precision != CLR_MLP_BF16_MFMA)
    return mlp_fail(net, CLR_ERR_ARG, "precision must be CLR_MLP_FP32 or CLR_MLP_BF16_MFMA");
  if (n == 0) return CLR_OK;
  MLP_HIP(net, hipSetDevice(net->device));
  float *dx = nullptr, *dy = nullptr;
  hipError_t e = hipMalloc((void **)&dx, (size_t)n * kIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&dy, (size_t)n * kOut * sizeof(float));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dx, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, net->stream);
  if (e == hipSuccess) {
    e = clr_mlp_predict_device(net, dx, n, dy, precision, net->stream) ? hipErrorUnknown : hipSuccess;
  }
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, (size_t)n * kOut * sizeof(float), hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) return mlp_fail(net, CLR_ERR_HIP, "prediction", e);
  return CLR_OK;
}

}  // extern "C"

// Device-side prediction on device buffers (x: n x 4, y: n x 2) on `stream`
// (ordered after the network's own stream).
int clr_mlp_predict_device(clr_mlp *net, const float *dx, int n, float *dy, int precision, hipStream_t stream) {
  if (n <= 0) return CLR_OK;
  if (stream != net->stream) {
    // order after training on the network's stream
    MLP_HIP(net, hipStreamSynchronize(net->stream));
  }
  if (precision == CLR_MLP_BF16_MFMA) {
    if (!net->packed) {
      hipLaunchKernelGGL(k_mlp_pack_bf16, dim3(1), dim3(64), 0, stream, net->P, net->pk);
      net->packed = 1;
    }
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, stream, net->pk, net->P,
                       net->nm, dx, n, dy);
  } else {
    hipLaunchKernelGGL(k_mlp_fwd_f32, dim3((n + 127) / 128), dim3(128), 0, stream, net->P, net->nm, dx, n, dy);
  }
  MLP_HIP(net, hipGetLastError());
  return CLR_OK;
}
