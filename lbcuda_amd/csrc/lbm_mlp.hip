/*
 * lbm_mlp.hip -- the near-contact closure network: a 4 -> 64 -> 64 -> 3 MLP
 * (tanh hidden layers, linear output) in FP32, with forward, backward and
 * Adam kernels for training on the GPU and a device inference routine the
 * pair-force kernel calls.  No reference counterpart: the reference's only
 * near-contact model is the lubrication term of addHertzian
 * (kernels_fluid_PART.CUF:1463-1501), which stays as the fallback.
 *
 * Parameter vector (LBM_MLP_NPARAM floats):
 *   W1[64][4], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]   (row = output)
 *
 * Training pass: a workgroup of 64 lanes takes 64 samples, one per lane, and
 * keeps their activations and back-propagated errors in LDS; each lane then
 * forms 1/64 of the parameter gradients as dot products over the 64 samples.
 * Workgroup partial gradients are summed over workgroups in a fixed order by
 * a second kernel, so gradients do not depend on scheduling.
 */
#include "lbm_common.h"

#include <math.h>

#include <vector>

namespace {

constexpr int NI = 4, NH = 64, NO = 3;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NO * NH, NP = OB3 + NO;
static_assert(NP == LBM_MLP_NPARAM, "parameter count");
constexpr int SB = 64;   /* samples per workgroup */

__global__ __launch_bounds__(SB) void k_mlp_grad(const float* __restrict__ prm, const float* __restrict__ x,
                                                 const float* __restrict__ t, const int* __restrict__ idx, int n,
                                                 float scale, float* __restrict__ part, double* __restrict__ lpart) {
  __shared__ float sx[SB][NI];
  __shared__ float sa1[SB][NH + 1];
  __shared__ float sa2[SB][NH + 1];
  __shared__ float sd1[SB][NH + 1];
  __shared__ float sd2[SB][NH + 1];
  __shared__ float sdy[SB][NO];
  __shared__ double sl[SB];
  const int lane = threadIdx.x;
  const int s = blockIdx.x * SB + lane;
  const bool valid = s < n;
  const int si = valid ? (idx ? idx[s] : s) : 0;
  float xi[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) xi[i] = valid ? x[NI * si + i] : 0.0f;
  /* forward */
  float a1[NH];
#pragma unroll 8
  for (int j = 0; j < NH; ++j) {
    float z = prm[OB1 + j];
#pragma unroll
    for (int i = 0; i < NI; ++i) z = fmaf(prm[OW1 + NI * j + i], xi[i], z);
    a1[j] = tanhf(z);
  }
  float a2[NH];
#pragma unroll 4
  for (int k = 0; k < NH; ++k) {
    float z = prm[OB2 + k];
#pragma unroll 16
    for (int j = 0; j < NH; ++j) z = fmaf(prm[OW2 + NH * k + j], a1[j], z);
    a2[k] = tanhf(z);
  }
  float dy[NO];
  double loss = 0.0;
#pragma unroll
  for (int m = 0; m < NO; ++m) {
    float y = prm[OB3 + m];
#pragma unroll 16
    for (int k = 0; k < NH; ++k) y = fmaf(prm[OW3 + NH * m + k], a2[k], y);
    const float e = valid ? y - t[NO * si + m] : 0.0f;
    loss += (double)e * (double)e;
    dy[m] = scale * 2.0f * e;   /* d(loss)/dy for loss = scale * sum e^2 */
  }
  /* backward through the hidden layers */
  float d2[NH];
#pragma unroll 4
  for (int k = 0; k < NH; ++k) {
    float g = 0.0f;
#pragma unroll
    for (int m = 0; m < NO; ++m) g = fmaf(prm[OW3 + NH * m + k], dy[m], g);
    d2[k] = g * (1.0f - a2[k] * a2[k]);
  }
#pragma unroll 4
  for (int j = 0; j < NH; ++j) {
    float g = 0.0f;
#pragma unroll 16
    for (int k = 0; k < NH; ++k) g = fmaf(prm[OW2 + NH * k + j], d2[k], g);
    sd1[lane][j] = g * (1.0f - a1[j] * a1[j]);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) sx[lane][i] = xi[i];
#pragma unroll
  for (int j = 0; j < NH; ++j) {
    sa1[lane][j] = a1[j];
    sa2[lane][j] = a2[j];
    sd2[lane][j] = d2[j];
  }
#pragma unroll
  for (int m = 0; m < NO; ++m) sdy[lane][m] = dy[m];
  sl[lane] = loss;
  __syncthreads();
  /* parameter gradients: sums over the 64 samples of this workgroup */
  float* out = part + (size_t)blockIdx.x * NP;
  for (int p = lane; p < NP; p += SB) {
    float g = 0.0f;
    if (p < OB1) {
      const int j = (p - OW1) / NI, i = (p - OW1) % NI;
      for (int q = 0; q < SB; ++q) g = fmaf(sd1[q][j], sx[q][i], g);
    } else if (p < OW2) {
      const int j = p - OB1;
      for (int q = 0; q < SB; ++q) g += sd1[q][j];
    } else if (p < OB2) {
      const int k = (p - OW2) / NH, j = (p - OW2) % NH;
      for (int q = 0; q < SB; ++q) g = fmaf(sd2[q][k], sa1[q][j], g);
    } else if (p < OW3) {
      const int k = p - OB2;
      for (int q = 0; q < SB; ++q) g += sd2[q][k];
    } else if (p < OB3) {
      const int m = (p - OW3) / NH, k = (p - OW3) % NH;
      for (int q = 0; q < SB; ++q) g = fmaf(sdy[q][m], sa2[q][k], g);
    } else {
      const int m = p - OB3;
      for (int q = 0; q < SB; ++q) g += sdy[q][m];
    }
    out[p] = g;
  }
  if (lane == 0) {
    double l = 0.0;
    for (int q = 0; q < SB; ++q) l += sl[q];
    lpart[blockIdx.x] = l;
  }
}

/* grad[p] = sum over workgroups b (in order) of part[b][p] */
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ grad,
                                                    const double* __restrict__ lpart, double* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NP) {
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += part[(size_t)b * NP + p];
    grad[p] = g;
  }
  if (p == 0 && loss) {
    double l = 0.0;
    for (int b = 0; b < nblk; ++b) l += lpart[b];
    *loss = l;
  }
}

/* Adam (Kingma & Ba 2015) with bias correction */
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ prm, const float* __restrict__ grad,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const float g = grad[p];
  const float a = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  prm[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ prm, const float* __restrict__ x,
                                                     float* __restrict__ y, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float in[NI], out[NO];
  for (int i = 0; i < NI; ++i) in[i] = x[NI * s + i];
  lbm_mlp_eval(prm, in, out);
  for (int m = 0; m < NO; ++m) y[NO * s + m] = out[m];
}

int ok(hipError_t e) { return e == hipSuccess ? 0 : LBM_ERR_HIP; }

template <class T>
int upload(T** d, const T* h, size_t count) {
  *d = nullptr;
  if (hipMalloc((void**)d, sizeof(T) * (count ? count : 1)) != hipSuccess) return LBM_ERR_ALLOC;
  if (count && h && hipMemcpy(*d, h, sizeof(T) * count, hipMemcpyHostToDevice) != hipSuccess) return LBM_ERR_HIP;
  return 0;
}

/* loss and gradient of one batch (idx = sample indices or NULL for 0..n-1) */
int grad_batch(const float* d_prm, const float* d_x, const float* d_t, const int* d_idx, int n, float scale,
               float* d_part, double* d_lpart, float* d_grad, double* d_loss, hipStream_t st) {
  const int nblk = (n + SB - 1) / SB;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(SB), 0, st, d_prm, d_x, d_t, d_idx, n, scale, d_part, d_lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, st, d_part, nblk, d_grad, d_lpart, d_loss);
  return ok(hipGetLastError());
}

}  // namespace

extern "C" {

int lbm_mlp_forward(const float* params, const float* x, float* y, int32_t n) {
  if (!params || !x || !y || n <= 0) return LBM_ERR_ARG;
  float *dp = nullptr, *dx = nullptr, *dy = nullptr;
  int rc = upload(&dp, params, NP);
  rc |= upload(&dx, x, (size_t)NI * n);
  rc |= upload(&dy, (const float*)nullptr, (size_t)NO * n);
  if (!rc) {
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, 0, dp, dx, dy, n);
    rc |= ok(hipGetLastError());
    rc |= ok(hipMemcpy(y, dy, sizeof(float) * NO * n, hipMemcpyDeviceToHost));
  }
  (void)hipFree(dp);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

int lbm_mlp_loss_grad(const float* params, const float* x, const float* t, int32_t n, double* loss, float* grad) {
  if (!params || !x || !t || n <= 0) return LBM_ERR_ARG;
  float *dp = nullptr, *dx = nullptr, *dt = nullptr, *dpart = nullptr, *dg = nullptr;
  double *dl = nullptr, *dlp = nullptr;
  const int nblk = (n + SB - 1) / SB;
  int rc = upload(&dp, params, NP);
  rc |= upload(&dx, x, (size_t)NI * n);
  rc |= upload(&dt, t, (size_t)NO * n);
  rc |= upload(&dpart, (const float*)nullptr, (size_t)nblk * NP);
  rc |= upload(&dg, (const float*)nullptr, NP);
  rc |= upload(&dl, (const double*)nullptr, 1);
  rc |= upload(&dlp, (const double*)nullptr, nblk);
  const float scale = 1.0f / (float)(NO * n);   /* mean squared error */
  if (!rc) rc |= grad_batch(dp, dx, dt, nullptr, n, scale, dpart, dlp, dg, dl, 0);
  if (!rc && grad) rc |= ok(hipMemcpy(grad, dg, sizeof(float) * NP, hipMemcpyDeviceToHost));
  if (!rc && loss) {
    double l = 0.0;
    rc |= ok(hipMemcpy(&l, dl, sizeof(double), hipMemcpyDeviceToHost));
    *loss = l / (double)(NO * n);
  }
  (void)hipFree(dp); (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dpart);
  (void)hipFree(dg); (void)hipFree(dl); (void)hipFree(dlp);
  return rc;
}

int lbm_mlp_train(float* params, const float* x, const float* t, int32_t n, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, double* final_loss) {
  if (!params || !x || !t || n <= 0 || epochs < 0 || batch <= 0 || !(lr > 0.0f)) return LBM_ERR_ARG;
  if (batch > n) batch = n;
  const int nblk = (batch + SB - 1) / SB;
  float *dp = nullptr, *dx = nullptr, *dt = nullptr, *dpart = nullptr, *dg = nullptr, *m1 = nullptr, *m2 = nullptr;
  double *dl = nullptr, *dlp = nullptr;
  int* didx = nullptr;
  int rc = upload(&dp, params, NP);
  rc |= upload(&dx, x, (size_t)NI * n);
  rc |= upload(&dt, t, (size_t)NO * n);
  rc |= upload(&dpart, (const float*)nullptr, (size_t)nblk * NP);
  rc |= upload(&dg, (const float*)nullptr, NP);
  std::vector<float> zeros(NP, 0.0f);
  rc |= upload(&m1, zeros.data(), NP);
  rc |= upload(&m2, zeros.data(), NP);
  rc |= upload(&dl, (const double*)nullptr, 1);
  rc |= upload(&dlp, (const double*)nullptr, nblk);
  rc |= upload(&didx, (const int*)nullptr, (size_t)n);
  hipStream_t st = nullptr;
  if (!rc) rc |= ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  /* each epoch visits the samples in a seeded random order (Fisher-Yates on
     a SplitMix64 stream), in batches of `batch` */
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long it = 0;
  for (int e = 0; e < epochs && !rc; ++e) {
    for (int i = n - 1; i > 0; --i) {
      const int j = (int)(next() % (uint64_t)(i + 1));
      const int tmp = order[i]; order[i] = order[j]; order[j] = tmp;
    }
    rc |= ok(hipMemcpyAsync(didx, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    for (int b0 = 0; b0 < n && !rc; b0 += batch) {
      const int nb = b0 + batch <= n ? batch : n - b0;
      rc |= grad_batch(dp, dx, dt, didx + b0, nb, 1.0f / (float)(NO * nb), dpart, dlp, dg, nullptr, st);
      ++it;
      const float c1 = 1.0f - powf(b1, (float)it), c2 = 1.0f - powf(b2, (float)it);
      hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, st, dp, dg, m1, m2, lr, b1, b2, eps, c1, c2);
      rc |= ok(hipGetLastError());
    }
    rc |= ok(hipStreamSynchronize(st));
  }
  if (!rc) {
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    rc |= ok(hipMemcpy(didx, all.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  }
  if (!rc) rc |= ok(hipMemcpy(params, dp, sizeof(float) * NP, hipMemcpyDeviceToHost));
  if (!rc && final_loss) {
    double l = 0.0;
    rc |= lbm_mlp_loss_grad(params, x, t, n, &l, nullptr);
    *final_loss = l;
  }
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(dp); (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dpart); (void)hipFree(dg);
  (void)hipFree(m1); (void)hipFree(m2); (void)hipFree(dl); (void)hipFree(dlp); (void)hipFree(didx);
  return rc;
}

}  /* extern "C" */
