// Small MLP surrogate trained and evaluated on the GPU: 5 inputs -> 64 SiLU -> 64 SiLU -> 3 outputs, FP32.
//
// It maps (T, A, W, field, log N) of a lattice to (S, E, chi) of the plain Metropolis method and is used only to
// choose where the plain method samples (temperature points near the chi peak) and to warm-start directors;
// every reported number still comes from the Metropolis runs.
//
// Kernels:
//   k_mlp_forward  one thread per sample, weights read through the cache, activations in registers;
//   k_mlp_grad     32 samples per workgroup: forward and backward of each sample by one lane into LDS, then the
//                  weight gradients of the workgroup as sums over its samples (every lane owns a slice of the
//                  parameters), one partial gradient per workgroup;
//   k_mlp_adam     fixed-order sum of the partial gradients and the Adam update (Kingma & Ba, 2015), one thread
//                  per parameter; the result does not depend on scheduling.
// Parameter layout: W1[64][5], b1[64], W2[64][64], b2[64], W3[3][64], b3[3] (4739 floats).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mclgpu.h"
#include "philox.h"

namespace {

constexpr int NI = 5, NH = 64, NO = 3;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NO * NH, NP = OB3 + NO;
constexpr int GS = 32;      // samples per gradient workgroup
constexpr int GT = 64;      // threads per gradient workgroup
constexpr int LP = NH + 1;  // padded LDS row (no bank conflicts on column walks)

thread_local std::string g_err;
int ferr(int code, const std::string &m) {
  g_err = m;
  return code;
}

#define MLP_HIP(call)                                                                                        \
  do {                                                                                                       \
    hipError_t e_ = (call);                                                                                  \
    if (e_ != hipSuccess) return ferr(MCL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, int n,
                                                    float *__restrict__ Y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NI], h1[NH];
  for (int i = 0; i < NI; i++) x[i] = X[(size_t)s * NI + i];
#pragma unroll 4
  for (int u = 0; u < NH; u++) {
    float z = P[OB1 + u];
    for (int i = 0; i < NI; i++) z += P[OW1 + u * NI + i] * x[i];
    h1[u] = z * sigm(z);
  }
  float y[NO];
  for (int o = 0; o < NO; o++) y[o] = P[OB3 + o];
  for (int u = 0; u < NH; u++) {
    float z = P[OB2 + u];
#pragma unroll 8
    for (int v = 0; v < NH; v++) z += P[OW2 + u * NH + v] * h1[v];
    const float h2 = z * sigm(z);
    for (int o = 0; o < NO; o++) y[o] += P[OW3 + o * NH + u] * h2;
  }
  for (int o = 0; o < NO; o++) Y[(size_t)s * NO + o] = y[o];
}

// Partial gradient of L = sum_s sum_o w_o (y_so - t_so)^2 * scale over the samples idx[b*GS .. b*GS + GS) (or the
// samples b*GS.. when idx is null); loss partial in lossp[b].
__global__ __launch_bounds__(GT) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                const float *__restrict__ Tg, const int *__restrict__ idx, int n,
                                                float w0, float w1, float w2, float scale,
                                                float *__restrict__ gpart, float *__restrict__ lossp) {
  __shared__ float xs[GS][NI];
  __shared__ float h1[GS][LP], a1[GS][LP];  // SiLU output and derivative, layer 1
  __shared__ float h2[GS][LP], a2[GS][LP];  // layer 2
  __shared__ float d2[GS][LP], d1[GS][LP];  // dL/dz2, dL/dz1
  __shared__ float dy[GS][NO];
  __shared__ float lsum[GT];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int first = b * GS;
  const float wo[NO] = {w0, w1, w2};
  float loss = 0.0f;
  if (tid < GS) {
    const int k = first + tid;
    const bool live = k < n;
    const int s = live ? (idx ? idx[k] : k) : 0;
    float x[NI];
    for (int i = 0; i < NI; i++) {
      x[i] = live ? X[(size_t)s * NI + i] : 0.0f;
      xs[tid][i] = x[i];
    }
    for (int u = 0; u < NH; u++) {
      float z = P[OB1 + u];
      for (int i = 0; i < NI; i++) z += P[OW1 + u * NI + i] * x[i];
      const float sg = sigm(z);
      h1[tid][u] = z * sg;
      a1[tid][u] = sg * (1.0f + z * (1.0f - sg));
    }
    float y[NO];
    for (int o = 0; o < NO; o++) y[o] = P[OB3 + o];
    for (int u = 0; u < NH; u++) {
      float z = P[OB2 + u];
      for (int v = 0; v < NH; v++) z += P[OW2 + u * NH + v] * h1[tid][v];
      const float sg = sigm(z);
      const float h = z * sg;
      h2[tid][u] = h;
      a2[tid][u] = sg * (1.0f + z * (1.0f - sg));
      for (int o = 0; o < NO; o++) y[o] += P[OW3 + o * NH + u] * h;
    }
    for (int o = 0; o < NO; o++) {
      const float r = live ? y[o] - Tg[(size_t)s * NO + o] : 0.0f;
      loss += wo[o] * r * r * scale;
      dy[tid][o] = 2.0f * wo[o] * r * scale;
    }
    // back through layer 3 and the SiLU of layer 2
    for (int u = 0; u < NH; u++) {
      float g = 0.0f;
      for (int o = 0; o < NO; o++) g += P[OW3 + o * NH + u] * dy[tid][o];
      d2[tid][u] = g * a2[tid][u];
    }
    // back through layer 2 and the SiLU of layer 1
    for (int v = 0; v < NH; v++) {
      float g = 0.0f;
      for (int u = 0; u < NH; u++) g += P[OW2 + u * NH + v] * d2[tid][u];
      d1[tid][v] = g * a1[tid][v];
    }
  }
  lsum[tid] = loss;
  __syncthreads();
  // weight gradients: sums over the workgroup's samples, parameters split over the lanes
  float *gp = gpart + (size_t)b * NP;
  for (int p = tid; p < NP; p += GT) {
    float g = 0.0f;
    if (p < OB1) {
      const int u = p / NI, i = p % NI;
      for (int q = 0; q < GS; q++) g += d1[q][u] * xs[q][i];
    } else if (p < OW2) {
      const int u = p - OB1;
      for (int q = 0; q < GS; q++) g += d1[q][u];
    } else if (p < OB2) {
      const int u = (p - OW2) / NH, v = (p - OW2) % NH;
      for (int q = 0; q < GS; q++) g += d2[q][u] * h1[q][v];
    } else if (p < OW3) {
      const int u = p - OB2;
      for (int q = 0; q < GS; q++) g += d2[q][u];
    } else if (p < OB3) {
      const int o = (p - OW3) / NH, u = (p - OW3) % NH;
      for (int q = 0; q < GS; q++) g += dy[q][o] * h2[q][u];
    } else {
      const int o = p - OB3;
      for (int q = 0; q < GS; q++) g += dy[q][o];
    }
    gp[p] = g;
  }
  if (tid == 0) {
    float l = 0.0f;
    for (int q = 0; q < GT; q++) l += lsum[q];
    lossp[b] = l;
  }
}

// Sum of the partial gradients (fixed order) and, when adam != 0, the Adam update.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                                 const float *__restrict__ gpart, int nblk, float *__restrict__ gout,
                                                 int adam, float lr, float b1, float b2, float eps, float c1,
                                                 float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; b++) g += gpart[(size_t)b * NP + p];
  if (gout) gout[p] = g;
  if (!adam) return;
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

struct Mlp {
  float *P = nullptr, *M = nullptr, *V = nullptr;
  float *X = nullptr, *T = nullptr;
  int *idx = nullptr;
  int n = 0;
  float *gpart = nullptr, *lossp = nullptr, *gsum = nullptr;
  int maxblk = 0, maxloss = 0;
  long long step = 0;
  hipStream_t st = nullptr;
};

template <typename T>
int mlp_alloc(T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return ferr(MCL_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  *p = static_cast<T *>(q);
  return MCL_OK;
}

// Partial gradients for ngrad workgroups and loss partials for nloss workgroups.
int ensure_partials(Mlp *m, int ngrad, int nloss) {
  int st;
  if (ngrad > m->maxblk) {
    (void)hipFree(m->gpart);
    m->gpart = nullptr;
    if ((st = mlp_alloc(&m->gpart, (size_t)ngrad * NP))) return st;
    m->maxblk = ngrad;
  }
  if (nloss > m->maxloss) {
    (void)hipFree(m->lossp);
    m->lossp = nullptr;
    if ((st = mlp_alloc(&m->lossp, nloss))) return st;
    m->maxloss = nloss;
  }
  return MCL_OK;
}

}  // namespace

extern "C" {

const char *mcl_mlp_last_error(void) { return g_err.c_str(); }

int mcl_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, void **out) {
  if (!out) return ferr(MCL_ERR_ARG, "null output");
  if (n_in != NI || n_hidden != NH || n_out != NO)
    return ferr(MCL_ERR_ARG, "this build supports the 5-64-64-3 network only");
  Mlp *m = new Mlp();
  int st;
  if ((st = mlp_alloc(&m->P, NP)) || (st = mlp_alloc(&m->M, NP)) || (st = mlp_alloc(&m->V, NP)) ||
      (st = mlp_alloc(&m->gsum, NP))) {
    delete m;
    return st;
  }
  // He-uniform initial weights (bound sqrt(6 / fan_in)) from Philox keyed by the seed; zero biases
  std::vector<float> p(NP, 0.0f), z(NP, 0.0f);
  auto fill = [&](int off, int rows, int cols) {
    const float bound = std::sqrt(6.0f / cols);
    for (int q = 0; q < rows * cols; q++) {
      const mcl::Philox4 R = mcl::philox4x32_10((uint32_t)(off + q), 0u, 0u, 0x4D4C5031u, (uint32_t)seed,
                                                (uint32_t)(seed >> 32));
      p[off + q] = (2.0f * mcl::u01_open(R.v[0]) - 1.0f) * bound;
    }
  };
  fill(OW1, NH, NI);
  fill(OW2, NH, NH);
  fill(OW3, NO, NH);
  hipError_t e = hipMemcpy(m->P, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->M, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->V, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete m;
    return ferr(MCL_ERR_DEVICE, hipGetErrorString(e));
  }
  *out = m;
  return MCL_OK;
}

void mcl_mlp_destroy(void *h) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m) return;
  (void)hipDeviceSynchronize();
  void *b[] = {m->P, m->M, m->V, m->X, m->T, m->idx, m->gpart, m->lossp, m->gsum};
  for (void *q : b)
    if (q) (void)hipFree(q);
  delete m;
}

int mcl_mlp_num_params(void) { return NP; }

int mcl_mlp_get_params(void *h, float *p, int32_t n) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || !p || n != NP) return ferr(MCL_ERR_ARG, "bad arguments");
  MLP_HIP(hipMemcpy(p, m->P, sizeof(float) * NP, hipMemcpyDeviceToHost));
  return MCL_OK;
}

int mcl_mlp_set_params(void *h, const float *p, int32_t n) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || !p || n != NP) return ferr(MCL_ERR_ARG, "bad arguments");
  MLP_HIP(hipMemcpy(m->P, p, sizeof(float) * NP, hipMemcpyHostToDevice));
  std::vector<float> z(NP, 0.0f);
  MLP_HIP(hipMemcpy(m->M, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->V, z.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
  m->step = 0;
  return MCL_OK;
}

int mcl_mlp_forward(void *h, const float *x, int32_t n, float *y) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || !x || !y || n < 0) return ferr(MCL_ERR_ARG, "bad arguments");
  if (n == 0) return MCL_OK;
  float *dx = nullptr, *dy = nullptr;
  int st;
  if ((st = mlp_alloc(&dx, (size_t)n * NI)) || (st = mlp_alloc(&dy, (size_t)n * NO))) {
    (void)hipFree(dx);
    return st;
  }
  hipError_t e = hipMemcpy(dx, x, sizeof(float) * n * NI, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, 0, m->P, dx, n, dy);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, dy, sizeof(float) * n * NO, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (e != hipSuccess) return ferr(MCL_ERR_DEVICE, std::string("forward: ") + hipGetErrorString(e));
  return MCL_OK;
}

int mcl_mlp_gradient(void *h, const float *x, const float *t, const float *w_out, int32_t n, float *grad,
                     float *loss) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || !x || !t || !grad || n <= 0) return ferr(MCL_ERR_ARG, "bad arguments");
  const float w[NO] = {w_out ? w_out[0] : 1.0f, w_out ? w_out[1] : 1.0f, w_out ? w_out[2] : 1.0f};
  const int nblk = (n + GS - 1) / GS;
  int st;
  if ((st = ensure_partials(m, nblk, nblk))) return st;
  float *dx = nullptr, *dt = nullptr;
  if ((st = mlp_alloc(&dx, (size_t)n * NI)) || (st = mlp_alloc(&dt, (size_t)n * NO))) {
    (void)hipFree(dx);
    return st;
  }
  hipError_t e = hipMemcpy(dx, x, sizeof(float) * n * NI, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt, t, sizeof(float) * n * NO, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(GT), 0, 0, m->P, dx, dt, nullptr, n, w[0], w[1], w[2],
                       1.0f / n, m->gpart, m->lossp);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NP + 255) / 256), dim3(256), 0, 0, m->P, m->M, m->V, m->gpart, nblk,
                       m->gsum, 0, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(grad, m->gsum, sizeof(float) * NP, hipMemcpyDeviceToHost);
  if (e == hipSuccess && loss) {
    std::vector<float> lp(nblk);
    e = hipMemcpy(lp.data(), m->lossp, sizeof(float) * nblk, hipMemcpyDeviceToHost);
    double l = 0.0;
    for (float v : lp) l += v;
    *loss = (float)l;
  }
  (void)hipFree(dx);
  (void)hipFree(dt);
  if (e != hipSuccess) return ferr(MCL_ERR_DEVICE, std::string("gradient: ") + hipGetErrorString(e));
  return MCL_OK;
}

int mcl_mlp_set_data(void *h, const float *x, const float *t, int32_t n) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || !x || !t || n <= 0) return ferr(MCL_ERR_ARG, "bad arguments");
  (void)hipFree(m->X);
  (void)hipFree(m->T);
  (void)hipFree(m->idx);
  m->X = nullptr;
  m->T = nullptr;
  m->idx = nullptr;
  int st;
  if ((st = mlp_alloc(&m->X, (size_t)n * NI)) || (st = mlp_alloc(&m->T, (size_t)n * NO)) ||
      (st = mlp_alloc(&m->idx, n)))
    return st;
  MLP_HIP(hipMemcpy(m->X, x, sizeof(float) * n * NI, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(m->T, t, sizeof(float) * n * NO, hipMemcpyHostToDevice));
  m->n = n;
  return MCL_OK;
}

// Mini-batch Adam over the uploaded data: each epoch visits the samples in a random order (Fisher-Yates with
// Philox keyed by the seed and the epoch); the mean loss of the last epoch is returned.
int mcl_mlp_train(void *h, int32_t epochs, int32_t batch, float lr, const float *w_out, uint64_t seed,
                  float *last_loss) {
  Mlp *m = static_cast<Mlp *>(h);
  if (!m || m->n <= 0 || epochs < 1 || batch < 1) return ferr(MCL_ERR_ARG, "bad arguments or no data");
  const float w[NO] = {w_out ? w_out[0] : 1.0f, w_out ? w_out[1] : 1.0f, w_out ? w_out[2] : 1.0f};
  const int n = m->n;
  const int bmax = (batch + GS - 1) / GS;  // gradient workgroups of one step
  const int nbatch = (n + batch - 1) / batch;
  int st;
  if ((st = ensure_partials(m, bmax, nbatch * bmax))) return st;
  std::vector<int> perm(n);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double epoch_loss = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    for (int q = 0; q < n; q++) perm[q] = q;
    for (int q = n - 1; q > 0; q--) {
      const mcl::Philox4 R = mcl::philox4x32_10((uint32_t)q, (uint32_t)ep, (uint32_t)(m->step >> 32), 0x53485546u,
                                                (uint32_t)seed, (uint32_t)(seed >> 32));
      const int r = (int)(((uint64_t)R.v[0] * (uint64_t)(q + 1)) >> 32);
      std::swap(perm[q], perm[r]);
    }
    MLP_HIP(hipMemcpy(m->idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    int nb_used = 0;
    for (int first = 0; first < n; first += batch) {
      const int cnt = std::min(batch, n - first);
      const int nblk = (cnt + GS - 1) / GS;
      m->step++;
      const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
      hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(GT), 0, 0, m->P, m->X, m->T, m->idx + first, cnt, w[0], w[1],
                         w[2], 1.0f / cnt, m->gpart + (size_t)0, m->lossp + nb_used);
      hipLaunchKernelGGL(k_mlp_adam, dim3((NP + 255) / 256), dim3(256), 0, 0, m->P, m->M, m->V, m->gpart, nblk,
                         nullptr, 1, lr, b1, b2, eps, c1, c2);
      nb_used += nblk;
    }
    MLP_HIP(hipGetLastError());
    if (ep == epochs - 1) {
      std::vector<float> lp(nb_used);
      MLP_HIP(hipMemcpy(lp.data(), m->lossp, sizeof(float) * nb_used, hipMemcpyDeviceToHost));
      // each batch's loss is its mean; weight by the batch sizes to get the epoch mean
      epoch_loss = 0.0;
      int pos = 0;
      for (int first = 0; first < n; first += batch) {
        const int cnt = std::min(batch, n - first), nblk = (cnt + GS - 1) / GS;
        double l = 0.0;
        for (int q = 0; q < nblk; q++) l += lp[pos + q];
        epoch_loss += l * cnt;
        pos += nblk;
      }
      epoch_loss /= n;
    }
  }
  MLP_HIP(hipDeviceSynchronize());
  if (last_loss) *last_loss = (float)epoch_loss;
  return MCL_OK;
}

}  // extern "C"
