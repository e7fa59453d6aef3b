// The trick-count net: an MLP that predicts NS tricks for (deal, strain, leader), used to seed
// the null-window search (DDX_SEED_CERT). Trained and evaluated on gfx950 by the kernels here:
// BF16 weights and activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16bf16_1k), FP32 master weights and Adam.
//
// Input (217 features, padded to 224): one-hot owner (4) of each of the 52 cards, one-hot
// strain (5), one-hot leader (4). Hidden: 256 ReLU, 256 ReLU. Output: 14 logits (padded to
// 16; the pad logits are masked) for NS tricks 0..13. 125,198 trained parameters.
//
// Parameter vector (FP32, this order; weights stored out x in, row-major):
//   W1t [256][224] | b1 [256] | W2t [256][256] | b2 [256] | W3t [16][256] | b3 [16]
// Input columns 217..223 and output rows 14..15 stay zero.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <random>
#include <vector>

#include "ddx.h"
#include "ddx_internal.h"

#define HIPCHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      ddx_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return DDX_E_HIP;                                                             \
    }                                                                               \
  } while (0)

namespace {

constexpr int IN = 224, IN_USED = 217, H = 256, OUT = 16, OUT_USED = 14;
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + (size_t)H * IN, OFF_W2 = OFF_B1 + H,
                 OFF_B2 = OFF_W2 + (size_t)H * H, OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + (size_t)OUT * H,
                 NPARAM = OFF_B3 + OUT;

typedef unsigned short bf16;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16 to_bf16(float f) {  // round to nearest even (finite inputs)
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16 b) {
  unsigned u = (unsigned)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Samples of deal d: row 20d + 4*strain + declarer; the leader is declarer + 1.
__global__ void encode_kernel(const uint64_t* hands, const uint8_t* tables, long long n, bf16* X,
                              int* labels) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n * 20) return;
  const long long d = s / 20;
  const int strain = (int)(s % 20) / 4, decl = (int)(s % 4), leader = (decl + 1) & 3;
  const uint64_t h0 = hands[4 * d], h1 = hands[4 * d + 1], h2 = hands[4 * d + 2], h3 = hands[4 * d + 3];
  unsigned* row = (unsigned*)(X + s * IN);
  for (int w = 0; w < IN / 2; w++) {
    unsigned v = 0;
    for (int half = 0; half < 2; half++) {
      const int f = 2 * w + half;
      bool on = false;
      if (f < 208) {
        const int c = f >> 2, o = f & 3;
        const int bitpos = 16 * (c / 13) + c % 13;
        const uint64_t hh = o == 0 ? h0 : o == 1 ? h1 : o == 2 ? h2 : h3;
        on = (hh >> bitpos) & 1;
      } else if (f < 213) {
        on = f - 208 == strain;
      } else if (f < 217) {
        on = f - 213 == leader;
      }
      if (on) v |= 0x3F80u << (16 * half);  // bf16 1.0
    }
    row[w] = v;
  }
  if (labels) {
    const int t = tables[s];  // 255: unknown (unsolved), masked out of the loss
    labels[s] = t == 255 ? -1 : (decl & 1) == 0 ? t : 13 - t;
  }
}

// C[M][N] (+)= op(A)[M][K] * op(B)[K][N], one wave per 16x16 tile, K split over blockIdx.y.
// TA: A is stored K x M (lda = row stride), else M x K. TB: B is stored N x K, else K x N.
// MFMA 16x16x16 bf16 operand layout (wave64): lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16]; the result D[4(l/16)+r][l%16] for r = 0..3.
template <bool TA, bool TB, bool ATOMIC>
__global__ __launch_bounds__(256) void gemm_kernel(const bf16* A, const bf16* B, float* C, int M,
                                                   int N, int K, int lda, int ldb, int ldc, int kchunk) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tiles_n = N / 16;
  const long long tile = (long long)blockIdx.x * 4 + wave;
  if (tile >= (long long)(M / 16) * tiles_n) return;
  const int mt = (int)(tile / tiles_n), nt = (int)(tile % tiles_n);
  const int k_begin = blockIdx.y * kchunk, k_end = min(K, k_begin + kchunk);
  const int r16 = l & 15, q = l >> 4;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_begin; k0 < k_end; k0 += 16) {
    const int kk = k0 + 4 * q;
    v4s a, b;
    if (!TA) {
      const bf16* p = A + (size_t)(mt * 16 + r16) * lda + kk;
      a = *(const v4s*)p;
    } else {
      const bf16* p = A + (size_t)kk * lda + mt * 16 + r16;
      a = v4s{(short)p[0], (short)p[lda], (short)p[2 * (size_t)lda], (short)p[3 * (size_t)lda]};
    }
    if (TB) {
      const bf16* p = B + (size_t)(nt * 16 + r16) * ldb + kk;
      b = *(const v4s*)p;
    } else {
      const bf16* p = B + (size_t)kk * ldb + nt * 16 + r16;
      b = v4s{(short)p[0], (short)p[ldb], (short)p[2 * (size_t)ldb], (short)p[3 * (size_t)ldb]};
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  float* c = C + (size_t)(mt * 16 + 4 * q) * ldc + nt * 16 + r16;
  for (int r = 0; r < 4; r++) {
    if (ATOMIC) atomicAdd(c + (size_t)r * ldc, acc[r]);
    else c[(size_t)r * ldc] = acc[r];
  }
}

// Y = relu(Z + bias) as bf16 (the next layer's input); Z is M x N fp32.
__global__ void bias_relu_kernel(const float* Z, const float* bias, bf16* Y, long long M, int N) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const float v = Z[i] + bias[i % N];
  Y[i] = to_bf16(v > 0.f ? v : 0.f);
}

// Softmax cross-entropy over the 14 used logits (+ bias); dZ = (p - y) / M as bf16; sums the
// loss and the count of right argmax predictions into stats[0], stats[1]. Rows with label -1
// (unknown) add no loss and get a zero gradient.
__global__ void softmax_ce_kernel(const float* Z, const float* b3, const int* labels, long long M,
                                  bf16* dZ, float* stats) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float loss = 0.f, right = 0.f;
  if (s < M) {
    float z[OUT];
    float mx = -1e30f;
    int arg = 0;
    for (int j = 0; j < OUT_USED; j++) {
      z[j] = Z[s * OUT + j] + b3[j];
      if (z[j] > mx) {
        mx = z[j];
        arg = j;
      }
    }
    float sum = 0.f;
    for (int j = 0; j < OUT_USED; j++) sum += expf(z[j] - mx);
    const int y = labels[s];
    if (y >= 0) {
      loss = logf(sum) - (z[y] - mx);
      right = arg == y ? 1.f : 0.f;
    }
    const float inv = y >= 0 ? 1.f / (float)M : 0.f;
    for (int j = 0; j < OUT; j++) {
      const float p = j < OUT_USED ? expf(z[j] - mx) / sum : 0.f;
      dZ[s * OUT + j] = to_bf16(((j == y ? p - 1.f : p)) * inv);
    }
  }
  // wave reduction, then one atomic per wave
  for (int m = 32; m; m >>= 1) {
    loss += __shfl_xor(loss, m);
    right += __shfl_xor(right, m);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats[0], loss);
    atomicAdd(&stats[1], right);
  }
}

// dY = dZbf16 (fp32 from the GEMM) * (Y > 0) as bf16.
__global__ void relu_grad_kernel(const float* G, const bf16* Y, bf16* dY, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dY[i] = to_bf16(Y[i] != 0 ? G[i] : 0.f);
}

// out[j] += sum over rows of G[M][N] (bf16), one block per column.
__global__ void colsum_kernel(const bf16* G, long long M, int N, float* out) {
  const int j = blockIdx.x;
  float s = 0.f;
  for (long long i = threadIdx.x; i < M; i += blockDim.x) s += from_bf16(G[i * N + j]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] += red[0];
}

// Zeros the gradients of the padding (input columns 217.., output rows 14..).
__global__ void mask_pad_kernel(float* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < H * IN && (i % IN) >= IN_USED) g[OFF_W1 + i] = 0.f;
  if (i < OUT * H && i / H >= OUT_USED) g[OFF_W3 + i] = 0.f;
  if (i < OUT && i >= OUT_USED) g[OFF_B3 + i] = 0.f;
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void to_bf16_kernel(const float* p, bf16* w, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = to_bf16(p[i]);
}

// NS-trick seed per sample: the most likely class.
__global__ void argmax_kernel(const float* Z, const float* b3, long long M, uint8_t* seeds) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  int arg = 0;
  float mx = -1e30f;
  for (int j = 0; j < OUT_USED; j++) {
    const float z = Z[s * OUT + j] + b3[j];
    if (z > mx) {
      mx = z;
      arg = j;
    }
  }
  seeds[s] = (uint8_t)arg;
}

__global__ void fill_f32_kernel(float* p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct ddx_mlp {
  int device = 0;
  float* p = nullptr;  // fp32 master parameters
  float* g = nullptr;  // gradients
  float* m = nullptr;  // Adam moments
  float* v = nullptr;
  bf16* w = nullptr;  // bf16 copy of p (weights used by the GEMMs)
  long long steps = 0;
  long long cap = 0;  // samples the activation buffers hold
  bf16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *dZ = nullptr, *dH2 = nullptr, *dH1 = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *stats = nullptr;
  int* labels = nullptr;
  uint64_t* d_hands = nullptr;
  uint8_t* d_tables = nullptr;
  long long host_cap = 0;
  uint8_t* d_seeds = nullptr;  // predictions for ddx_mlp_predict_host
  long long seeds_cap = 0;
  hipStream_t s = nullptr;
};

static void mlp_free_acts(ddx_mlp* h) {
  void* ptrs[] = {h->X, h->H1, h->H2, h->dZ, h->dH2, h->dH1, h->Z1, h->Z2, h->Z3, h->labels};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  h->X = h->H1 = h->H2 = h->dZ = h->dH2 = h->dH1 = nullptr;
  h->Z1 = h->Z2 = h->Z3 = nullptr;
  h->labels = nullptr;
  h->cap = 0;
}

static int mlp_reserve(ddx_mlp* h, long long samples) {
  if (samples <= h->cap) return DDX_OK;
  mlp_free_acts(h);
  const long long M = (samples + 15) / 16 * 16;
  HIPCHECK(hipMalloc(&h->X, M * IN * 2));
  HIPCHECK(hipMalloc(&h->H1, M * H * 2));
  HIPCHECK(hipMalloc(&h->H2, M * H * 2));
  HIPCHECK(hipMalloc(&h->dZ, M * OUT * 2));
  HIPCHECK(hipMalloc(&h->dH2, M * H * 2));
  HIPCHECK(hipMalloc(&h->dH1, M * H * 2));
  HIPCHECK(hipMalloc(&h->Z1, M * H * 4));
  HIPCHECK(hipMalloc(&h->Z2, M * H * 4));
  HIPCHECK(hipMalloc(&h->Z3, M * OUT * 4));
  HIPCHECK(hipMalloc(&h->labels, M * 4));
  HIPCHECK(hipMemsetAsync(h->X, 0, M * IN * 2, h->s));  // rows past the batch stay zero
  h->cap = M;
  return DDX_OK;
}

static int mlp_stage_host(ddx_mlp* h, const uint64_t* hands, const uint8_t* tables, long long n) {
  if (n > h->host_cap) {
    if (h->d_hands) (void)hipFree(h->d_hands);
    if (h->d_tables) (void)hipFree(h->d_tables);
    h->d_hands = nullptr;
    h->d_tables = nullptr;
    HIPCHECK(hipMalloc(&h->d_hands, n * 32));
    HIPCHECK(hipMalloc(&h->d_tables, n * 20));
    h->host_cap = n;
  }
  HIPCHECK(hipMemcpyAsync(h->d_hands, hands, n * 32, hipMemcpyHostToDevice, h->s));
  if (tables) HIPCHECK(hipMemcpyAsync(h->d_tables, tables, n * 20, hipMemcpyHostToDevice, h->s));
  return DDX_OK;
}

template <bool TA, bool TB, bool ATOMIC>
static void gemm(hipStream_t s, const bf16* A, const bf16* B, float* C, int M, int N, int K, int lda,
                 int ldb, int ldc, int kchunk) {
  const long long tiles = (long long)(M / 16) * (N / 16);
  dim3 grid(blocks(tiles, 4), (K + kchunk - 1) / kchunk);
  hipLaunchKernelGGL((gemm_kernel<TA, TB, ATOMIC>), grid, dim3(256), 0, s, A, B, C, M, N, K, lda, ldb,
                     ldc, kchunk);
}

// Activations of one forward pass (rows padded to a multiple of 16). Z2 may share Z1's memory:
// Z1 is dead once H1 is computed.
struct Acts {
  bf16* X;
  float* Z1;
  bf16* H1;
  float* Z2;
  bf16* H2;
  float* Z3;
  int* labels;
};

// Forward pass for n deals in d_hands on stream s (labels when tables != null): Z3 = logits
// without the output bias.
static void forward_into(const ddx_mlp* h, hipStream_t s, const Acts& a, const uint64_t* d_hands,
                         const uint8_t* d_tables, long long n) {
  const long long S = n * 20, M = (S + 15) / 16 * 16;
  hipLaunchKernelGGL(encode_kernel, dim3(blocks(S)), dim3(256), 0, s, d_hands, d_tables, n, a.X,
                     d_tables ? a.labels : nullptr);
  const bf16* W1 = h->w + OFF_W1;
  const bf16* W2 = h->w + OFF_W2;
  const bf16* W3 = h->w + OFF_W3;
  gemm<false, true, false>(s, a.X, W1, a.Z1, (int)M, H, IN, IN, IN, H, IN);
  hipLaunchKernelGGL(bias_relu_kernel, dim3(blocks(M * H)), dim3(256), 0, s, a.Z1, h->p + OFF_B1, a.H1, M, H);
  gemm<false, true, false>(s, a.H1, W2, a.Z2, (int)M, H, H, H, H, H, H);
  hipLaunchKernelGGL(bias_relu_kernel, dim3(blocks(M * H)), dim3(256), 0, s, a.Z2, h->p + OFF_B2, a.H2, M, H);
  gemm<false, true, false>(s, a.H2, W3, a.Z3, (int)M, OUT, H, H, H, OUT, H);
}

static void forward(ddx_mlp* h, const uint64_t* d_hands, const uint8_t* d_tables, long long n) {
  forward_into(h, h->s, Acts{h->X, h->Z1, h->H1, h->Z2, h->H2, h->Z3, h->labels}, d_hands, d_tables, n);
}

// ---- scratch of the allocation-free prediction ----------------------------------------------
// Activations for at most kNoAllocDeals deals (20 samples each) in one file-scope device array:
// X | Z (Z1, then Z2) | H1 | H2 | Z3, every part a multiple of 256 bytes. 13.1 MB, held by every
// GPU the library runs on until the process ends. In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU comes from hipGetSymbolAddress
// (resolved once per device by ddx_mlp_create) and only that address is ever used.
constexpr long long kNoAllocDeals = 256, kNoAllocRows = kNoAllocDeals * 20;
constexpr size_t kNaX = 0, kNaZ = kNaX + (size_t)kNoAllocRows * IN * 2,
                 kNaH1 = kNaZ + (size_t)kNoAllocRows * H * 4, kNaH2 = kNaH1 + (size_t)kNoAllocRows * H * 2,
                 kNaZ3 = kNaH2 + (size_t)kNoAllocRows * H * 2, kNaBytes = kNaZ3 + (size_t)kNoAllocRows * OUT * 4;
static_assert(kNaZ % 256 == 0 && kNaH1 % 256 == 0 && kNaH2 % 256 == 0 && kNaZ3 % 256 == 0, "scratch parts");
static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per GPU

// With `device` current: looks up the device address of `scratch` on it, once.
static int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) {
    ddx_set_error("device %d is beyond the %d devices the scratch table covers", device, kMaxDevices);
    return DDX_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return DDX_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    ddx_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return DDX_E_HIP;
  }
  g_scratch_ptr[device] = (unsigned char*)p;
  return DDX_OK;
}

static unsigned char* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Gradients of the mean loss over the staged batch into h->g; loss/accuracy into stats.
static void backward(ddx_mlp* h, long long n) {
  const long long S = n * 20, M = (S + 15) / 16 * 16;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks(NPARAM)), dim3(256), 0, h->s, h->g, (long long)NPARAM, 0.f);
  hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, h->s, h->stats, 2LL, 0.f);
  // rows S..M of dZ must be zero: they are padding
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks(M * OUT / 2)), dim3(256), 0, h->s, (float*)h->dZ,
                     M * OUT / 2, 0.f);
  hipLaunchKernelGGL(softmax_ce_kernel, dim3(blocks(S)), dim3(256), 0, h->s, h->Z3, h->p + OFF_B3,
                     h->labels, S, h->dZ, h->stats);
  const int kc = 1024;  // rows per split-K chunk of the weight gradients
  // layer 3
  gemm<true, false, true>(h->s, h->dZ, h->H2, h->g + OFF_W3, OUT, H, (int)M, OUT, H, H, kc);
  hipLaunchKernelGGL(colsum_kernel, dim3(OUT), dim3(256), 0, h->s, h->dZ, M, OUT, h->g + OFF_B3);
  gemm<false, false, false>(h->s, h->dZ, h->w + OFF_W3, h->Z2, (int)M, H, OUT, OUT, H, H, OUT);
  hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks(M * H)), dim3(256), 0, h->s, h->Z2, h->H2, h->dH2, M * H);
  // layer 2
  gemm<true, false, true>(h->s, h->dH2, h->H1, h->g + OFF_W2, H, H, (int)M, H, H, H, kc);
  hipLaunchKernelGGL(colsum_kernel, dim3(H), dim3(256), 0, h->s, h->dH2, M, H, h->g + OFF_B2);
  gemm<false, false, false>(h->s, h->dH2, h->w + OFF_W2, h->Z1, (int)M, H, H, H, H, H, H);
  hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks(M * H)), dim3(256), 0, h->s, h->Z1, h->H1, h->dH1, M * H);
  // layer 1
  gemm<true, false, true>(h->s, h->dH1, h->X, h->g + OFF_W1, H, IN, (int)M, H, IN, IN, kc);
  hipLaunchKernelGGL(colsum_kernel, dim3(H), dim3(256), 0, h->s, h->dH1, M, H, h->g + OFF_B1);
  hipLaunchKernelGGL(mask_pad_kernel, dim3(blocks(H * IN)), dim3(256), 0, h->s, h->g);
}

static void sync_bf16(ddx_mlp* h) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(NPARAM)), dim3(256), 0, h->s, h->p, h->w, (long long)NPARAM);
}

extern "C" int ddx_mlp_create(int device, uint64_t seed, ddx_mlp** out) {
  if (!out) return DDX_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
    ddx_set_error("ddx_mlp_create: no HIP device %d", device);
    return DDX_E_NODEV;
  }
  HIPCHECK(hipSetDevice(device));
  ddx_mlp* h = new (std::nothrow) ddx_mlp;
  if (!h) return DDX_E_NOMEM;
  h->device = device;
  if (int rc = resolve_scratch(device)) {
    delete h;
    return rc;
  }
  HIPCHECK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
  HIPCHECK(hipMalloc(&h->p, NPARAM * 4));
  HIPCHECK(hipMalloc(&h->g, NPARAM * 4));
  HIPCHECK(hipMalloc(&h->m, NPARAM * 4));
  HIPCHECK(hipMalloc(&h->v, NPARAM * 4));
  HIPCHECK(hipMalloc(&h->w, NPARAM * 2));
  HIPCHECK(hipMalloc(&h->stats, 64));
  // He initialisation of the used weights, zero biases and padding.
  std::vector<float> p(NPARAM, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](size_t off, int rows, int cols, int used_rows, int used_cols, int fan_in) {
    std::normal_distribution<float> nd(0.f, sqrtf(2.f / fan_in));
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) p[off + (size_t)r * cols + c] = (r < used_rows && c < used_cols) ? nd(rng) : 0.f;
  };
  init(OFF_W1, H, IN, H, IN_USED, IN_USED);
  init(OFF_W2, H, H, H, H, H);
  init(OFF_W3, OUT, H, OUT_USED, H, H);
  HIPCHECK(hipMemcpy(h->p, p.data(), NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(h->m, 0, NPARAM * 4));
  HIPCHECK(hipMemset(h->v, 0, NPARAM * 4));
  sync_bf16(h);
  HIPCHECK(hipStreamSynchronize(h->s));
  *out = h;
  return DDX_OK;
}

extern "C" void ddx_mlp_destroy(ddx_mlp* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->s) (void)hipStreamSynchronize(h->s);
  mlp_free_acts(h);
  void* ptrs[] = {h->p, h->g, h->m, h->v, h->w, h->stats, h->d_hands, h->d_tables, h->d_seeds};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

extern "C" int64_t ddx_mlp_param_count(void) { return (int64_t)NPARAM; }

extern "C" int ddx_mlp_params(ddx_mlp* h, float* buf, int get) {
  if (!h || !buf) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->device));
  if (get) {
    HIPCHECK(hipMemcpy(buf, h->p, NPARAM * 4, hipMemcpyDeviceToHost));
  } else {
    HIPCHECK(hipMemcpy(h->p, buf, NPARAM * 4, hipMemcpyHostToDevice));
    sync_bf16(h);
    HIPCHECK(hipStreamSynchronize(h->s));
  }
  return DDX_OK;
}

extern "C" int ddx_mlp_grads(ddx_mlp* h, const uint64_t* hands, const uint8_t* tables, int64_t n,
                             float* grads, float* loss_acc) {
  if (!h || !hands || !tables || n <= 0) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->device));
  int rc;
  if ((rc = mlp_reserve(h, n * 20)) != DDX_OK || (rc = mlp_stage_host(h, hands, tables, n)) != DDX_OK)
    return rc;
  forward(h, h->d_hands, h->d_tables, n);
  backward(h, n);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(h->s));
  if (grads) HIPCHECK(hipMemcpy(grads, h->g, NPARAM * 4, hipMemcpyDeviceToHost));
  if (loss_acc) {
    float st[2];
    HIPCHECK(hipMemcpy(st, h->stats, 8, hipMemcpyDeviceToHost));
    loss_acc[0] = st[0] / (float)(n * 20);
    loss_acc[1] = st[1] / (float)(n * 20);
  }
  return DDX_OK;
}

extern "C" int ddx_mlp_train(ddx_mlp* h, const uint64_t* hands, const uint8_t* tables, int64_t n,
                             int epochs, int batch_deals, float lr, uint64_t shuffle_seed,
                             float* loss_acc) {
  if (!h || !hands || !tables || n <= 0 || epochs < 1 || batch_deals < 1) {
    ddx_set_error("ddx_mlp_train: bad arguments");
    return DDX_E_ARG;
  }
  HIPCHECK(hipSetDevice(h->device));
  int rc;
  if ((rc = mlp_reserve(h, (long long)batch_deals * 20)) != DDX_OK) return rc;
  std::vector<long long> order(n);
  for (long long i = 0; i < n; i++) order[i] = i;
  std::mt19937_64 rng(shuffle_seed);
  std::vector<uint64_t> bh((size_t)batch_deals * 4);
  std::vector<uint8_t> bt((size_t)batch_deals * 20);
  double loss_sum = 0, acc_sum = 0, samples = 0;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(order.begin(), order.end(), rng);
    loss_sum = acc_sum = samples = 0;
    for (long long b0 = 0; b0 + batch_deals <= n; b0 += batch_deals) {
      for (int i = 0; i < batch_deals; i++) {
        memcpy(&bh[4 * (size_t)i], hands + 4 * order[b0 + i], 32);
        memcpy(&bt[20 * (size_t)i], tables + 20 * order[b0 + i], 20);
      }
      if ((rc = mlp_stage_host(h, bh.data(), bt.data(), batch_deals)) != DDX_OK) return rc;
      forward(h, h->d_hands, h->d_tables, batch_deals);
      backward(h, batch_deals);
      h->steps++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - powf(b1, (float)h->steps), c2 = 1.f - powf(b2, (float)h->steps);
      hipLaunchKernelGGL(adam_kernel, dim3(blocks(NPARAM)), dim3(256), 0, h->s, h->p, h->g, h->m, h->v,
                         (long long)NPARAM, lr, b1, b2, c1, c2, 1e-8f);
      sync_bf16(h);
      float st[2];
      HIPCHECK(hipMemcpyAsync(st, h->stats, 8, hipMemcpyDeviceToHost, h->s));
      HIPCHECK(hipStreamSynchronize(h->s));
      loss_sum += st[0];
      acc_sum += st[1];
      samples += batch_deals * 20.0;
    }
  }
  HIPCHECK(hipGetLastError());
  if (loss_acc && samples > 0) {
    loss_acc[0] = (float)(loss_sum / samples);
    loss_acc[1] = (float)(acc_sum / samples);
  }
  return DDX_OK;
}

extern "C" int ddx_mlp_predict(ddx_mlp* h, const uint64_t* d_hands, int64_t n, uint8_t* d_seeds,
                               void* stream) {
  if (!h || !d_hands || !d_seeds || n < 0) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->device));
  int rc;
  const long long chunk = 1 << 16;  // deals per pass through the activation buffers
  if ((rc = mlp_reserve(h, (n < chunk ? n : chunk) * 20)) != DDX_OK) return rc;
  hipEvent_t ev;
  if (stream) {  // order after the caller's stream
    HIPCHECK(hipEventCreate(&ev));
    HIPCHECK(hipEventRecord(ev, (hipStream_t)stream));
    HIPCHECK(hipStreamWaitEvent(h->s, ev, 0));
    HIPCHECK(hipEventDestroy(ev));
  }
  for (long long d0 = 0; d0 < n; d0 += chunk) {
    const long long m = n - d0 < chunk ? n - d0 : chunk;
    forward(h, d_hands + 4 * d0, nullptr, m);
    hipLaunchKernelGGL(argmax_kernel, dim3(blocks(m * 20)), dim3(256), 0, h->s, h->Z3, h->p + OFF_B3,
                       m * 20, d_seeds + 20 * d0);
  }
  HIPCHECK(hipGetLastError());
  if (stream) {  // the caller's stream waits for the seeds
    HIPCHECK(hipEventCreate(&ev));
    HIPCHECK(hipEventRecord(ev, h->s));
    HIPCHECK(hipStreamWaitEvent((hipStream_t)stream, ev, 0));
    HIPCHECK(hipEventDestroy(ev));
  } else {
    HIPCHECK(hipStreamSynchronize(h->s));
  }
  return DDX_OK;
}

// ddx_mlp_predict without any allocation: the activations live in the file-scope `scratch` (at
// most kNoAllocDeals deals per call), every kernel runs on the caller's stream (NULL: the null
// stream, and the call waits for the seeds). Calls that overlap on one GPU share that scratch
// and race; the caller orders them.
extern "C" int ddx_mlp_predict_noalloc(const ddx_mlp* h, const uint64_t* d_hands, int64_t n,
                                       uint8_t* d_seeds, void* stream) {
  if (!h || n < 0 || (n > 0 && (!d_hands || !d_seeds))) {
    ddx_set_error("ddx_mlp_predict_noalloc: bad arguments");
    return DDX_E_ARG;
  }
  if (n > kNoAllocDeals) {
    ddx_set_error("ddx_mlp_predict_noalloc: %lld deals exceed its scratch of %lld deals", (long long)n,
                  kNoAllocDeals);
    return DDX_E_ARG;
  }
  HIPCHECK(hipSetDevice(h->device));
  unsigned char* base = scratch_of(h->device);
  if (!base) {
    ddx_set_error("ddx_mlp_predict_noalloc: no scratch on device %d", h->device);
    return DDX_E_HIP;
  }
  if (n == 0) return DDX_OK;
  hipStream_t s = (hipStream_t)stream;
  float* Z = (float*)(base + kNaZ);
  const Acts a{(bf16*)(base + kNaX), Z, (bf16*)(base + kNaH1), Z, (bf16*)(base + kNaH2),
               (float*)(base + kNaZ3), nullptr};
  const long long S = n * 20, M = (S + 15) / 16 * 16;
  if (M > S)  // padding rows of the input: zero (the scratch keeps the last call's contents)
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks((M - S) * IN / 2)), dim3(256), 0, s, (float*)(a.X + S * IN),
                       (M - S) * IN / 2, 0.f);
  forward_into(h, s, a, d_hands, nullptr, n);
  hipLaunchKernelGGL(argmax_kernel, dim3(blocks(S)), dim3(256), 0, s, a.Z3, h->p + OFF_B3, S, d_seeds);
  HIPCHECK(hipGetLastError());
  if (!stream) HIPCHECK(hipStreamSynchronize(nullptr));
  return DDX_OK;
}

// Host in, host out. Every copy and kernel runs on the handle's non-blocking stream and
// buffers are reused (no per-call allocation or free, which could wait for the whole device),
// so predictions can be made while a persistent solver kernel runs on another stream.
extern "C" int ddx_mlp_predict_host(ddx_mlp* h, const uint64_t* hands, int64_t n, uint8_t* seeds) {
  if (!h || !hands || !seeds || n < 0) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->device));
  int rc;
  if ((rc = mlp_stage_host(h, hands, nullptr, n)) != DDX_OK) return rc;
  if (n > h->seeds_cap) {
    if (h->d_seeds) (void)hipFree(h->d_seeds);
    h->d_seeds = nullptr;
    h->seeds_cap = 0;
    HIPCHECK(hipMalloc(&h->d_seeds, n * 20 + 16));
    h->seeds_cap = n;
  }
  rc = ddx_mlp_predict(h, h->d_hands, n, h->d_seeds, nullptr);
  if (rc != DDX_OK) return rc;
  HIPCHECK(hipMemcpyAsync(seeds, h->d_seeds, n * 20, hipMemcpyDeviceToHost, h->s));
  HIPCHECK(hipStreamSynchronize(h->s));
  return DDX_OK;
}

// File: "DDXMLP1\0", int64 count, then count fp32 parameters.
extern "C" int ddx_mlp_save(ddx_mlp* h, const char* path) {
  std::vector<float> p(NPARAM);
  int rc = ddx_mlp_params(h, p.data(), 1);
  if (rc != DDX_OK) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) {
    ddx_set_error("cannot write %s", path);
    return DDX_E_IO;
  }
  const int64_t cnt = NPARAM;
  bool ok = fwrite("DDXMLP1", 1, 8, f) == 8 && fwrite(&cnt, 8, 1, f) == 1 &&
            fwrite(p.data(), 4, NPARAM, f) == NPARAM;
  fclose(f);
  return ok ? DDX_OK : DDX_E_IO;
}

extern "C" int ddx_mlp_load(ddx_mlp* h, const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) {
    ddx_set_error("cannot open %s", path);
    return DDX_E_IO;
  }
  char magic[8];
  int64_t cnt = 0;
  std::vector<float> p(NPARAM);
  bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "DDXMLP1", 8) && fread(&cnt, 8, 1, f) == 1 &&
            cnt == (int64_t)NPARAM && fread(p.data(), 4, NPARAM, f) == NPARAM;
  fclose(f);
  if (!ok) {
    ddx_set_error("%s is not a ddx MLP file", path);
    return DDX_E_IO;
  }
  return ddx_mlp_params(h, p.data(), 0);
}
