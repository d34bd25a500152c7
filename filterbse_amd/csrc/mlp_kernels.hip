// mlp_kernels.hip -- the warm-start network: BF16 MFMA forward/backward,
// Adam, grid features, band-edge densities and the warm start itself.
//
// Dense layers are point-major GEMMs with at most 128 columns: a wavefront
// owns 16 points and all column tiles of the layer (v_mfma_f32_16x16x16_bf16,
// one 16x16 FP32 accumulator tile per 16 columns), the layer's BF16 weights sit
// in LDS.  Weight gradients reduce over points: each work-group reduces a
// chunk of points for every 16x16 tile of dW with the points as the MFMA's K
// dimension, writes its partial tile set, and a second kernel sums the
// partials in a fixed order (deterministic).  Activations are rounded to BF16
// after each layer, pre-activations kept in FP32 for the SiLU derivative.
//
// MFMA operand layout (16x16x16, wave64): lane l holds A[l%16][4*(l/16)+q]
// and B[4*(l/16)+q][l%16] for q < 4; the result D[4*(l/16)+q][l%16].
#include "fbse_internal.h"
#include "../../include/fbse_mlp.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

struct fbse_mlp {
  int device = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  std::string err;
  int n_in = 0, k_in = 0, width = 0, n_out = 0;
  long n_params = 0;
  long off_w[4] = {}, off_b[4] = {};  // packed FP32 parameter offsets
  int lk[4] = {}, ln[4] = {};         // padded layer shapes (K x N)
  float *d_params = nullptr, *d_grads = nullptr, *d_m = nullptr, *d_v = nullptr;
  unsigned short *d_wb[4] = {};       // BF16 weights, padded K x N
  float *d_bias[4] = {};              // FP32 biases, padded to N
  long adam_t = 0;
  // training buffers (grown on demand)
  long cap = 0;
  unsigned short *d_act[4] = {};      // BF16 activations of layers 1..3 (and dZ staging)
  float *d_z[4] = {};                 // FP32 pre-activations of layers 1..3
  float *d_y = nullptr;               // FP32 outputs
  unsigned short *d_dz = nullptr, *d_dz2 = nullptr;  // BF16 dZ of the current / previous layer
  float *d_part = nullptr;            // per-work-group partial gradients
  long part_cap = 0;
  double *d_red = nullptr;            // reductions (loss, weight sum)
};

namespace fbse {

static int mlp_err(fbse_mlp *m, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  m->err = buf;
  return code;
}
#define MLP_HIP(m, call)                                                                                  \
  do {                                                                                                    \
    hipError_t e_ = (call);                                                                               \
    if (e_ != hipSuccess) return mlp_err((m), FBSE_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));      \
  } while (0)

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even (finite inputs)
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

constexpr int kMlpThreads = 256;  // 4 wavefronts, 16 points each
constexpr int kMlpPoints = 64;    // points per work-group in the point GEMMs
constexpr int kWgChunk = 128;     // points per chunk of the weight-gradient reduction
constexpr int kWgBlocks = 1024;   // work-groups of the weight-gradient kernel (at most)

// Y = act(X W + b) for n points.  X: n x K BF16 (ld K), W: K x N BF16 in LDS,
// NT = N / 16.  OUT_F32: write FP32 Y (the linear output layer), else BF16
// Y = silu(Z) and the FP32 pre-activation Z.
template <int NT, bool OUT_F32>
__global__ __launch_bounds__(kMlpThreads) void dense_fwd_kernel(const unsigned short *__restrict__ X, int K,
                                                                const unsigned short *__restrict__ W,
                                                                const float *__restrict__ bias, long n,
                                                                unsigned short *__restrict__ Yb,
                                                                float *__restrict__ Z, float *__restrict__ Yf) {
  constexpr int N = NT * 16;
  extern __shared__ unsigned short s_w[];  // K x N
  for (int i = threadIdx.x; i < K * N; i += kMlpThreads) s_w[i] = W[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long p0 = (long)blockIdx.x * kMlpPoints + wave * 16;
  const int r = lane & 15, kq = 4 * (lane >> 4);
  f32x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const long pa = p0 + r;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a = bf16x4_t{0, 0, 0, 0};
    if (pa < n) {
      const unsigned short *xp = X + pa * K + k0 + kq;
      a = bf16x4_t{(short)xp[0], (short)xp[1], (short)xp[2], (short)xp[3]};
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const unsigned short *wp = s_w + (k0 + kq) * N + t * 16 + r;
      bf16x4_t b = bf16x4_t{(short)wp[0], (short)wp[N], (short)wp[2 * N], (short)wp[3 * N]};
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = t * 16 + r;
    const float bc = bias[col];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long p = p0 + kq + q;
      if (p >= n) continue;
      const float z = acc[t][q] + bc;
      if (OUT_F32) {
        Yf[p * N + col] = z;
      } else {
        Z[p * N + col] = z;
        Yb[p * N + col] = f2bf(silu(z));
      }
    }
  }
}

// dZprev = (dZ W^T) * silu'(Zprev), BF16: dZ n x N (BF16), W K x N (LDS),
// Zprev n x K FP32; output n x K.  KT = K / 16 output tiles, reduction over N.
template <int KT>
__global__ __launch_bounds__(kMlpThreads) void dense_bwd_data_kernel(const unsigned short *__restrict__ dZ, int N,
                                                                     const unsigned short *__restrict__ W,
                                                                     const float *__restrict__ Zprev, long n,
                                                                     unsigned short *__restrict__ dZprev) {
  constexpr int K = KT * 16;
  extern __shared__ unsigned short s_w[];  // K x N
  for (int i = threadIdx.x; i < K * N; i += kMlpThreads) s_w[i] = W[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long p0 = (long)blockIdx.x * kMlpPoints + wave * 16;
  const int r = lane & 15, kq = 4 * (lane >> 4);
  f32x4_t acc[KT];
#pragma unroll
  for (int t = 0; t < KT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const long pa = p0 + r;
  for (int j0 = 0; j0 < N; j0 += 16) {
    bf16x4_t a = bf16x4_t{0, 0, 0, 0};
    if (pa < n) {
      const unsigned short *dp = dZ + pa * N + j0 + kq;
      a = bf16x4_t{(short)dp[0], (short)dp[1], (short)dp[2], (short)dp[3]};
    }
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      // B[j][i] = W^T[j][i] = W[i][j], i = t*16 + r, j = j0 + kq + q
      const unsigned short *wp = s_w + (t * 16 + r) * N + j0 + kq;
      bf16x4_t b = bf16x4_t{(short)wp[0], (short)wp[1], (short)wp[2], (short)wp[3]};
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    const int col = t * 16 + r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long p = p0 + kq + q;
      if (p >= n) continue;
      dZprev[p * K + col] = f2bf(acc[t][q] * silu_grad(Zprev[p * K + col]));
    }
  }
}

// Partial weight and bias gradients of the chunks of kWgChunk points a
// work-group takes (chunks blockIdx.x, blockIdx.x + gridDim.x, ... in order):
// part[blk][k*N + j] = sum_p X[p][k] dZ[p][j], part[blk][K*N + j] = sum_p dZ[p][j].
// Every 16x16 tile of dW is an MFMA accumulation with the points as K; the
// accumulators stay in registers across the work-group's chunks, so the
// partial buffer has one tile per work-group, not one per chunk.
template <int KT, int NT>
__global__ __launch_bounds__(kMlpThreads) void dense_wgrad_kernel(const unsigned short *__restrict__ X,
                                                                  const unsigned short *__restrict__ dZ, long n,
                                                                  float *__restrict__ part) {
  constexpr int K = KT * 16, N = NT * 16, TILES = KT * NT;
  constexpr int TPW = (TILES + 3) / 4;  // tiles per wavefront
  __shared__ unsigned short s_x[kWgChunk * K];
  __shared__ unsigned short s_d[kWgChunk * N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kq = 4 * (lane >> 4);
  f32x4_t acc[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) acc[u] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;  // bias gradient of column threadIdx.x (< N)
  const long nblk = (n + kWgChunk - 1) / kWgChunk;
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long base = blk * kWgChunk;
    const int np = (int)((n - base) < kWgChunk ? (n - base) : kWgChunk);
    __syncthreads();  // the previous chunk's tiles are consumed
    for (int i = threadIdx.x; i < kWgChunk * K; i += kMlpThreads) s_x[i] = (i / K) < np ? X[base * K + i] : 0;
    for (int i = threadIdx.x; i < kWgChunk * N; i += kMlpThreads) s_d[i] = (i / N) < np ? dZ[base * N + i] : 0;
    __syncthreads();
    for (int p0 = 0; p0 < kWgChunk; p0 += 16) {
#pragma unroll
      for (int u = 0; u < TPW; ++u) {
        const int tile = wave * TPW + u;
        if (tile >= TILES) continue;
        const int kt = tile / NT, nt = tile - kt * NT;
        // A = X^T tile: A[i][pp] = X[p0+pp][kt*16+i]; B = dZ tile: B[pp][j] = dZ[p0+pp][nt*16+j]
        const unsigned short *xp = s_x + (p0 + kq) * K + kt * 16 + r;
        const unsigned short *dp = s_d + (p0 + kq) * N + nt * 16 + r;
        bf16x4_t a = bf16x4_t{(short)xp[0], (short)xp[K], (short)xp[2 * K], (short)xp[3 * K]};
        bf16x4_t b = bf16x4_t{(short)dp[0], (short)dp[N], (short)dp[2 * N], (short)dp[3 * N]};
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[u], 0, 0, 0);
      }
    }
    if (threadIdx.x < N) {
      float s = 0.f;
      for (int p = 0; p < kWgChunk; ++p) s += bf2f(s_d[p * N + threadIdx.x]);
      bsum += s;
    }
  }
  float *out = part + (long)blockIdx.x * (K * N + N);
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int tile = wave * TPW + u;
    if (tile >= TILES) continue;
    const int kt = tile / NT, nt = tile - kt * NT;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(kt * 16 + kq + q) * N + nt * 16 + r] = acc[u][q];
  }
  if (threadIdx.x < N) out[K * N + threadIdx.x] = bsum;
}

// grad (packed layout) = sum over blocks of the partials (fixed order); the
// padded rows/columns of the layer are dropped.
__global__ void reduce_part_kernel(const float *__restrict__ part, int nblk, int K, int N, int k_real, int n_real,
                                   float *__restrict__ gw, float *__restrict__ gb) {
  const int stride = K * N + N;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < stride; i += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(long)b * stride + i];
    if (i < K * N) {
      const int k = i / N, j = i - k * N;
      if (k < k_real && j < n_real) gw[k * n_real + j] = s;
    } else {
      const int j = i - K * N;
      if (j < n_real) gb[j] = s;
    }
  }
}

// sum of the weights (or the count) and the weighted squared error
__global__ void loss_sums_kernel(const float *__restrict__ y, const float *__restrict__ t,
                                 const float *__restrict__ w, long n, int n_out, int ld, double *__restrict__ red) {
  __shared__ double s_w[kMlpThreads], s_e[kMlpThreads];
  double sw = 0.0, se = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * n_out; i += (long)gridDim.x * blockDim.x) {
    const long p = i / n_out;
    const int o = (int)(i - p * n_out);
    const double wi = w ? (double)w[i] : 1.0;
    const double d = (double)y[p * ld + o] - (double)t[i];
    sw += wi;
    se += wi * d * d;
  }
  s_w[threadIdx.x] = sw;
  s_e[threadIdx.x] = se;
  __syncthreads();
  for (int o = kMlpThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_w[threadIdx.x] += s_w[threadIdx.x + o];
      s_e[threadIdx.x] += s_e[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[2 + 2 * blockIdx.x] = s_w[0];
    red[3 + 2 * blockIdx.x] = s_e[0];
  }
}
__global__ void loss_final_kernel(double *red, int nblk) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double sw = 0.0, se = 0.0;
    for (int b = 0; b < nblk; ++b) {
      sw += red[2 + 2 * b];
      se += red[3 + 2 * b];
    }
    red[0] = sw;
    red[1] = se;
  }
}
// dZ of the output layer: 2 w (y - t) / sum w, BF16, padded columns 0
__global__ void loss_grad_kernel(const float *__restrict__ y, const float *__restrict__ t,
                                 const float *__restrict__ w, long n, int n_out, int ld,
                                 const double *__restrict__ red, unsigned short *__restrict__ dz) {
  const double inv = red[0] > 0.0 ? 2.0 / red[0] : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n * ld; i += (long)gridDim.x * blockDim.x) {
    const long p = i / ld;
    const int o = (int)(i - p * ld);
    float g = 0.f;
    if (o < n_out) {
      const long it = p * n_out + o;
      const double wi = w ? (double)w[it] : 1.0;
      g = (float)(inv * wi * ((double)y[i] - (double)t[it]));
    }
    dz[i] = f2bf(g);
  }
}

// Adam on the packed FP32 parameters, then the BF16 / padded device copies
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, float lr, float c1, float c2) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  }
}
__global__ void pack_layer_kernel(const float *__restrict__ w, const float *__restrict__ b, int k_real, int n_real,
                                  int K, int N, unsigned short *__restrict__ wb, float *__restrict__ bb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < K * N + N; i += gridDim.x * blockDim.x) {
    if (i < K * N) {
      const int k = i / N, j = i - k * N;
      wb[i] = (k < k_real && j < n_real) ? f2bf(w[k * n_real + j]) : 0;
    } else {
      const int j = i - K * N;
      bb[j] = j < n_real ? b[j] : 0.f;
    }
  }
}

// ---------------------------------------------------------------------------
// features, densities, warm start
// ---------------------------------------------------------------------------
constexpr int kMaxSpecies = 4;
constexpr int kNearest = 8;
constexpr int kAtomTile = 512;

__global__ __launch_bounds__(256) void features_kernel(long nx, long ny, long nz, double ox, double oy, double oz,
                                                       double dx, const float4 *__restrict__ atoms, long n_atoms,
                                                       int n_species, const double *__restrict__ vloc, double cx,
                                                       double cy, double cz, double inv_rs, int k_in,
                                                       unsigned short *__restrict__ out) {
  __shared__ float4 s_at[kAtomTile];
  const long ng = nx * ny * nz;
  const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (g < ng) {
    const long jx = g % nx, jy = (g / nx) % ny, jz = g / (nx * ny);
    px = (float)(ox + dx * (double)jx);
    py = (float)(oy + dx * (double)jy);
    pz = (float)(oz + dx * (double)jz);
  }
  float best[kMaxSpecies][kNearest];
#pragma unroll
  for (int s = 0; s < kMaxSpecies; ++s)
#pragma unroll
    for (int q = 0; q < kNearest; ++q) best[s][q] = 3.0e38f;
  for (long a0 = 0; a0 < n_atoms; a0 += kAtomTile) {
    const int na = (int)((n_atoms - a0) < kAtomTile ? (n_atoms - a0) : kAtomTile);
    __syncthreads();
    for (int i = threadIdx.x; i < na; i += blockDim.x) s_at[i] = atoms[a0 + i];
    __syncthreads();
    for (int i = 0; i < na; ++i) {
      const float4 at = s_at[i];
      const float ddx = px - at.x, ddy = py - at.y, ddz = pz - at.z;
      float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
      const int s = (int)at.w;
#pragma unroll
      for (int ss = 0; ss < kMaxSpecies; ++ss) {
        if (ss != s) continue;
        // insertion into the ascending list of the 8 smallest squared distances
#pragma unroll
        for (int q = 0; q < kNearest; ++q) {
          const float cur = best[ss][q];
          const bool sw = d2 < cur;
          best[ss][q] = sw ? d2 : cur;
          d2 = sw ? cur : d2;
        }
      }
    }
  }
  if (g >= ng) return;
  unsigned short *o = out + g * k_in;
  int c = 0;
#pragma unroll
  for (int s = 0; s < kMaxSpecies; ++s) {
    if (s >= n_species) break;
#pragma unroll
    for (int q = 0; q < kNearest; ++q) {
      const float d2 = best[s][q];
      o[c++] = f2bf(d2 < 1.0e38f ? 1.0f / (1.0f + 0.5f * sqrtf(d2)) : 0.0f);
    }
  }
  o[c++] = f2bf((float)vloc[g]);
  const double rx = px - cx, ry = py - cy, rz = pz - cz;
  o[c++] = f2bf((float)(sqrt(rx * rx + ry * ry + rz * rz) * inv_rs));
  for (; c < k_in; ++c) o[c] = 0;
}

// rho[g] += sum over spins |psi(g, sigma)|^2 of one state (complex or real layout)
__global__ void density_add_kernel(const double *__restrict__ psi, long ngrid, int nspin, int cplx,
                                   double *__restrict__ rho) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < ngrid; g += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int sp = 0; sp < nspin; ++sp) {
      const long i = sp * ngrid + g;
      if (cplx) {
        const double re = psi[2 * i], im = psi[2 * i + 1];
        s += re * re + im * im;
      } else {
        s += psi[i] * psi[i];
      }
    }
    rho[g] += s;
  }
}

// start state (blockIdx.y) *= envelope(g); partial sums of |psi|^2 per block
constexpr int kWarmBlocks = 256;
__global__ void warm_envelope_kernel(const float *__restrict__ y, int ld, int channel, double *__restrict__ psi,
                                     long state_stride, long ngrid, int nspin, int cplx,
                                     double *__restrict__ part) {
  __shared__ double s_sum[256];
  double *st = psi + blockIdx.y * state_stride;
  double acc = 0.0;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < ngrid; g += (long)gridDim.x * blockDim.x) {
    float lr = y[g * ld + channel];
    lr = lr > 30.f ? 30.f : (lr < -30.f ? -30.f : lr);
    const double env = exp(0.5 * (double)lr);
    for (int sp = 0; sp < nspin; ++sp) {
      const long i = sp * ngrid + g;
      if (cplx) {
        const double re = st[2 * i] * env, im = st[2 * i + 1] * env;
        st[2 * i] = re;
        st[2 * i + 1] = im;
        acc += re * re + im * im;
      } else {
        const double v = st[i] * env;
        st[i] = v;
        acc += v * v;
      }
    }
  }
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = s_sum[0];
}
__global__ void warm_scale_kernel(double *__restrict__ psi, long state_stride, long elems, int nblk,
                                  const double *__restrict__ part, double dv) {
  double *st = psi + blockIdx.y * state_stride;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[blockIdx.y * nblk + b];  // same order in every thread
  const double f = s > 0.0 ? 1.0 / sqrt(s * dv) : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < elems; i += (long)gridDim.x * blockDim.x)
    st[i] *= f;
}

}  // namespace fbse

using namespace fbse;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static int layer_dims(const fbse_mlp *m, int l, int *k_real, int *n_real) {
  *k_real = l == 0 ? m->n_in : m->width;
  *n_real = l == 3 ? m->n_out : m->width;
  return 0;
}

static int mlp_pack(fbse_mlp *m) {
  for (int l = 0; l < 4; ++l) {
    int kr, nr;
    layer_dims(m, l, &kr, &nr);
    hipLaunchKernelGGL(pack_layer_kernel, dim3(64), dim3(256), 0, m->stream, m->d_params + m->off_w[l],
                       m->d_params + m->off_b[l], kr, nr, m->lk[l], m->ln[l], m->d_wb[l], m->d_bias[l]);
  }
  MLP_HIP(m, hipGetLastError());
  return FBSE_OK;
}

extern "C" int fbse_mlp_create(int device, int n_in, int width, int n_out, unsigned long long seed,
                               fbse_mlp **out) {
  if (!out || n_in < 1 || n_in > 64 || width < 16 || width > 128 || width % 16 || n_out < 1 || n_out > 16)
    return FBSE_ERR_ARG;
  if (width != 32 && width != 64 && width != 128) return FBSE_ERR_UNSUPPORTED;
  fbse_mlp *m = new fbse_mlp();
  m->device = device;
  m->n_in = n_in;
  m->k_in = (n_in + 15) / 16 * 16;
  m->width = width;
  m->n_out = n_out;
  const int kr[4] = {n_in, width, width, width}, nr[4] = {width, width, width, n_out};
  long off = 0;
  for (int l = 0; l < 4; ++l) {
    m->off_w[l] = off;
    off += (long)kr[l] * nr[l];
    m->off_b[l] = off;
    off += nr[l];
    m->lk[l] = l == 0 ? m->k_in : width;
    m->ln[l] = l == 3 ? 16 : width;
  }
  m->n_params = off;
  auto fail = [&](int rc) {
    fbse_mlp_destroy(m);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(FBSE_ERR_HIP);
  m->stream = m->own_stream;
  const size_t pb = (size_t)m->n_params * sizeof(float);
  if (hipMalloc(&m->d_params, pb) != hipSuccess || hipMalloc(&m->d_grads, pb) != hipSuccess ||
      hipMalloc(&m->d_m, pb) != hipSuccess || hipMalloc(&m->d_v, pb) != hipSuccess ||
      hipMalloc(&m->d_red, 2 * (2 + 2 * 1024) * sizeof(double)) != hipSuccess)
    return fail(FBSE_ERR_NOMEM);
  for (int l = 0; l < 4; ++l)
    if (hipMalloc(&m->d_wb[l], (size_t)m->lk[l] * m->ln[l] * 2) != hipSuccess ||
        hipMalloc(&m->d_bias[l], (size_t)m->ln[l] * sizeof(float)) != hipSuccess)
      return fail(FBSE_ERR_NOMEM);
  // Glorot-uniform weights from a 64-bit LCG (host), biases 0
  std::vector<float> h(m->n_params, 0.f);
  unsigned long long st = seed * 6364136223846793005ull + 1442695040888963407ull;
  for (int l = 0; l < 4; ++l) {
    const double lim = std::sqrt(6.0 / (double)(kr[l] + nr[l]));
    for (long i = 0; i < (long)kr[l] * nr[l]; ++i) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      const double u = (double)(st >> 11) * (1.0 / 9007199254740992.0);
      h[m->off_w[l] + i] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  if (fbse_mlp_set_params(m, h.data()) != FBSE_OK) return fail(FBSE_ERR_HIP);
  *out = m;
  return FBSE_OK;
}

extern "C" void fbse_mlp_destroy(fbse_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->d_params, m->d_grads, m->d_m, m->d_v, m->d_y, m->d_dz, m->d_dz2, m->d_part, m->d_red};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int l = 0; l < 4; ++l) {
    if (m->d_wb[l]) (void)hipFree(m->d_wb[l]);
    if (m->d_bias[l]) (void)hipFree(m->d_bias[l]);
    if (m->d_act[l]) (void)hipFree(m->d_act[l]);
    if (m->d_z[l]) (void)hipFree(m->d_z[l]);
  }
  if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
  delete m;
}
extern "C" const char *fbse_mlp_last_error(const fbse_mlp *m) { return m ? m->err.c_str() : "null network"; }
extern "C" int fbse_mlp_set_stream(fbse_mlp *m, void *stream) {
  if (!m) return FBSE_ERR_ARG;
  m->stream = stream ? (hipStream_t)stream : m->own_stream;
  return FBSE_OK;
}
extern "C" int fbse_mlp_input_width(const fbse_mlp *m) { return m ? m->k_in : -1; }
extern "C" long fbse_mlp_num_params(const fbse_mlp *m) { return m ? m->n_params : -1; }
extern "C" int fbse_mlp_get_params(fbse_mlp *m, float *params) {
  if (!m || !params) return FBSE_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->device));
  MLP_HIP(m, hipMemcpyAsync(params, m->d_params, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return FBSE_OK;
}
extern "C" int fbse_mlp_set_params(fbse_mlp *m, const float *params) {
  if (!m || !params) return FBSE_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->device));
  MLP_HIP(m, hipMemcpyAsync(m->d_params, params, m->n_params * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(m, hipMemsetAsync(m->d_m, 0, m->n_params * sizeof(float), m->stream));
  MLP_HIP(m, hipMemsetAsync(m->d_v, 0, m->n_params * sizeof(float), m->stream));
  m->adam_t = 0;
  int rc = mlp_pack(m);
  if (rc) return rc;
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return FBSE_OK;
}

static int ensure_train(fbse_mlp *m, long n) {
  if (n <= m->cap) return FBSE_OK;
  for (int l = 0; l < 3; ++l) {
    if (m->d_act[l]) (void)hipFree(m->d_act[l]);
    if (m->d_z[l]) (void)hipFree(m->d_z[l]);
    m->d_act[l] = nullptr;
    m->d_z[l] = nullptr;
  }
  if (m->d_y) (void)hipFree(m->d_y);
  if (m->d_dz) (void)hipFree(m->d_dz);
  if (m->d_dz2) (void)hipFree(m->d_dz2);
  m->d_y = nullptr;
  m->d_dz = m->d_dz2 = nullptr;
  m->cap = 0;
  const size_t w = (size_t)n * m->width;
  for (int l = 0; l < 3; ++l)
    if (hipMalloc(&m->d_act[l], w * 2) != hipSuccess || hipMalloc(&m->d_z[l], w * 4) != hipSuccess)
      return mlp_err(m, FBSE_ERR_NOMEM, "activations for %ld points", n);
  if (hipMalloc(&m->d_y, (size_t)n * 16 * 4) != hipSuccess || hipMalloc(&m->d_dz, w * 2) != hipSuccess ||
      hipMalloc(&m->d_dz2, w * 2) != hipSuccess)
    return mlp_err(m, FBSE_ERR_NOMEM, "outputs for %ld points", n);
  m->cap = n;
  return FBSE_OK;
}

// launchers with the column-tile count as a template argument
static void launch_fwd(fbse_mlp *m, int l, const unsigned short *x, long n, unsigned short *yb, float *z,
                       float *yf) {
  const dim3 g((unsigned)((n + kMlpPoints - 1) / kMlpPoints)), b(kMlpThreads);
  const size_t lds = (size_t)m->lk[l] * m->ln[l] * 2;
  const int K = m->lk[l];
  switch (m->ln[l] / 16) {
    case 1: hipLaunchKernelGGL((dense_fwd_kernel<1, true>), g, b, lds, m->stream, x, K, m->d_wb[l], m->d_bias[l], n, yb, z, yf); break;
    case 2: hipLaunchKernelGGL((dense_fwd_kernel<2, false>), g, b, lds, m->stream, x, K, m->d_wb[l], m->d_bias[l], n, yb, z, yf); break;
    case 4: hipLaunchKernelGGL((dense_fwd_kernel<4, false>), g, b, lds, m->stream, x, K, m->d_wb[l], m->d_bias[l], n, yb, z, yf); break;
    default: hipLaunchKernelGGL((dense_fwd_kernel<8, false>), g, b, lds, m->stream, x, K, m->d_wb[l], m->d_bias[l], n, yb, z, yf); break;
  }
}
static void launch_bwd_data(fbse_mlp *m, int l, const unsigned short *dz, const float *zprev, long n,
                            unsigned short *dzprev) {
  const dim3 g((unsigned)((n + kMlpPoints - 1) / kMlpPoints)), b(kMlpThreads);
  const size_t lds = (size_t)m->lk[l] * m->ln[l] * 2;
  const int N = m->ln[l];
  switch (m->lk[l] / 16) {
    case 2: hipLaunchKernelGGL((dense_bwd_data_kernel<2>), g, b, lds, m->stream, dz, N, m->d_wb[l], zprev, n, dzprev); break;
    case 4: hipLaunchKernelGGL((dense_bwd_data_kernel<4>), g, b, lds, m->stream, dz, N, m->d_wb[l], zprev, n, dzprev); break;
    default: hipLaunchKernelGGL((dense_bwd_data_kernel<8>), g, b, lds, m->stream, dz, N, m->d_wb[l], zprev, n, dzprev); break;
  }
}
template <int KT>
static void launch_wgrad_k(fbse_mlp *m, int nt, dim3 g, const unsigned short *x, const unsigned short *dz, long n) {
  switch (nt) {
    case 1: hipLaunchKernelGGL((dense_wgrad_kernel<KT, 1>), g, dim3(kMlpThreads), 0, m->stream, x, dz, n, m->d_part); break;
    case 2: hipLaunchKernelGGL((dense_wgrad_kernel<KT, 2>), g, dim3(kMlpThreads), 0, m->stream, x, dz, n, m->d_part); break;
    case 4: hipLaunchKernelGGL((dense_wgrad_kernel<KT, 4>), g, dim3(kMlpThreads), 0, m->stream, x, dz, n, m->d_part); break;
    default: hipLaunchKernelGGL((dense_wgrad_kernel<KT, 8>), g, dim3(kMlpThreads), 0, m->stream, x, dz, n, m->d_part); break;
  }
}
static int launch_wgrad(fbse_mlp *m, int l, const unsigned short *x, const unsigned short *dz, long n) {
  // a fixed number of work-groups (so the summation order does not depend on
  // the device), each over many chunks
  const long nblk = std::min((n + kWgChunk - 1) / kWgChunk, (long)kWgBlocks);
  const long need = nblk * ((long)m->lk[l] * m->ln[l] + m->ln[l]);
  if (need > m->part_cap) {
    if (m->d_part) (void)hipFree(m->d_part);
    m->d_part = nullptr;
    m->part_cap = 0;
    MLP_HIP(m, hipMalloc(&m->d_part, (size_t)need * sizeof(float)));
    m->part_cap = need;
  }
  const dim3 g((unsigned)nblk);
  const int nt = m->ln[l] / 16;
  switch (m->lk[l] / 16) {
    case 1: launch_wgrad_k<1>(m, nt, g, x, dz, n); break;
    case 2: launch_wgrad_k<2>(m, nt, g, x, dz, n); break;
    case 3: launch_wgrad_k<3>(m, nt, g, x, dz, n); break;
    case 4: launch_wgrad_k<4>(m, nt, g, x, dz, n); break;
    default: launch_wgrad_k<8>(m, nt, g, x, dz, n); break;
  }
  int kr, nr;
  layer_dims(m, l, &kr, &nr);
  const int stride = m->lk[l] * m->ln[l] + m->ln[l];
  hipLaunchKernelGGL(reduce_part_kernel, dim3((stride + 255) / 256), dim3(256), 0, m->stream, m->d_part, (int)nblk,
                     m->lk[l], m->ln[l], kr, nr, m->d_grads + m->off_w[l], m->d_grads + m->off_b[l]);
  MLP_HIP(m, hipGetLastError());
  return FBSE_OK;
}

static int forward_all(fbse_mlp *m, const unsigned short *x, long n) {
  launch_fwd(m, 0, x, n, m->d_act[0], m->d_z[0], nullptr);
  launch_fwd(m, 1, m->d_act[0], n, m->d_act[1], m->d_z[1], nullptr);
  launch_fwd(m, 2, m->d_act[1], n, m->d_act[2], m->d_z[2], nullptr);
  launch_fwd(m, 3, m->d_act[2], n, nullptr, nullptr, m->d_y);
  MLP_HIP(m, hipGetLastError());
  return FBSE_OK;
}

extern "C" int fbse_mlp_forward_dev(fbse_mlp *m, const unsigned short *d_x, long n, float *d_y) {
  if (!m || !d_x || !d_y || n < 0) return FBSE_ERR_ARG;
  if (n == 0) return FBSE_OK;
  MLP_HIP(m, hipSetDevice(m->device));
  // inference in chunks of the training buffers
  const long ch = std::min(n, std::max(m->cap, 1L << 20));
  int rc = ensure_train(m, ch);
  if (rc) return rc;
  for (long p0 = 0; p0 < n; p0 += ch) {
    const long np = std::min(ch, n - p0);
    if ((rc = forward_all(m, d_x + p0 * m->k_in, np))) return rc;
    // outputs: n_out of the 16 padded columns
    MLP_HIP(m, hipMemcpy2DAsync(d_y + p0 * m->n_out, m->n_out * sizeof(float), m->d_y, 16 * sizeof(float),
                                m->n_out * sizeof(float), np, hipMemcpyDeviceToDevice, m->stream));
  }
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return FBSE_OK;
}

extern "C" int fbse_mlp_loss_grad_dev(fbse_mlp *m, const unsigned short *d_x, const float *d_t, const float *d_w,
                                      long n, double *loss, float *grads) {
  if (!m || !d_x || !d_t || n < 1) return FBSE_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->device));
  int rc = ensure_train(m, n);
  if (rc) return rc;
  if ((rc = forward_all(m, d_x, n))) return rc;
  const int nblk = 1024;
  hipLaunchKernelGGL(loss_sums_kernel, dim3(nblk), dim3(kMlpThreads), 0, m->stream, m->d_y, d_t, d_w, n, m->n_out,
                     16, m->d_red);
  hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(64), 0, m->stream, m->d_red, nblk);
  hipLaunchKernelGGL(loss_grad_kernel, dim3(2048), dim3(256), 0, m->stream, m->d_y, d_t, d_w, n, m->n_out, 16,
                     m->d_red, m->d_dz);
  MLP_HIP(m, hipGetLastError());
  // backward: layer 3 (linear output, dZ in d_dz, n x 16)
  if ((rc = launch_wgrad(m, 3, m->d_act[2], m->d_dz, n))) return rc;
  launch_bwd_data(m, 3, m->d_dz, m->d_z[2], n, m->d_dz2);  // dZ of layer 2 (n x width)
  for (int l = 2; l >= 1; --l) {
    if ((rc = launch_wgrad(m, l, m->d_act[l - 1], m->d_dz2, n))) return rc;
    launch_bwd_data(m, l, m->d_dz2, m->d_z[l - 1], n, m->d_dz);
    std::swap(m->d_dz, m->d_dz2);
  }
  if ((rc = launch_wgrad(m, 0, d_x, m->d_dz2, n))) return rc;
  MLP_HIP(m, hipGetLastError());
  double red[2] = {0, 0};
  MLP_HIP(m, hipMemcpyAsync(red, m->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grads)
    MLP_HIP(m, hipMemcpyAsync(grads, m->d_grads, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  if (loss) *loss = red[0] > 0.0 ? red[1] / red[0] : 0.0;
  return FBSE_OK;
}

extern "C" int fbse_mlp_adam_step(fbse_mlp *m, float lr) {
  if (!m || !(lr > 0.f)) return FBSE_ERR_ARG;
  MLP_HIP(m, hipSetDevice(m->device));
  ++m->adam_t;
  const float c1 = 1.0f - std::pow(0.9f, (float)m->adam_t), c2 = 1.0f - std::pow(0.999f, (float)m->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, m->stream, m->d_params, m->d_grads, m->d_m, m->d_v,
                     m->n_params, lr, c1, c2);
  MLP_HIP(m, hipGetLastError());
  return mlp_pack(m);
}

extern "C" int fbse_mlp_features_dev(fbse_mlp *m, long nx, long ny, long nz, const double *origin, double dx,
                                     const double *atoms_xyz, const int *species, long n_atoms, int n_species,
                                     const double *d_vloc, const double *center, double r_scale,
                                     unsigned short *d_x) {
  if (!m || !origin || !atoms_xyz || !species || n_atoms < 1 || n_species < 1 || n_species > kMaxSpecies ||
      !d_vloc || !center || !(r_scale > 0.0) || !d_x || nx < 1 || ny < 1 || nz < 1)
    return FBSE_ERR_ARG;
  if (kNearest * n_species + 2 > m->k_in)
    return mlp_err(m, FBSE_ERR_ARG, "%d species need %d inputs, the network has %d", n_species,
                   kNearest * n_species + 2, m->k_in);
  MLP_HIP(m, hipSetDevice(m->device));
  std::vector<float4> at(n_atoms);
  for (long a = 0; a < n_atoms; ++a) {
    if (species[a] < 0 || species[a] >= n_species) return mlp_err(m, FBSE_ERR_ARG, "atom %ld species %d", a, species[a]);
    at[a] = make_float4((float)atoms_xyz[3 * a], (float)atoms_xyz[3 * a + 1], (float)atoms_xyz[3 * a + 2],
                        (float)species[a]);
  }
  float4 *d_at = nullptr;
  MLP_HIP(m, hipMalloc(&d_at, n_atoms * sizeof(float4)));
  hipError_t e = hipMemcpyAsync(d_at, at.data(), n_atoms * sizeof(float4), hipMemcpyHostToDevice, m->stream);
  const long ng = nx * ny * nz;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, m->stream, nx, ny, nz,
                       origin[0], origin[1], origin[2], dx, d_at, n_atoms, n_species, d_vloc, center[0], center[1],
                       center[2], 1.0 / r_scale, m->k_in, d_x);
    e = hipGetLastError();
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(d_at);
  if (e != hipSuccess) return mlp_err(m, FBSE_ERR_HIP, "features: %s", hipGetErrorString(e));
  return FBSE_OK;
}

extern "C" int fbse_state_density_dev(fbse_ctx *c, const double *d_psi, const long *idx, long n, double *d_rho) {
  if (!c || !d_psi || (n > 0 && !idx) || n < 0 || !d_rho) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = c->nsg * (c->is_complex + 1);
  int rc = fill_zero(c, d_rho, (size_t)c->ngrid * sizeof(double));
  if (rc) return rc;
  for (long i = 0; i < n; ++i) {
    if (idx[i] < 0) return set_err(c, FBSE_ERR_ARG, "state index %ld", idx[i]);
    hipLaunchKernelGGL(density_add_kernel, dim3(1024), dim3(256), 0, c->stream, d_psi + idx[i] * sl, c->ngrid,
                       (int)c->nspin, c->is_complex, d_rho);
  }
  FBSE_HIP(c, hipGetLastError());
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  return FBSE_OK;
}

extern "C" int fbse_warm_start_dev(fbse_ctx *c, const float *d_y, int ld, int channel, double smooth_ep,
                                   double *d_psi_rank, long n_starts, long ms) {
  if (!c || !d_y || ld < 1 || channel < 0 || channel >= ld || !d_psi_rank || n_starts < 0 || ms < 1 ||
      smooth_ep < 0.0)
    return FBSE_ERR_ARG;
  if (n_starts == 0) return FBSE_OK;
  if (n_starts > 65535) return set_err(c, FBSE_ERR_UNSUPPORTED, "%ld start states in one call", n_starts);
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = c->nsg * (c->is_complex + 1);
  if (smooth_ep > 0.0) {  // low-pass the start states first (one state at a time: they are ms*stlen apart)
    for (long j = 0; j < n_starts; ++j) {
      int rc = precondition_dev(c, d_psi_rank + j * ms * sl, d_psi_rank + j * ms * sl, 1, smooth_ep);
      if (rc) return rc;
    }
  }
  double *part = nullptr;
  FBSE_HIP(c, hipMalloc(&part, (size_t)n_starts * kWarmBlocks * sizeof(double)));
  hipLaunchKernelGGL(warm_envelope_kernel, dim3(kWarmBlocks, (unsigned)n_starts), dim3(256), 0, c->stream, d_y, ld,
                     channel, d_psi_rank, ms * sl, c->ngrid, (int)c->nspin, c->is_complex, part);
  hipLaunchKernelGGL(warm_scale_kernel, dim3(512, (unsigned)n_starts), dim3(256), 0, c->stream, d_psi_rank, ms * sl,
                     sl, kWarmBlocks, part, c->dv);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(part);
  if (e != hipSuccess) return set_err(c, FBSE_ERR_HIP, "warm start: %s", hipGetErrorString(e));
  return FBSE_OK;
}
