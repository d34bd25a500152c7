// gemm_kernels.hip -- FP64 matrix-core (v_mfma_f64_16x16x4f64) products of
// the subspace step: the overlap / Hamiltonian Grams C = A^H B of tall state
// blocks (orthogonalisation, ortho.c:54,150; Rayleigh-Ritz matrix, Hmat.c:141-164)
// and the rotations Out = A V of a state block by a small matrix (ortho's
// U = A V Sigma^-1, Rayleigh-Ritz's psi_new = psi Y, Hmat.c:230-281).
//
// States are columns of m grid values (complex: double2 re,im; real: double),
// leading dimension lda.  Both kernels stage 16-row (Gram) or 16-state
// (rotation) slices of 64 columns in LDS, double buffered, and give each of
// four waves a 32 x 32 block of the 64 x 64 output tile as 2 x 2 MFMA tiles of
// 16 x 16.  A complex product takes four real MFMAs per 16 x 16 x 4 step
// (re*re, im*im into the real part, re*im, -im*re into the imaginary part).
//
// Lane layout of v_mfma_f64_16x16x4f64 (64 lanes, l):
//   A operand (16 x 4):  row l % 16, k = l / 16
//   B operand (4 x 16):  k = l / 16, column l % 16
//   D (16 x 16), 4 values v: column l % 16, row 4 v + l / 16   (measured on gfx950)
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fbse_internal.h"

namespace fbse {

typedef double dbl4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kT = 64;   // output tile edge
constexpr int kK = 16;   // reduction slice staged per iteration
constexpr int kThreads = 256;

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <bool CPLX> using elem_t = typename std::conditional<CPLX, double2, double>::type;

template <bool CPLX> __device__ __forceinline__ elem_t<CPLX> zero_elem();
template <> __device__ __forceinline__ double2 zero_elem<true>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ double zero_elem<false>() { return 0.0; }

// part[p*pstride + i + j*ldp] = sum over the rows of slab p of conj(A[r, i])
// B[r, j] (i < n1, j < n2), one 64 x 64 output tile per work-group and one
// slab per grid row y.  Rows of slab p: [p*slab, min(m, (p+1)*slab)).
template <bool CPLX>
__global__ __launch_bounds__(kThreads, 2) void gram_mfma_kernel(const elem_t<CPLX> *__restrict__ A, long lda,
                                                                const elem_t<CPLX> *__restrict__ B, long ldb, long m,
                                                                int n1, int n2, long slab, int upper,
                                                                elem_t<CPLX> *__restrict__ part, long ldp,
                                                                long pstride) {
  using E = elem_t<CPLX>;
  __shared__ E As[2][kK][kT + 1];
  __shared__ E Bs[2][kK][kT + 1];
  // output tile (ta, tb): column-major over the tile grid, or, for a
  // Hermitian product (upper), the tiles with ta <= tb in column order
  int ta, tb;
  {
    const int t1 = (n1 + kT - 1) / kT;
    const int tl = blockIdx.x;
    if (upper) {
      int b = (int)((sqrtf(8.0f * tl + 1.0f) - 1.0f) * 0.5f);
      while (b * (b + 1) / 2 > tl) --b;
      while ((b + 1) * (b + 2) / 2 <= tl) ++b;
      tb = b;
      ta = tl - b * (b + 1) / 2;
    } else {
      ta = tl % t1;
      tb = tl / t1;
    }
  }
  const int i0 = ta * kT, j0 = tb * kT;
  const long r_begin = (long)blockIdx.y * slab;
  const long r_end = min(m, r_begin + slab);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
  // loader: element e = t + 256 q -> column e / 16, row e % 16 (16 threads
  // read 16 consecutive rows of one column)
  const int lc = t >> 4, lk = t & 15;
  E ra[4], rb[4];
  auto load = [&](long r0) {
    const long r = r0 + lk;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lc + 16 * q;
      ra[q] = (r < r_end && i0 + c < n1) ? A[(long)(i0 + c) * lda + r] : zero_elem<CPLX>();
      rb[q] = (r < r_end && j0 + c < n2) ? B[(long)(j0 + c) * ldb + r] : zero_elem<CPLX>();
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][lk][lc + 16 * q] = ra[q];
      Bs[buf][lk][lc + 16 * q] = rb[q];
    }
  };
  dbl4 cre[2][2], cim[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cre[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
      cim[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
  const long nchunk = (r_end - r_begin + kK - 1) / kK;
  if (nchunk > 0) {
    load(r_begin);
    store(0);
  }
  __syncthreads();
  const int fl = lane & 15, fk = lane >> 4;
  for (long ch = 0; ch < nchunk; ++ch) {
    const int buf = (int)(ch & 1);
    if (ch + 1 < nchunk) load(r_begin + (ch + 1) * kK);
#pragma unroll
    for (int g = 0; g < kK / 4; ++g) {
      const int k = g * 4 + fk;
      E a[2], b[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        a[s] = As[buf][k][wi + s * 16 + fl];
        b[s] = Bs[buf][k][wj + s * 16 + fl];
      }
      if constexpr (CPLX) {
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) {
            cre[si][sj] = mfma(a[si].x, b[sj].x, cre[si][sj]);
            cim[si][sj] = mfma(a[si].x, b[sj].y, cim[si][sj]);
          }
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) {
            cre[si][sj] = mfma(a[si].y, b[sj].y, cre[si][sj]);
            cim[si][sj] = mfma(-a[si].y, b[sj].x, cim[si][sj]);
          }
      } else {
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) cre[si][sj] = mfma(a[si], b[sj], cre[si][sj]);
      }
    }
    if (ch + 1 < nchunk) store(buf ^ 1);
    __syncthreads();
  }
  E *out = part + (size_t)blockIdx.y * pstride;
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
      const int j = j0 + wj + sj * 16 + fl;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + si * 16 + 4 * v + fk;
        if (i < n1 && j < n2) {
          if constexpr (CPLX)
            out[i + (long)j * ldp] = make_double2(cre[si][sj][v], cim[si][sj][v]);
          else
            out[i + (long)j * ldp] = cre[si][sj][v];
        }
      }
    }
}

// Out[r, j] (+)= sum_k A[r, k] V[k, j] for r < nrows, j < nout, k < nin (added
// to Out when accumulate is set).  The
// MFMA computes the transposed tile (V^T A^T), so that a lane's four outputs
// are one grid row of four columns and 16 lanes write 16 consecutive rows.
template <bool CPLX>
__global__ __launch_bounds__(kThreads, 2) void rot_mfma_kernel(const elem_t<CPLX> *__restrict__ A, long lda,
                                                               long nrows, int nin,
                                                               const elem_t<CPLX> *__restrict__ V, long ldv,
                                                               int nout, elem_t<CPLX> *__restrict__ Out, long ldo,
                                                               int accumulate) {
  using E = elem_t<CPLX>;
  __shared__ E As[2][kK][kT];
  __shared__ E Vs[2][kK][kT + 1];
  // one-dimensional grid, column tiles fastest: the work-groups of one row
  // tile run together and share its rows of A in the cache
  const int nct = (nout + kT - 1) / kT;
  const int j0 = (int)(blockIdx.x % nct) * kT;
  const long r0 = (long)(blockIdx.x / nct) * kT;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
  // loaders: A slice (16 states x 64 rows): e -> state e / 64, row e % 64;
  // V slice (16 k x 64 columns): e -> column e / 16, k e % 16
  E ra[4], rv[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + kThreads * q;
      const int ak = e >> 6, ar = e & 63;
      ra[q] = (k0 + ak < nin && r0 + ar < nrows) ? A[(long)(k0 + ak) * lda + r0 + ar] : zero_elem<CPLX>();
      const int vj = e >> 4, vk = e & 15;
      rv[q] = (k0 + vk < nin && j0 + vj < nout) ? V[(long)(j0 + vj) * ldv + k0 + vk] : zero_elem<CPLX>();
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + kThreads * q;
      As[buf][e >> 6][e & 63] = ra[q];
      Vs[buf][e & 15][e >> 4] = rv[q];
    }
  };
  dbl4 cre[2][2], cim[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cre[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
      cim[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
  const int nchunk = (nin + kK - 1) / kK;
  load(0);
  store(0);
  __syncthreads();
  const int fl = lane & 15, fk = lane >> 4;
  for (int ch = 0; ch < nchunk; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nchunk) load((ch + 1) * kK);
#pragma unroll
    for (int g = 0; g < kK / 4; ++g) {
      const int k = g * 4 + fk;
      E a[2], v[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        a[s] = As[buf][k][wi + s * 16 + fl];  // B operand: A^T[k][row]
        v[s] = Vs[buf][k][wj + s * 16 + fl];  // A operand: V^T[col][k]
      }
      if constexpr (CPLX) {
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) {
            cre[si][sj] = mfma(v[sj].x, a[si].x, cre[si][sj]);
            cim[si][sj] = mfma(v[sj].y, a[si].x, cim[si][sj]);
          }
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) {
            cre[si][sj] = mfma(-v[sj].y, a[si].y, cre[si][sj]);
            cim[si][sj] = mfma(v[sj].x, a[si].y, cim[si][sj]);
          }
      } else {
#pragma unroll
        for (int si = 0; si < 2; ++si)
#pragma unroll
          for (int sj = 0; sj < 2; ++sj) cre[si][sj] = mfma(v[sj], a[si], cre[si][sj]);
      }
    }
    if (ch + 1 < nchunk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int si = 0; si < 2; ++si)
#pragma unroll
    for (int sj = 0; sj < 2; ++sj) {
      const long r = r0 + wi + si * 16 + fl;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = j0 + wj + sj * 16 + 4 * v + fk;
        if (r < nrows && j < nout) {
          E val;
          if constexpr (CPLX)
            val = make_double2(cre[si][sj][v], cim[si][sj][v]);
          else
            val = cre[si][sj][v];
          if (accumulate) val += Out[(long)j * ldo + r];
          Out[(long)j * ldo + r] = val;
        }
      }
    }
}

}  // namespace

long gram_tile_count(long n1, long n2, bool upper) {
  const long t1 = (n1 + kT - 1) / kT, t2 = (n2 + kT - 1) / kT;
  return upper ? t1 * (t1 + 1) / 2 : t1 * t2;
}

int launch_gram_mfma(hipStream_t s, bool cplx, const double *A, long lda, const double *B, long ldb, long m, long n1,
                     long n2, bool upper, long slab, long parts, double *part, long ldp, long pstride) {
  if (upper && n1 != n2) return -1;
  const long tiles = gram_tile_count(n1, n2, upper);
  if (tiles > 0x7fffffffL || parts > 65535) return -1;
  const dim3 grid((unsigned)tiles, (unsigned)parts);
  if (cplx)
    hipLaunchKernelGGL(gram_mfma_kernel<true>, grid, dim3(kThreads), 0, s, (const double2 *)A, lda,
                       (const double2 *)B, ldb, m, (int)n1, (int)n2, slab, upper ? 1 : 0, (double2 *)part, ldp,
                       pstride);
  else
    hipLaunchKernelGGL(gram_mfma_kernel<false>, grid, dim3(kThreads), 0, s, A, lda, B, ldb, m, (int)n1, (int)n2,
                       slab, upper ? 1 : 0, part, ldp, pstride);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_rot_mfma(hipStream_t s, bool cplx, const double *A, long lda, long nrows, long nin, const double *V,
                    long ldv, long nout, double *Out, long ldo, bool accumulate) {
  const long tiles = ((nout + kT - 1) / kT) * ((nrows + kT - 1) / kT);
  if (tiles > 0x7fffffffL) return -1;
  const dim3 grid((unsigned)tiles);
  if (cplx)
    hipLaunchKernelGGL(rot_mfma_kernel<true>, grid, dim3(kThreads), 0, s, (const double2 *)A, lda, nrows, (int)nin,
                       (const double2 *)V, ldv, (int)nout, (double2 *)Out, ldo, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL(rot_mfma_kernel<false>, grid, dim3(kThreads), 0, s, A, lda, nrows, (int)nin, V, ldv,
                       (int)nout, Out, ldo, accumulate ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace fbse
