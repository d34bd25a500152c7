// fft_kernels.hip -- batched 3D complex FFT for gfx950, one kernel per axis pass.
//
// A work-group loads a tile of `lines` 1D lines of length n into LDS, runs a
// Stockham self-sorting mixed-radix transform there (ping-pong between two LDS
// buffers, radices 2/3/4/5/8 hard-coded, any other prime by direct DFT), and
// writes the tile back.  x lines are contiguous rows; y and z tiles take
// `lines` consecutive inner indices so every row of the tile is a contiguous
// global segment.  Fused stages replace separate HBM passes:
//   z pass  : forward z FFT -> x 0.5|k|^2/N (ksqr) -> backward z FFT
//             (kinetic of hamiltonian.c:62-99, FFTW's unnormalised pair)
//   x pass  : backward x FFT -> out = a1*(T psi + V psi_in) + a0*psi_in
//             (V_loc of phamiltonian.c:121-132 and the Newton shift of
//              filter.c:402-407 in the same sweep)
// Sign convention is FFTW's: forward exp(-2 pi i jk/n), backward exp(+...).
#include "fbse_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>

namespace fbse {

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C> __device__ __forceinline__ C cconj(C a) {
  a.y = -a.y;
  return a;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b) {
  a.x -= b.x;
  a.y -= b.y;
  return a;
}
// multiply by s*i (s = +-1)
template <typename C, typename T> __device__ __forceinline__ C cmul_i(C a, T s) {
  C r;
  r.x = -s * a.y;
  r.y = s * a.x;
  return r;
}

// Butterfly of radix R on values x[0..R-1] (already twiddled); sgn = -1 forward.
template <typename T, int R> struct Bfly;

template <typename T> struct Bfly<T, 2> {
  using C = cplx_t<T>;
  __device__ static void run(C *x, T) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};
template <typename T> struct Bfly<T, 3> {
  using C = cplx_t<T>;
  __device__ static void run(C *x, T sgn) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    C t1 = cadd(x[1], x[2]);
    C t2;
    t2.x = x[0].x - (T)0.5 * t1.x;
    t2.y = x[0].y - (T)0.5 * t1.y;
    C d = csub(x[1], x[2]);
    C t3 = cmul_i<C, T>(d, sgn * h);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, t3);
    x[2] = csub(t2, t3);
  }
};
template <typename T> struct Bfly<T, 4> {
  using C = cplx_t<T>;
  __device__ static void run(C *x, T sgn) {
    C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    C c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
    C di = cmul_i<C, T>(d, sgn);
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, di);
    x[3] = csub(b, di);
  }
};
template <typename T> struct Bfly<T, 5> {
  using C = cplx_t<T>;
  __device__ static void run(C *x, T sgn) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    C t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    C t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
    C a1, a2, u1, u2;
    a1.x = x[0].x + c1 * t1.x + c2 * t2.x;
    a1.y = x[0].y + c1 * t1.y + c2 * t2.y;
    a2.x = x[0].x + c2 * t1.x + c1 * t2.x;
    a2.y = x[0].y + c2 * t1.y + c1 * t2.y;
    u1.x = s1 * t3.x + s2 * t4.x;
    u1.y = s1 * t3.y + s2 * t4.y;
    u2.x = s2 * t3.x - s1 * t4.x;
    u2.y = s2 * t3.y - s1 * t4.y;
    C b1 = cmul_i<C, T>(u1, sgn), b2 = cmul_i<C, T>(u2, sgn);
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, b1);
    x[4] = csub(a1, b1);
    x[2] = cadd(a2, b2);
    x[3] = csub(a2, b2);
  }
};
template <typename T> struct Bfly<T, 8> {
  using C = cplx_t<T>;
  __device__ static void run(C *x, T sgn) {
    C e[4] = {x[0], x[2], x[4], x[6]};
    C o[4] = {x[1], x[3], x[5], x[7]};
    Bfly<T, 4>::run(e, sgn);
    Bfly<T, 4>::run(o, sgn);
    const T r = (T)0.70710678118654752440;
    // w^1 = r(1 + sgn i), w^2 = sgn i, w^3 = r(-1 + sgn i)
    C w1o, w2o, w3o;
    w1o.x = r * (o[1].x - sgn * o[1].y);
    w1o.y = r * (o[1].y + sgn * o[1].x);
    w2o = cmul_i<C, T>(o[2], sgn);
    w3o.x = r * (-o[3].x - sgn * o[3].y);
    w3o.y = r * (-o[3].y + sgn * o[3].x);
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], w1o);
    x[5] = csub(e[1], w1o);
    x[2] = cadd(e[2], w2o);
    x[6] = csub(e[2], w2o);
    x[3] = cadd(e[3], w3o);
    x[7] = csub(e[3], w3o);
  }
};

// One Stockham pass of compile-time radix R over all lines of the tile.
template <typename T, int R>
__device__ __forceinline__ void pass_fixed(const cplx_t<T> *in, cplx_t<T> *out, const cplx_t<T> *tw, int n,
                                           int p, int LS, int nlines, bool inv) {
  using C = cplx_t<T>;
  const int m = n / R;
  const int tstep = n / (p * R);
  const int nb = m * nlines;
  const T sgn = inv ? (T)1 : (T)-1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const int l = b / m;
    const int i = b - l * m;
    const int k = i % p;
    const C *src = in + l * LS;
    C *dst = out + l * LS;
    C x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C v = src[i + r * m];
      if (r > 0 && k > 0) {
        C w = tw[r * k * tstep];
        if (inv) w = cconj(w);
        v = cmul(v, w);
      }
      x[r] = v;
    }
    Bfly<T, R>::run(x, sgn);
    const int j = (i - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[j + q * p] = x[q];
  }
}

// Generic radix (any R): direct DFT from LDS, twiddles from the length-n table.
template <typename T>
__device__ void pass_generic(const cplx_t<T> *in, cplx_t<T> *out, const cplx_t<T> *tw, int n, int R, int p,
                             int LS, int nlines, bool inv) {
  using C = cplx_t<T>;
  const int m = n / R;
  const int tstep = n / (p * R);
  const int rstep = n / R;
  const int nb = m * nlines;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const int l = b / m;
    const int i = b - l * m;
    const int k = i % p;
    const C *src = in + l * LS;
    C *dst = out + l * LS;
    const int j = (i - k) * R + k;
    for (int q = 0; q < R; ++q) {
      C acc;
      acc.x = 0;
      acc.y = 0;
      for (int r = 0; r < R; ++r) {
        C v = src[i + r * m];
        if (r > 0 && k > 0) {
          C w = tw[r * k * tstep];
          if (inv) w = cconj(w);
          v = cmul(v, w);
        }
        C w2 = tw[((q * r) % R) * rstep];
        if (inv) w2 = cconj(w2);
        acc = cadd(acc, cmul(v, w2));
      }
      dst[j + q * p] = acc;
    }
  }
}

// Runs all radix passes; returns the buffer holding the result.
template <typename T>
__device__ cplx_t<T> *run_passes(cplx_t<T> *a, cplx_t<T> *b, const cplx_t<T> *tw, const AxisArgs<T> &A, int LS,
                                 int nlines, bool inv) {
  int p = 1;
  for (int s = 0; s < A.nf; ++s) {
    const int R = A.f[s];
    switch (R) {
      case 2: pass_fixed<T, 2>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 3: pass_fixed<T, 3>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 4: pass_fixed<T, 4>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 5: pass_fixed<T, 5>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 8: pass_fixed<T, 8>(a, b, tw, A.n, p, LS, nlines, inv); break;
      default: pass_generic<T>(a, b, tw, A.n, R, p, LS, nlines, inv); break;
    }
    __syncthreads();
    cplx_t<T> *t = a;
    a = b;
    b = t;
    p *= R;
  }
  return a;
}

// Global loads a thread keeps in flight while it fills or drains a tile.  The
// batch is held as separate real and imaginary arrays: an array of 16-byte
// complex<double> structs is not split into registers by the compiler and
// would live in scratch memory.
template <typename T>
constexpr int load_batch() { return 8; }

template <typename T, int KIND>
__global__ __launch_bounds__(kFftThreads) void fft_axis_kernel(AxisArgs<T> A) {
  using C = cplx_t<T>;
  constexpr int NB = load_batch<T>();
  extern __shared__ unsigned char smem_raw[];
  const int n = A.n;
  const int L = A.lines;
  const int LS = n | 1;  // odd line stride (in elements) against bank conflicts
  C *buf0 = reinterpret_cast<C *>(smem_raw);
  C *buf1 = buf0 + L * LS;
  C *tw = buf1 + L * LS;

  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = A.tw[t];

  const long tile = blockIdx.x;
  long base, i0 = 0, line0 = 0;
  int nl;
  if (A.xaxis) {
    line0 = tile * L;
    base = line0 * n;
    long rem = A.total_lines - line0;
    nl = rem < L ? (int)rem : L;
  } else {
    const long o = tile / A.tiles_per_outer;
    i0 = (tile - o * A.tiles_per_outer) * L;
    base = o * A.outer_stride + i0;
    long rem = A.inner_count - i0;
    nl = rem < L ? (int)rem : L;
  }

  const int tot = L * n;
  if (A.xaxis) {
    const int totv = nl * n;
    for (int e0 = threadIdx.x; e0 < totv; e0 += kFftThreads * NB) {
      T vr[NB], vi[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        if (e < totv) {
          const C t = A.src[base + (long)e];
          vr[u] = t.x;
          vi[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          buf0[l * LS + p] = C{vr[u], vi[u]};
        }
      }
    }
  } else {
    for (int e0 = threadIdx.x; e0 < tot; e0 += kFftThreads * NB) {
      T vr[NB], vi[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        const int p = e / L, l = e - p * L;
        if (e < tot && l < nl) {
          const C t = A.src[base + (long)p * A.inner + l];
          vr[u] = t.x;
          vi[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        const int p = e / L, l = e - p * L;
        if (e < tot && l < nl) buf0[l * LS + p] = C{vr[u], vi[u]};
      }
    }
  }
  __syncthreads();

  C *res;
  if (KIND == kPassFwd) {
    res = run_passes<T>(buf0, buf1, tw, A, LS, nl, false);
  } else if (KIND == kPassBwd || KIND == kPassBwdCombine) {
    res = run_passes<T>(buf0, buf1, tw, A, LS, nl, true);
  } else {  // kinetic (x k^2) or preconditioner (x g(k^2)): forward, scale, backward
    res = run_passes<T>(buf0, buf1, tw, A, LS, nl, false);
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      const int l = e / n, p = e - l * n;
      const long inner = i0 + l;
      T k2;
      if (A.ksqr) {
        k2 = A.ksqr[(long)p * A.inner + inner];
      } else {
        const long jy = inner / A.nx, jx = inner - jy * A.nx;
        k2 = A.kx[jx] + A.ky[jy] + A.kz[p];
        if (k2 > A.kcap) k2 = A.kcap;
      }
      if (KIND == kPassPrec) k2 = A.pa / ((T)1 + k2 * A.pb);
      C v = res[l * LS + p];
      v.x *= k2;
      v.y *= k2;
      res[l * LS + p] = v;
    }
    __syncthreads();
    C *other = (res == buf0) ? buf1 : buf0;
    res = run_passes<T>(res, other, tw, A, LS, nl, true);
  }

  if (KIND == kPassBwdCombine) {
    const int totv = nl * n;
    for (int e0 = threadIdx.x; e0 < totv; e0 += kFftThreads * NB) {
      T pr[NB], pim[NB];
      T vv[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          const long row = (line0 + l) % A.vloc_rows;
          vv[u] = A.vloc[row * n + p];
          const C t = A.psi_in[base + (long)e];
          pr[u] = t.x;
          pim[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kFftThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          const C t = res[l * LS + p];
          C o;
          o.x = A.a1 * (t.x + vv[u] * pr[u]) + A.a0 * pr[u];
          o.y = A.a1 * (t.y + vv[u] * pim[u]) + A.a0 * pim[u];
          A.dst[base + (long)e] = o;
        }
      }
    }
  } else if (A.xaxis) {
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int l = e / n, p = e - l * n;
      if (l < nl) A.dst[base + (long)e] = res[l * LS + p];
    }
  } else {
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int p = e / L, l = e - p * L;
      if (l < nl) A.dst[base + (long)p * A.inner + l] = res[l * LS + p];
    }
  }
}

constexpr int kIpThreads = 512;  // threads of the in-place kernel

// One in-place Stockham pass: every thread first reads (and twiddles) the R
// inputs of each of its butterflies into registers, the work-group syncs, and
// then the outputs are written to their self-sorted positions.  One LDS buffer
// instead of two lets the tile hold twice the lines.
template <typename T, int R, int KB>
__device__ __forceinline__ void pass_inplace(cplx_t<T> *buf, const cplx_t<T> *tw, int n, int p, int LS, int nlines,
                                             bool inv) {
  using C = cplx_t<T>;
  const int m = n / R;
  const int tstep = n / (p * R);
  const int nb = m * nlines;
  const T sgn = inv ? (T)1 : (T)-1;
  C x[KB][R];
  int dsti[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int b = threadIdx.x + k * kIpThreads;
    dsti[k] = -1;
    if (b < nb) {
      const int l = b / m;
      const int i = b - l * m;
      const int kk = i % p;
      const C *src = buf + l * LS + i;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        C v = src[r * m];
        if (r > 0 && kk > 0) {
          C w = tw[r * kk * tstep];
          if (inv) w = cconj(w);
          v = cmul(v, w);
        }
        x[k][r] = v;
      }
      dsti[k] = l * LS + (i - kk) * R + kk;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    if (dsti[k] >= 0) {
      Bfly<T, R>::run(x[k], sgn);
#pragma unroll
      for (int q = 0; q < R; ++q) buf[dsti[k] + q * p] = x[k][q];
    }
  }
  __syncthreads();
}

// KE = elements of the tile per thread (tile <= kIpThreads * KE elements).
template <typename T, int KE>
__device__ void run_inplace(cplx_t<T> *buf, const cplx_t<T> *tw, const AxisArgs<T> &A, int LS, int nlines, bool inv) {
  int p = 1;
  for (int s = 0; s < A.nf; ++s) {
    const int R = A.f[s];
    switch (R) {
      case 2: pass_inplace<T, 2, (KE + 1) / 2>(buf, tw, A.n, p, LS, nlines, inv); break;
      case 3: pass_inplace<T, 3, (KE + 2) / 3>(buf, tw, A.n, p, LS, nlines, inv); break;
      case 4: pass_inplace<T, 4, (KE + 3) / 4>(buf, tw, A.n, p, LS, nlines, inv); break;
      case 5: pass_inplace<T, 5, (KE + 4) / 5>(buf, tw, A.n, p, LS, nlines, inv); break;
      default: pass_inplace<T, 8, (KE + 7) / 8>(buf, tw, A.n, p, LS, nlines, inv); break;
    }
    p *= R;
  }
}

// Same contract as fft_axis_kernel, one LDS buffer, tile of up to
// kIpThreads*KE elements; radices limited to 2, 3, 4, 5 and 8.
template <typename T, int KIND, int KE>
__global__ __launch_bounds__(kIpThreads) void fft_axis_inplace_kernel(AxisArgs<T> A) {
  using C = cplx_t<T>;
  constexpr int NB = load_batch<T>();
  extern __shared__ unsigned char smem_raw[];
  const int n = A.n;
  const int L = A.lines;
  const int LS = n | 1;
  C *buf = reinterpret_cast<C *>(smem_raw);
  C *tw = buf + L * LS;
  for (int t = threadIdx.x; t < n; t += kIpThreads) tw[t] = A.tw[t];

  const long tile = blockIdx.x;
  long base, i0 = 0, line0 = 0;
  int nl;
  if (A.xaxis) {
    line0 = tile * L;
    base = line0 * n;
    const long rem = A.total_lines - line0;
    nl = rem < L ? (int)rem : L;
  } else {
    const long o = tile / A.tiles_per_outer;
    i0 = (tile - o * A.tiles_per_outer) * L;
    base = o * A.outer_stride + i0;
    const long rem = A.inner_count - i0;
    nl = rem < L ? (int)rem : L;
  }
  const int tot = L * n;
  // tile load: NB independent global loads in flight per thread
  if (A.xaxis) {
    const int totv = nl * n;
    for (int e0 = threadIdx.x; e0 < totv; e0 += kIpThreads * NB) {
      T vr[NB], vi[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        if (e < totv) {
          const C t = A.src[base + (long)e];
          vr[u] = t.x;
          vi[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          buf[l * LS + p] = C{vr[u], vi[u]};
        }
      }
    }
  } else {
    // L is a power of two for the strided passes
    const int lsh = __ffs(L) - 1;
    for (int e0 = threadIdx.x; e0 < tot; e0 += kIpThreads * NB) {
      T vr[NB], vi[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        const int p = e >> lsh, l = e & (L - 1);
        if (e < tot && l < nl) {
          const C t = A.src[base + (long)p * A.inner + l];
          vr[u] = t.x;
          vi[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        const int p = e >> lsh, l = e & (L - 1);
        if (e < tot && l < nl) buf[l * LS + p] = C{vr[u], vi[u]};
      }
    }
  }
  __syncthreads();
  const bool inv = (KIND == kPassBwd || KIND == kPassBwdCombine);
  run_inplace<T, KE>(buf, tw, A, LS, nl, inv);
  if (KIND == kPassKin || KIND == kPassPrec) {
    const int lsh = __ffs(L) - 1;
    for (int e = threadIdx.x; e < nl * n; e += kIpThreads) {
      const int l = e % nl, p = e / nl;
      const long inner = i0 + l;
      T k2;
      if (A.ksqr) {
        k2 = A.ksqr[(long)p * A.inner + inner];
      } else {
        const long jy = inner / A.nx, jx = inner - jy * A.nx;
        k2 = A.kx[jx] + A.ky[jy] + A.kz[p];
        if (k2 > A.kcap) k2 = A.kcap;
      }
      if (KIND == kPassPrec) k2 = A.pa / ((T)1 + k2 * A.pb);
      C v = buf[l * LS + p];
      v.x *= k2;
      v.y *= k2;
      buf[l * LS + p] = v;
    }
    (void)lsh;
    __syncthreads();
    run_inplace<T, KE>(buf, tw, A, LS, nl, true);
  }
  if (KIND == kPassBwdCombine) {
    const int totv = nl * n;
    for (int e0 = threadIdx.x; e0 < totv; e0 += kIpThreads * NB) {
      T pr[NB], pim[NB];
      T vv[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          const long row = (line0 + l) % A.vloc_rows;
          vv[u] = A.vloc[row * n + p];
          const C t = A.psi_in[base + (long)e];
          pr[u] = t.x;
          pim[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int e = e0 + u * kIpThreads;
        if (e < totv) {
          const int l = e / n, p = e - l * n;
          const C t = buf[l * LS + p];
          C o;
          o.x = A.a1 * (t.x + vv[u] * pr[u]) + A.a0 * pr[u];
          o.y = A.a1 * (t.y + vv[u] * pim[u]) + A.a0 * pim[u];
          A.dst[base + (long)e] = o;
        }
      }
    }
  } else if (A.xaxis) {
    const int totv = nl * n;
    for (int e = threadIdx.x; e < totv; e += kIpThreads) {
      const int l = e / n, p = e - l * n;
      A.dst[base + (long)e] = buf[l * LS + p];
    }
  } else {
    const int lsh = __ffs(L) - 1;
    for (int e = threadIdx.x; e < tot; e += kIpThreads) {
      const int p = e >> lsh, l = e & (L - 1);
      if (l < nl) A.dst[base + (long)p * A.inner + l] = buf[l * LS + p];
    }
  }
}

static bool inplace_supported(const AxisPlan &pl) {
  for (int i = 0; i < pl.nf; ++i) {
    const int r = pl.f[i];
    if (!(r == 2 || r == 3 || r == 4 || r == 5 || r == 8)) return false;
  }
  return true;
}

static size_t axis_lds_bytes(int n, int lines, size_t elem) {
  const int LS = n | 1;
  return (size_t)(2 * lines * LS + n) * elem;
}

// Tile shape of the in-place kernel: the largest tile within 64 KiB of LDS
// and kIpThreads*KE elements; y/z tiles take a power-of-two number of
// consecutive inner indices (512-byte rows when the length allows).
template <typename T>
static int fft_pass_inplace(fbse_ctx *c, int axis, int kind, AxisArgs<T> A, long nbatch, int lines_req) {
  constexpr int KE = sizeof(T) == 4 ? 16 : 8;
  const int n = A.n, LS = n | 1;
  const long cap_lds = (65536 / (long)sizeof(cplx_t<T>) - n) / LS;
  const long cap_e = (long)kIpThreads * KE / n;
  long L = std::min(cap_lds, cap_e);
  if (L < 1) return set_err(c, FBSE_ERR_UNSUPPORTED, "axis length %d too long for one tile", n);
  if (lines_req > 0) L = std::min(L, (long)lines_req);
  long ntiles;
  if (axis == 0) {
    L = std::min(L, 64L);
    A.xaxis = 1;
    A.inner = 1;
    A.total_lines = nbatch * c->ny * c->nz;
    if (L > A.total_lines) L = A.total_lines;
    ntiles = (A.total_lines + L - 1) / L;
  } else {
    const long lmax = sizeof(T) == 4 ? 64 : 32;
    L = std::min(L, lmax);
    long pw = 1;
    while (pw * 2 <= L) pw *= 2;
    L = pw;
    A.xaxis = 0;
    A.inner = axis == 1 ? c->nx : c->nx * c->ny;
    A.inner_count = A.inner;
    A.outer_stride = A.inner * n;
    A.tiles_per_outer = (A.inner_count + L - 1) / L;
    ntiles = (axis == 1 ? nbatch * c->nz : nbatch) * A.tiles_per_outer;
  }
  A.lines = (int)L;
  const size_t lds = (size_t)(L * LS + n) * sizeof(cplx_t<T>);
  hipEvent_t ev = nullptr;
  const int pk = (kind == kPassBwdCombine) ? 1 : 0;
  prof_begin(c, pk, &ev);
  switch (kind) {
    case kPassFwd:
      hipLaunchKernelGGL((fft_axis_inplace_kernel<T, kPassFwd, KE>), dim3(ntiles), dim3(kIpThreads), lds, c->stream, A);
      break;
    case kPassBwd:
      hipLaunchKernelGGL((fft_axis_inplace_kernel<T, kPassBwd, KE>), dim3(ntiles), dim3(kIpThreads), lds, c->stream, A);
      break;
    case kPassKin:
      hipLaunchKernelGGL((fft_axis_inplace_kernel<T, kPassKin, KE>), dim3(ntiles), dim3(kIpThreads), lds, c->stream, A);
      break;
    case kPassPrec:
      hipLaunchKernelGGL((fft_axis_inplace_kernel<T, kPassPrec, KE>), dim3(ntiles), dim3(kIpThreads), lds, c->stream,
                         A);
      break;
    default:
      hipLaunchKernelGGL((fft_axis_inplace_kernel<T, kPassBwdCombine, KE>), dim3(ntiles), dim3(kIpThreads), lds,
                         c->stream, A);
      break;
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, pk, ev);
  return FBSE_OK;
}

template <typename T>
static int fft_pass_variant(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
                            const cplx_t<T> *psi_in, T a1, T a0, int variant, int lines_req) {
  const AxisPlan &pl = c->plan[axis];
  AxisArgs<T> A{};
  A.src = src;
  A.dst = dst;
  if constexpr (sizeof(T) == 8) {
    A.tw = c->d_tw[axis];
    A.kx = c->d_kx;
    A.ky = c->d_ky;
    A.kz = c->d_kz;
    A.ksqr = c->ksqr_sep ? nullptr : c->d_ksqr;
    A.vloc = c->d_vloc;
  } else {
    A.tw = c->d_tw_f[axis];
    A.kx = c->d_kx_f;
    A.ky = c->d_ky_f;
    A.kz = c->d_kz_f;
    A.ksqr = c->ksqr_sep ? nullptr : c->d_ksqr_f;
    A.vloc = c->d_vloc_f;
  }
  A.kcap = (T)c->kcap;
  A.nx = (int)c->nx;
  A.n = pl.n;
  A.nf = pl.nf;
  for (int i = 0; i < pl.nf; ++i) A.f[i] = pl.f[i];
  // lines per tile: the requested value, clamped so the tile never exceeds the data
  int lines = lines_req > 0 ? lines_req : pl.lines;
  A.psi_in = psi_in;
  A.a1 = a1;
  A.a0 = a0;
  A.pa = a1;  // the preconditioner pass takes (pa, pb) in the (a1, a0) slots
  A.pb = a0;
  A.vloc_rows = c->nspin * c->nz * c->ny;
  if (variant == 1) return fft_pass_inplace<T>(c, axis, kind, A, nbatch, lines_req);
  long ntiles;
  if (axis == 0) {
    A.xaxis = 1;
    A.inner = 1;
    A.total_lines = nbatch * c->ny * c->nz;
    if (lines > A.total_lines) lines = (int)A.total_lines;
    ntiles = (A.total_lines + lines - 1) / lines;
  } else {
    A.xaxis = 0;
    A.inner = axis == 1 ? c->nx : c->nx * c->ny;
    A.inner_count = A.inner;
    A.outer_stride = A.inner * pl.n;
    if (lines > A.inner_count) lines = (int)A.inner_count;
    A.tiles_per_outer = (A.inner_count + lines - 1) / lines;
    const long nouter = axis == 1 ? nbatch * c->nz : nbatch;
    ntiles = nouter * A.tiles_per_outer;
  }
  // keep the tile within 64 KiB of LDS
  while (lines > 1 && axis_lds_bytes(pl.n, lines, sizeof(cplx_t<T>)) > 65536) --lines;
  if (axis == 0) {
    ntiles = (A.total_lines + lines - 1) / lines;
  } else {
    A.tiles_per_outer = (A.inner_count + lines - 1) / lines;
    ntiles = (axis == 1 ? nbatch * c->nz : nbatch) * A.tiles_per_outer;
  }
  A.lines = lines;
  const size_t lds = axis_lds_bytes(pl.n, lines, sizeof(cplx_t<T>));
  if (lds > 65536) return set_err(c, FBSE_ERR_UNSUPPORTED, "axis length %d needs %zu bytes of LDS", pl.n, lds);
  hipEvent_t ev = nullptr;
  const int pk = (kind == kPassBwdCombine) ? 1 : 0;
  prof_begin(c, pk, &ev);
  switch (kind) {
    case kPassFwd:
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassFwd>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassBwd:
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassBwd>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassKin:
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassKin>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassPrec:
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassPrec>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    default:
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassBwdCombine>), dim3(ntiles), dim3(kFftThreads), lds, c->stream,
                         A);
      break;
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, pk, ev);
  return FBSE_OK;
}


static int pass_class(int kind) { return kind == kPassKin || kind == kPassPrec ? 1 : (kind == kPassBwdCombine ? 2 : 0); }

// Autotuner: the first call of each (precision, axis, pass class) times every
// candidate kernel/tile on the real batch, writing into a scratch buffer so
// the caller's data is untouched, and keeps the fastest.
template <typename T>
static int tune_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, long nbatch, const cplx_t<T> *psi_in,
                     T a1, T a0) {
  const int pi = sizeof(T) == 8 ? 0 : 1, cls = pass_class(kind);
  const int bk_ = nbatch * c->ngrid >= (32L << 20) ? 1 : 0;
  const AxisPlan &pl = c->plan[axis];
  std::vector<std::pair<int, int>> cand;
  const int base_lines[] = {4, 8, 16, 32, 64};
  for (int L : base_lines) cand.push_back({0, L});
  if (inplace_supported(pl))
    for (int L : base_lines) cand.push_back({1, L});
  const size_t bytes = (size_t)nbatch * c->ngrid * sizeof(cplx_t<T>);
  cplx_t<T> *scratch = nullptr;
  if (hipMalloc(&scratch, bytes) != hipSuccess) {
    hipGetLastError();
    c->tune_kernel[pi][bk_][axis][cls] = 0;
    c->tune_lines[pi][bk_][axis][cls] = pl.lines;
    return FBSE_OK;  // no room to tune: keep the default ping-pong tile
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int saved_profile = c->profile;
  c->profile = 0;
  double best = 1e30;
  int bk = 0, bl = pl.lines;
  for (auto &cd : cand) {
    double tmin = 1e30;
    bool ok = true;
    for (int rep = 0; rep < 3 && ok; ++rep) {
      hipEventRecord(e0, c->stream);
      int rc = fft_pass_variant<T>(c, axis, kind, src, scratch, nbatch, psi_in, a1, a0, cd.first, cd.second);
      hipEventRecord(e1, c->stream);
      if (rc) {
        ok = false;
        c->err.clear();
        break;
      }
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      tmin = std::min(tmin, (double)ms);
    }
    if (ok && tmin < best) {
      best = tmin;
      bk = cd.first;
      bl = cd.second;
    }
  }
  c->profile = saved_profile;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamSynchronize(c->stream);
  hipFree(scratch);
  c->tune_kernel[pi][bk_][axis][cls] = bk;
  c->tune_lines[pi][bk_][axis][cls] = bl;
  c->tune_ms[pi][bk_][axis][cls] = best;
  return FBSE_OK;
}

template <typename T>
int fft_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
             const cplx_t<T> *psi_in, T a1, T a0) {
  const int pi = sizeof(T) == 8 ? 0 : 1, cls = pass_class(kind);
  const int bk = nbatch * c->ngrid >= (32L << 20) ? 1 : 0;
  if (c->force_variant >= 0) {
    const int v = (c->force_variant == 1 && inplace_supported(c->plan[axis])) ? 1 : 0;
    return fft_pass_variant<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0, v, c->force_lines);
  }
  if (c->tune_lines[pi][bk][axis][cls] == 0) {
    if (c->autotune && nbatch * c->ngrid >= (1L << 20)) {
      int rc = tune_pass<T>(c, axis, kind, src, nbatch, psi_in, a1, a0);
      if (rc) return rc;
    } else {
      // untuned default: the in-place kernel when its radices allow
      return fft_pass_variant<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0,
                                 inplace_supported(c->plan[axis]) ? 1 : 0, 0);
    }
  }
  return fft_pass_variant<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0, c->tune_kernel[pi][bk][axis][cls],
                             c->tune_lines[pi][bk][axis][cls]);
}

template <typename T>
int apply_h_cplx(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1, T a0) {
  const long nb = nstates * c->nspin;
  int rc;
  if ((rc = fft_pass<T>(c, 0, kPassFwd, psi, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 1, kPassFwd, work, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 2, kPassKin, work, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 1, kPassBwd, work, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 0, kPassBwdCombine, work, out, nb, psi, a1, a0))) return rc;
  if (c->has_proj) {
    if ((rc = proj_apply<T>(c, psi, out, nstates, (double)a1))) return rc;
  }
  return FBSE_OK;
}

template <typename T>
int precondition_cplx(fbse_ctx *c, const cplx_t<T> *r, cplx_t<T> *out, cplx_t<T> *work, long nstates, double e_p) {
  const long nb = nstates * c->nspin;
  const double pa = 1.0 / (double)c->ngrid, pb = (double)c->ngrid / e_p;
  int rc;
  if ((rc = fft_pass<T>(c, 0, kPassFwd, r, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 1, kPassFwd, work, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = fft_pass<T>(c, 2, kPassPrec, work, work, nb, nullptr, (T)pa, (T)pb))) return rc;
  if ((rc = fft_pass<T>(c, 1, kPassBwd, work, work, nb, nullptr, 0, 0))) return rc;
  return fft_pass<T>(c, 0, kPassBwd, work, out, nb, nullptr, 0, 0);
}
template int precondition_cplx<double>(fbse_ctx *, const double2 *, double2 *, double2 *, long, double);

template int fft_pass<double>(fbse_ctx *, int, int, const double2 *, double2 *, long, const double2 *, double,
                              double);
template int fft_pass<float>(fbse_ctx *, int, int, const float2 *, float2 *, long, const float2 *, float, float);
template int apply_h_cplx<double>(fbse_ctx *, const double2 *, double2 *, double2 *, long, double, double);
template int apply_h_cplx<float>(fbse_ctx *, const float2 *, float2 *, float2 *, long, float, float);

}  // namespace fbse
