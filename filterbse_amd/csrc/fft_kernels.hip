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

template <typename T, int KIND>
__global__ __launch_bounds__(kFftThreads) void fft_axis_kernel(AxisArgs<T> A) {
  using C = cplx_t<T>;
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
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int l = e / n, p = e - l * n;
      if (l < nl) buf0[l * LS + p] = A.src[base + (long)e];
    }
  } else {
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int p = e / L, l = e - p * L;
      if (l < nl) buf0[l * LS + p] = A.src[base + (long)p * A.inner + l];
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
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int l = e / n, p = e - l * n;
      if (l >= nl) continue;
      const long g = base + (long)e;
      const long row = (line0 + l) % A.vloc_rows;
      const T v = A.vloc[row * n + p];
      const C pin = A.psi_in[g];
      C t = res[l * LS + p];
      C o;
      o.x = A.a1 * (t.x + v * pin.x) + A.a0 * pin.x;
      o.y = A.a1 * (t.y + v * pin.y) + A.a0 * pin.y;
      A.dst[g] = o;
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

static size_t axis_lds_bytes(int n, int lines, size_t elem) {
  const int LS = n | 1;
  return (size_t)(2 * lines * LS + n) * elem;
}

template <typename T>
int fft_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
             const cplx_t<T> *psi_in, T a1, T a0) {
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
  // lines per tile: the tuned value, clamped so the tile never exceeds the data
  int lines = pl.lines;
  A.psi_in = psi_in;
  A.a1 = a1;
  A.a0 = a0;
  A.pa = a1;  // the preconditioner pass takes (pa, pb) in the (a1, a0) slots
  A.pb = a0;
  A.vloc_rows = c->nspin * c->nz * c->ny;
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
