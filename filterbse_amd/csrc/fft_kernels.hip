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
#include "fft_common.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>

namespace fbse {

// q = i / d, r = i % d for 0 <= i < 2^22, 1 <= d: float reciprocal, one correction
__device__ __forceinline__ void divmod_small(int i, int d, float inv_d, int &q, int &r) {
  q = (int)((float)i * inv_d);
  r = i - q * d;
  if (r >= d) {
    r -= d;
    ++q;
  }
  if (r < 0) {
    r += d;
    --q;
  }
}

// One Stockham pass of compile-time radix R over all lines of the tile (one
// butterfly per loop iteration: nothing is held across a barrier).
template <typename T, int R>
__device__ __forceinline__ void pass_fixed(const cplx_t<T> *in, cplx_t<T> *out, const cplx_t<T> *tw,
                                                     int n, int p, int LS, int nlines, bool inv) {
  using C = cplx_t<T>;
  const int m = n / R;
  const int tstep = n / (p * R);
  const int nb = m * nlines;
  const float inv_m = 1.0f / (float)m, inv_p = 1.0f / (float)p;
  const T sgn = inv ? (T)1 : (T)-1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int l, i, q0, k;
    divmod_small(b, m, inv_m, l, i);
    divmod_small(i, p, inv_p, q0, k);
    const C *src = in + l * LS;
    C *dst = out + l * LS;
    C x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = src[i + r * m];
    if (k > 0) twiddle_powers<T, R>(x, tw, k, tstep, inv);
    Bfly<T, R>::run(x, sgn);
    const int j = (i - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[j + q * p] = x[Bfly<T, R>::perm(q)];
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
      case 6: pass_fixed<T, 6>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 8: pass_fixed<T, 8>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 9: pass_fixed<T, 9>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 10: pass_fixed<T, 10>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 12: pass_fixed<T, 12>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 15: pass_fixed<T, 15>(a, b, tw, A.n, p, LS, nlines, inv); break;
      case 16: pass_fixed<T, 16>(a, b, tw, A.n, p, LS, nlines, inv); break;
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
  T *kzl = reinterpret_cast<T *>(tw + n);  // kinetic pass: k^2 of z, then x+y per line
  T *kxyl = kzl + n;

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
  if ((KIND == kPassKin || KIND == kPassPrec) && !A.ksqr) {
    // separable k^2 into LDS now, so no global load sits between the transforms
    for (int t = threadIdx.x; t < n; t += blockDim.x) kzl[t] = A.kz[t];
    for (int l = threadIdx.x; l < nl; l += blockDim.x) {
      const long inner = i0 + l;
      const long jy = inner / A.nx, jx = inner - jy * A.nx;
      kxyl[l] = A.kx[jx] + A.ky[jy];
    }
  }
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
    const float inv_n = 1.0f / (float)n;
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      int l, p;
      divmod_small(e, n, inv_n, l, p);
      T k2;
      if (A.ksqr) {
        k2 = A.ksqr[(long)p * A.inner + i0 + l];
      } else {
        k2 = kxyl[l] + kzl[p];
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
    const T a0 = A.a0_tab ? (T)A.a0_tab[*A.step_base + A.step_off] : A.a0;
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
          o.x = A.a1 * (t.x + vv[u] * pr[u]) + a0 * pr[u];
          o.y = A.a1 * (t.y + vv[u] * pim[u]) + a0 * pim[u];
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

// LDS of a two-buffer tile: 2 x lines x (n | 1) elements, n twiddles, and the
// kinetic pass's real k^2 tables (n + lines)
static size_t axis_lds_bytes(int n, int lines, size_t elem) {
  const int LS = n | 1;
  return (size_t)(2 * lines * LS + n) * elem + (size_t)(n + lines) * (elem / 2);
}

// LDS a per-axis tile may take (dynamic LDS above 64 KiB is enabled per kernel)
constexpr size_t kAxisLdsCap = 160 * 1024;

template <typename T, int KIND> static void axis_lds_limit(int device) {
  static bool done[64] = {};
  if (device < 0 || device >= 64 || !done[device]) {
    (void)hipFuncSetAttribute((const void *)fft_axis_kernel<T, KIND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kAxisLdsCap);
    if (device >= 0 && device < 64) done[device] = true;
  }
}

template <typename T>
static int fft_pass_lines(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
                          const cplx_t<T> *psi_in, T a1, T a0, int lines_req) {
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
  const AxisPlan &sch = c->aplan[axis];  // the per-axis radix schedule
  A.n = pl.n;
  A.nf = sch.nf;
  for (int i = 0; i < sch.nf; ++i) A.f[i] = sch.f[i];
  // lines per tile: the requested value, clamped so the tile never exceeds the
  // data; 1000 + L asks for the fixed-schedule instance with L lines, which
  // keeps L (a last partial tile is masked) where the axis length has one
  const bool kin = kind == kPassKin || kind == kPassPrec;
  const bool fx = lines_req >= 1000 && axis_fx_length(pl.n) && !(kin && !c->ksqr_sep) &&
                  !(sizeof(T) == 8 && lines_req == 1064);  // 64-line instances are FP32 only
  int lines = lines_req >= 1000 ? lines_req - 1000 : (lines_req > 0 ? lines_req : pl.lines);
  A.psi_in = psi_in;
  A.a1 = a1;
  A.a0 = a0;
  if (kind == kPassBwdCombine && c->dyn_a0_tab) {
    A.a0_tab = c->dyn_a0_tab;
    A.step_base = c->dyn_step_base;
    A.step_off = c->dyn_step_off;
  }
  A.pa = a1;  // the preconditioner pass takes (pa, pb) in the (a1, a0) slots
  A.pb = a0;
  A.vloc_rows = c->nspin * c->nz * c->ny;
  long ntiles;
  if (axis == 0) {
    A.xaxis = 1;
    A.inner = 1;
    A.total_lines = nbatch * c->ny * c->nz;
    if (!fx && lines > A.total_lines) lines = (int)A.total_lines;
    ntiles = (A.total_lines + lines - 1) / lines;
  } else {
    A.xaxis = 0;
    A.inner = axis == 1 ? c->nx : c->nx * c->ny;
    A.inner_count = A.inner;
    A.outer_stride = A.inner * pl.n;
    if (!fx && lines > A.inner_count) lines = (int)A.inner_count;
    A.tiles_per_outer = (A.inner_count + lines - 1) / lines;
    const long nouter = axis == 1 ? nbatch * c->nz : nbatch;
    ntiles = nouter * A.tiles_per_outer;
  }
  // keep the tile within the LDS a work-group may take
  while (!fx && lines > 1 && axis_lds_bytes(pl.n, lines, sizeof(cplx_t<T>)) > kAxisLdsCap) --lines;
  if (axis == 0) {
    ntiles = (A.total_lines + lines - 1) / lines;
  } else {
    A.tiles_per_outer = (A.inner_count + lines - 1) / lines;
    ntiles = (axis == 1 ? nbatch * c->nz : nbatch) * A.tiles_per_outer;
  }
  A.lines = lines;
  A.ntiles = ntiles;
  A.ny = (int)c->ny;
  A.nz = (int)c->nz;
  if (kind == kPassBwdCombine && (c->ax_pjd || c->ax_fwd)) {
    // the fused combine: fixed instances only, tiles within one xy plane
    if (!fx || c->ny % lines != 0)
      return set_err(c, FBSE_ERR_UNSUPPORTED, "fused combine needs a fixed instance with ny %% lines == 0");
    if (c->ax_pjd) {
      A.rec_yoff = c->d_rec_yoff;
      A.rec_pa = c->d_rec_pa;
      if constexpr (sizeof(T) == 8) A.rec_val = c->d_rec_val;
      else A.rec_val = c->d_rec_val_f;
      A.pjd = (const cplx_t<T> *)c->ax_pjd;
      A.pj_natoms = c->pj.n_atoms;
      A.pt_ranks = c->pt_ranks;
    }
    A.fwd_dst = (cplx_t<T> *)c->ax_fwd;
    // spinors in the automatic mode: tiles of 8 rows of both spins of a state
    A.spin_pair = c->axis_fuse == 1 && c->nspin == 2 && lines == 16 && c->ny % 8 == 0 ? 1 : 0;
    if (c->ax_pp) {
      if (!A.spin_pair || !c->ax_pjd)
        return set_err(c, FBSE_ERR_UNSUPPORTED, "projections in the combine need its spin-paired tiles");
      A.pp_lane_off = c->d_pp_lane_off;
      A.pp_lanes = c->d_pp_lanes;
      A.pp_span = c->d_pp_span;
      A.pp_meta = c->d_pp_meta;
      if constexpr (sizeof(T) == 8) A.pp_val = c->d_pp_val;
      else A.pp_val = c->d_pp_val_f;
      A.pp_part = (cplx_t<T> *)c->ax_pp;
      A.pp_nslots = c->pp_nslots;
    }
  }
  // y and z passes of the Newton step in the z-tile layout (AxisArgs::zt_*)
  if (c->ax_zt) {
    if (axis == 0 || !fx || lines != 16 || c->nx % 16 != 0)
      return set_err(c, FBSE_ERR_UNSUPPORTED, "z-tile layout needs the 16-line fixed y / z instances");
    A.zt_src = c->ax_zt & 1;
    A.zt_dst = (c->ax_zt >> 1) & 1;
    A.zt_q = axis == 1 ? c->nz : 1;
    A.zt_inner = axis == 1 ? c->nx * c->nz : 16;
    A.zt_grid = c->ngrid;
    A.zt_nz16 = c->nz * 16;
  }
  // the combine's tiles of one row block for all batch items on one XCD: its
  // V_loc rows (and expansion records) are then fetched into one L2 once
  const int block_rows = A.spin_pair ? lines / 2 : lines;
  if (kind == kPassBwdCombine && fx && c->ny % block_rows == 0)
    A.xcd_order = (c->ny * c->nz / block_rows) % 8 == 0 ? 1 : 0;
  if (fx) {
    hipEvent_t ev = nullptr;
    const int pk = (kind == kPassBwdCombine) ? 1 : 0;
    prof_begin(c, pk, &ev);
    const int rc = axis_fx_pass<T>(c, kind, A, lines);
    if (rc) {
      if (ev) c->ev_pool.push_back(ev);
      return rc;
    }
    prof_end(c, pk, ev);
    return FBSE_OK;
  }
  const size_t lds = axis_lds_bytes(pl.n, lines, sizeof(cplx_t<T>));
  if (lds > kAxisLdsCap) return set_err(c, FBSE_ERR_UNSUPPORTED, "axis length %d needs %zu bytes of LDS", pl.n, lds);
  hipEvent_t ev = nullptr;
  const int pk = (kind == kPassBwdCombine) ? 1 : 0;
  prof_begin(c, pk, &ev);
  switch (kind) {
    case kPassFwd:
      axis_lds_limit<T, kPassFwd>(c->device);
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassFwd>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassBwd:
      axis_lds_limit<T, kPassBwd>(c->device);
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassBwd>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassKin:
      axis_lds_limit<T, kPassKin>(c->device);
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassKin>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    case kPassPrec:
      axis_lds_limit<T, kPassPrec>(c->device);
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassPrec>), dim3(ntiles), dim3(kFftThreads), lds, c->stream, A);
      break;
    default:
      axis_lds_limit<T, kPassBwdCombine>(c->device);
      hipLaunchKernelGGL((fft_axis_kernel<T, kPassBwdCombine>), dim3(ntiles), dim3(kFftThreads), lds, c->stream,
                         A);
      break;
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, pk, ev);
  return FBSE_OK;
}


static int pass_class(int kind) { return kind == kPassKin || kind == kPassPrec ? 1 : (kind == kPassBwdCombine ? 2 : 0); }

// Autotuner: the first call of each (precision, batch class, axis, pass
// class) times every tile height on the real batch, writing into a scratch
// buffer so the caller's data is untouched, and keeps the fastest.
template <typename T>
static int tune_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, long nbatch, const cplx_t<T> *psi_in,
                     T a1, T a0) {
  const int pi = sizeof(T) == 8 ? 0 : 1, cls = pass_class(kind);
  const int bk_ = nbatch * c->ngrid >= (32L << 20) ? 1 : 0;
  const AxisPlan &pl = c->plan[axis];
  // tile heights of the generic kernel, then the fixed-schedule instances
  // (1000 + lines) where the axis length has one
  const int cand[] = {4, 8, 16, 32, 64, 1016, 1032, 1064};
  const size_t bytes = (size_t)nbatch * c->ngrid * sizeof(cplx_t<T>);
  cplx_t<T> *scratch = nullptr;
  if (hipMalloc(&scratch, bytes) != hipSuccess) {
    (void)hipGetLastError();
    c->tune_lines[pi][bk_][axis][cls] = pl.lines;
    return FBSE_OK;  // no room to tune: keep the default tile
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int saved_profile = c->profile;
  c->profile = 0;
  double best = 1e30;
  int bl = pl.lines;
  for (int L : cand) {
    if (L >= 1000 && !axis_fx_length(pl.n)) continue;
    double tmin = 1e30;
    bool ok = true;
    for (int rep = 0; rep < 3 && ok; ++rep) {
      (void)hipEventRecord(e0, c->stream);
      int rc = fft_pass_lines<T>(c, axis, kind, src, scratch, nbatch, psi_in, a1, a0, L);
      (void)hipEventRecord(e1, c->stream);
      if (rc) {
        ok = false;
        c->err.clear();
        break;
      }
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      tmin = std::min(tmin, (double)ms);
    }
    if (ok && tmin < best) {
      best = tmin;
      bl = L;
    }
  }
  c->profile = saved_profile;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(scratch);
  c->tune_lines[pi][bk_][axis][cls] = bl;
  c->tune_ms[pi][bk_][axis][cls] = best;
  return FBSE_OK;
}

template <typename T>
int fft_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
             const cplx_t<T> *psi_in, T a1, T a0) {
  const int pi = sizeof(T) == 8 ? 0 : 1, cls = pass_class(kind);
  const int bk = nbatch * c->ngrid >= (32L << 20) ? 1 : 0;
  if (c->force_lines > 0) return fft_pass_lines<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0, c->force_lines);
  if (c->tune_lines[pi][bk][axis][cls] == 0) {
    if (!(c->autotune && nbatch * c->ngrid >= (1L << 20)))
      return fft_pass_lines<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0, 0);
    int rc = tune_pass<T>(c, axis, kind, src, nbatch, psi_in, a1, a0);
    if (rc) return rc;
  }
  return fft_pass_lines<T>(c, axis, kind, src, dst, nbatch, psi_in, a1, a0, c->tune_lines[pi][bk][axis][cls]);
}

// Plane/slab passes (3 sweeps) when the grid's planes fit a work-group's LDS,
// per-axis passes (5 sweeps) otherwise or when forced.
template <typename T> static bool use_tiles(const fbse_ctx *c) {
  if (c->force_variant == 0) return false;
  return tiles_supported<T>(c);
}

int fft_path(const fbse_ctx *c, int f32, int *instance) {
  const bool tiles = f32 ? use_tiles<float>(c) : use_tiles<double>(c);
  if (instance) {
    if (!tiles)
      *instance = 0;
    else if (c->force_variant == 1 && c->force_lines > 0)
      *instance = c->force_lines;
    else
      *instance = c->tile_variant[f32 ? 1 : 0] + 1;  // 0 while not tuned yet
  }
  return tiles ? 1 : 0;
}

// projection coefficients (column x atom x 6) for nstates states
template <typename T> static int ensure_pjd(fbse_ctx *c, long nstates) {
  const size_t need = (size_t)c->pj.n_atoms * nstates * c->nspin * 6 * sizeof(cplx_t<T>);
  if (need > c->pjd_bytes) {
    if (c->no_alloc)
      return set_err(c, FBSE_ERR_NOMEM, "projection coefficients (%zu bytes) inside the no-allocation entry point",
                     need);
    if (c->d_pjd) (void)hipFree(c->d_pjd);
    c->d_pjd = nullptr;
    c->pjd_bytes = 0;
    FBSE_HIP(c, hipMalloc(&c->d_pjd, need));
    c->pjd_bytes = need;
  }
  return FBSE_OK;
}

// The projection pass (it reads psi only) on the side stream, forked from the
// main stream, so that it can run beside the FFT passes that precede the
// combine; *joined = false: the main stream must wait for ev_join before
// anything reads d.  Runs on the main stream when profiling (every launch
// timed alone) or when the side stream is off.
// Not while the main stream is being captured into a window graph: a fork
// captured there (event record and wait nodes, the projection launched on the
// side stream) kept the memory it touched referenced after the graph was
// destroyed, so the filter's ring was not returned by hipFree (measured: 11 GB
// lost per graph-replayed filter of a 160^3 test grid, 60 GB at 256^3;
// nothing with the projection on the main stream).
template <typename T>
static int project_forked(fbse_ctx *c, const cplx_t<T> *psi, long nstates, double a1, cplx_t<T> *d, bool *joined) {
  *joined = true;
  if (!(c->concurrent_proj && c->side && !c->profile) || c->capturing)
    return proj_project<T>(c, psi, nstates, a1, d);
  hipStream_t main_stream = c->stream;
  FBSE_HIP(c, hipEventRecord(c->ev_fork, main_stream));
  FBSE_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
  c->stream = c->side;
  int rc = proj_project<T>(c, psi, nstates, a1, d);
  c->stream = main_stream;
  hipError_t e = hipEventRecord(c->ev_join, c->side);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(c->side);
    return rc ? rc : set_err(c, FBSE_ERR_HIP, "event record: %s", hipGetErrorString(e));
  }
  if (rc) {
    (void)hipStreamWaitEvent(main_stream, c->ev_join, 0);  // joined again before the error returns
    return rc;
  }
  *joined = false;
  return FBSE_OK;
}

template <typename T>
int apply_h_cplx(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1, T a0) {
  const long nb = nstates * c->nspin;
  int rc;
  if (use_tiles<T>(c)) {
    if (proj_fusable(c)) {
      // projections first (they read psi), the expansion rides on the combine pass
      if ((rc = ensure_pjd<T>(c, nstates))) return rc;
      cplx_t<T> *d = (cplx_t<T> *)c->d_pjd;
      if ((rc = proj_project<T>(c, psi, nstates, (double)a1, d))) return rc;
      return kinetic_local_tiles<T>(c, psi, out, work, nstates, a1, a0, d);
    }
    if ((rc = kinetic_local_tiles<T>(c, psi, out, work, nstates, a1, a0))) return rc;
  } else {
    // with one radial projector: projections before the passes (they read
    // psi), the expansion from the rank records after the combine
    cplx_t<T> *d = nullptr;
    bool joined = true;
    if (proj_fusable(c)) {
      if ((rc = ensure_pjd<T>(c, nstates))) return rc;
      d = (cplx_t<T> *)c->d_pjd;
      if ((rc = project_forked<T>(c, psi, nstates, (double)a1, d, &joined))) return rc;
    }
    rc = fft_pass<T>(c, 0, kPassFwd, psi, work, nb, nullptr, 0, 0);
    if (!rc) rc = fft_pass<T>(c, 1, kPassFwd, work, work, nb, nullptr, 0, 0);
    if (!rc) rc = fft_pass<T>(c, 2, kPassKin, work, work, nb, nullptr, 0, 0);
    if (!rc) rc = fft_pass<T>(c, 1, kPassBwd, work, work, nb, nullptr, 0, 0);
    if (!joined) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);  // the projections are done
    if (rc) return rc;
    if ((rc = fft_pass<T>(c, 0, kPassBwdCombine, work, out, nb, psi, a1, a0))) return rc;
    if (d) return proj_expand<T>(c, d, out, nstates);
  }
  if (c->has_proj) {
    if ((rc = proj_apply<T>(c, psi, out, nstates, (double)a1))) return rc;
  }
  return FBSE_OK;
}

// Per-axis passes with the fused combine: the x combine pass (fixed 16-line
// instance) adds the projector expansion and writes the forward x transform
// of its output into the work buffer, where the next Newton step's y pass
// starts (first = true: this step's forward x pass runs first).  For axis
// lengths with a fixed instance (x) and ny a multiple of 16, with no
// projectors or projectors the record layout covers, when the per-axis tile
// heights are automatic or forced to the fixed 16-line instance.
// lines per tile of the fused combine: fbse_set_axis_fusion's 16, 32 or 64
// (64: FP32 only), else 16
template <typename T> static int fused_lines(const fbse_ctx *c) {
  const int f = c->axis_fuse;
  return (f == 32 || (f == 64 && sizeof(T) == 4)) ? f : 16;
}
template <typename T> static bool axis_fused(const fbse_ctx *c) {
  // (the combine keeps 2 * pt_ranks rank-segment bounds in its 256 + 16
  // element k^2 tables, unused by the combine)
  return c->axis_fuse && axis_fx_length((int)c->nx) && c->ny % fused_lines<T>(c) == 0 &&
         (!c->has_proj || (proj_fusable(c) && c->pt_ranks <= 128)) &&
         (c->force_lines == 0 || c->force_lines == 1000 + fused_lines<T>(c));
}

// The Newton step's y and z passes in the z-tile layout (automatic mode of
// fbse_set_axis_fusion): 16-line fixed instances on y and z (lengths 160 or
// 256, separable k^2), nx a multiple of 16, tile heights not forced otherwise.
template <typename T> static bool axis_zt(const fbse_ctx *c) {
  return c->axis_fuse == 1 && c->nx % 16 == 0 && axis_fx_length((int)c->ny) && axis_fx_length((int)c->nz) &&
         c->ksqr_sep && (c->force_lines == 0 || c->force_lines == 1016);
}

// Projections inside the spin-paired combine (fbse_set_combine_projection):
// within a filter run's chain of Newton steps, with one radial projector of
// each kind and the combine in 8-row tiles of both spins.
template <typename T> static bool pp_usable(const fbse_ctx *c) {
  return c->pp_on && c->pp_chain && !c->no_alloc && c->d_pp_lanes && c->has_proj && proj_fusable(c) &&
         c->axis_fuse == 1 && c->nspin == 2 && c->ny % 8 == 0 && fused_lines<T>(c) == 16;
}
template <typename T> static int ensure_pp_part(fbse_ctx *c, long nstates) {
  const size_t need = (size_t)nstates * c->pp_nslots * 12 * sizeof(cplx_t<T>);
  if (need > c->pp_part_bytes) {
    if (c->d_pp_part) (void)hipFree(c->d_pp_part);
    c->d_pp_part = nullptr;
    c->pp_part_bytes = 0;
    c->pp_ready = 0;
    FBSE_HIP(c, hipMalloc(&c->d_pp_part, need));
    c->pp_part_bytes = need;
  }
  return FBSE_OK;
}

template <typename T>
static int apply_h_axes_chain(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates,
                              T a1, T a0, bool first) {
  const long nb = nstates * c->nspin;
  cplx_t<T> *d = nullptr;
  bool joined = true;
  int rc;
  const bool pp = pp_usable<T>(c);
  if (c->has_proj) {
    if ((rc = ensure_pjd<T>(c, nstates))) return rc;
    d = (cplx_t<T> *)c->d_pjd;
    // the previous step's combine left psi's per-tile projection sums
    const bool sums = pp && !first && c->pp_ready && c->pp_for == (const void *)psi && c->pp_for_n == nstates;
    c->pp_ready = 0;
    if (sums) rc = proj_reduce<T>(c, nstates, (double)a1, d);
    else rc = project_forked<T>(c, psi, nstates, (double)a1, d, &joined);
    if (rc) return rc;
  }
  if (pp && (rc = ensure_pp_part<T>(c, nstates))) {
    if (!joined) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    return rc;
  }
  rc = first ? fft_pass<T>(c, 0, kPassFwd, psi, work, nb, nullptr, 0, 0) : FBSE_OK;
  if (axis_zt<T>(c)) {
    // y forward from work into out in the z-tile layout, z in place there
    // (contiguous tiles), y backward back into work: out (the ring slot of
    // this step's result) is free until the combine writes it
    c->ax_zt = 2;
    if (!rc) rc = fft_pass_lines<T>(c, 1, kPassFwd, work, out, nb, nullptr, 0, 0, 1016);
    c->ax_zt = 3;
    if (!rc) rc = fft_pass_lines<T>(c, 2, kPassKin, out, out, nb, nullptr, 0, 0, 1016);
    c->ax_zt = 1;
    if (!rc) rc = fft_pass_lines<T>(c, 1, kPassBwd, out, work, nb, nullptr, 0, 0, 1016);
    c->ax_zt = 0;
  } else {
    if (!rc) rc = fft_pass<T>(c, 1, kPassFwd, work, work, nb, nullptr, 0, 0);
    if (!rc) rc = fft_pass<T>(c, 2, kPassKin, work, work, nb, nullptr, 0, 0);
    if (!rc) rc = fft_pass<T>(c, 1, kPassBwd, work, work, nb, nullptr, 0, 0);
  }
  if (!joined) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);  // the projections are done
  if (rc) return rc;
  c->ax_pjd = d;
  c->ax_fwd = work;
  c->ax_pp = pp ? c->d_pp_part : nullptr;
  rc = fft_pass_lines<T>(c, 0, kPassBwdCombine, work, out, nb, psi, a1, a0, 1000 + fused_lines<T>(c));
  c->ax_pjd = nullptr;
  c->ax_fwd = nullptr;
  c->ax_pp = nullptr;
  if (pp && !rc) {
    c->pp_ready = 1;
    c->pp_for = out;
    c->pp_for_n = nstates;
  }
  return rc;
}

template <typename T>
int apply_h_chain(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1, T a0,
                  bool first) {
  if (!use_tiles<T>(c)) {
    if (axis_fused<T>(c)) return apply_h_axes_chain<T>(c, psi, out, work, nstates, a1, a0, first);
    return apply_h_cplx<T>(c, psi, out, work, nstates, a1, a0);
  }
  const cplx_t<T> *d = nullptr;
  hipEvent_t join = nullptr;
  int rc;
  if (proj_fusable(c)) {
    if ((rc = ensure_pjd<T>(c, nstates))) return rc;
    d = (cplx_t<T> *)c->d_pjd;
    bool joined = true;
    // beside the fixed-schedule passes the side stream loses: measured on
    // 128 grids of 96^3, the z-slab pass took 803 us next to the projection
    // pass instead of 368 + 306 us one after the other
    if (tiles_fixed<T>(c)) {
      if ((rc = proj_project<T>(c, psi, nstates, (double)a1, (cplx_t<T> *)d))) return rc;
    } else if ((rc = project_forked<T>(c, psi, nstates, (double)a1, (cplx_t<T> *)d, &joined))) {
      return rc;
    }
    if (!joined) join = c->ev_join;
  }
  // the expansion of non-fusable projectors is added to out after the
  // combine pass, so then out's forward transform cannot ride on it
  const bool chain = !c->has_proj || d;
  rc = kinetic_local_tiles<T>(c, psi, out, work, nstates, a1, a0, d, chain && !first, chain, join);
  if (rc) {
    if (join) (void)hipStreamWaitEvent(c->stream, join, 0);  // never leave the side stream unjoined
    return rc;
  }
  if (!chain) return proj_apply<T>(c, psi, out, nstates, (double)a1);
  return FBSE_OK;
}
// One Newton step's H application for a group of contexts of one grid size
// that all run the FP32 fixed plane/slab instance (tile_groupable) with
// projectors the plane passes expand (or none): the projections per context,
// then the three passes as grouped launches on cs[0]'s stream (all contexts'
// streams must be that stream).  Same chain as apply_h_chain: the combine
// writes the forward xy transform of out into work for the next step.
int apply_h_group(int n, fbse_ctx *const *cs, const float2 *const *psi, float2 *const *out, float2 *const *work,
                  const long *nstates, const float *a1, const float *a0, bool first) {
  const float2 *pjd[kTileGroupMax] = {};
  long nb[kTileGroupMax];
  // contexts with projectors: one grouped projection launch
  fbse_ctx *pc[kTileGroupMax];
  const float2 *pp[kTileGroupMax];
  float2 *pd[kTileGroupMax];
  long pn[kTileGroupMax];
  float pa1[kTileGroupMax];
  int np = 0, rc;
  for (int i = 0; i < n; ++i) {
    fbse_ctx *c = cs[i];
    nb[i] = nstates[i] * c->nspin;
    if (c->has_proj) {
      if ((rc = ensure_pjd<float>(c, nstates[i]))) return rc;
      pjd[i] = (const float2 *)c->d_pjd;
      pc[np] = c;
      pp[np] = psi[i];
      pd[np] = (float2 *)c->d_pjd;
      pn[np] = nstates[i];
      pa1[np] = a1[i];
      ++np;
    }
  }
  if (np && (rc = proj_project_group(np, pc, pp, pn, pa1, pd))) {
    if (pc[0] != cs[0]) set_err(cs[0], rc, "%s", pc[0]->err.c_str());
    return rc;
  }
  if (first && (rc = tile_group_pass(n, cs, kTileFwdXY, psi, work, nb, nullptr, nullptr, nullptr, nullptr, nullptr)))
    return rc;
  const float2 *wsrc[kTileGroupMax];
  for (int i = 0; i < n; ++i) wsrc[i] = work[i];
  if ((rc = tile_group_pass(n, cs, kTileKinZ, wsrc, work, nb, nullptr, nullptr, nullptr, nullptr, nullptr))) return rc;
  return tile_group_pass(n, cs, kTileBwdXYCombine, wsrc, out, nb, psi, a1, a0, pjd, work);
}

template int apply_h_chain<double>(fbse_ctx *, const double2 *, double2 *, double2 *, long, double, double, bool);
template int apply_h_chain<float>(fbse_ctx *, const float2 *, float2 *, float2 *, long, float, float, bool);

template <typename T>
int precondition_cplx(fbse_ctx *c, const cplx_t<T> *r, cplx_t<T> *out, cplx_t<T> *work, long nstates, double e_p) {
  const long nb = nstates * c->nspin;
  const double pa = 1.0 / (double)c->ngrid, pb = (double)c->ngrid / e_p;
  int rc;
  if (use_tiles<T>(c)) return precondition_tiles<T>(c, r, out, work, nstates, (T)pa, (T)pb);
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
