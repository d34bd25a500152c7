// fft_tile_kernels.hip -- the kinetic part of H in three HBM passes.
//
// A gfx950 work-group can hold 160 KiB of LDS, enough for a whole nx x ny
// plane of a 96^3 grid in FP32 or FP64.  The 3D FFT pair of the kinetic
// operator (hamiltonian.c:62-99) therefore runs as
//   pass 1  xy planes : forward x FFT, forward y FFT          (plane in, plane out)
//   pass 2  xz slabs  : forward z FFT, x k^2 (or the preconditioner), backward z FFT
//   pass 3  xy planes : backward y FFT, backward x FFT, then
//                       out = a1*(T psi + V psi_in) + a0*psi_in
// instead of the five single-axis passes of fft_kernels.hip.  Each work-group
// loads one tile into LDS with rows of nx contiguous elements, transforms it
// in place with register-staged Stockham passes (every thread reads its
// butterflies' inputs into registers, the group syncs, outputs are written to
// their self-sorted positions), and writes it back.  Radices 2, 3, 4, 5, 8 and
// the composite register butterflies 6, 9, 10, 12, 15, 16.
//
// Two kinds of instance: generic ones (any line lengths and radix schedule,
// chosen at run time, one function call per radix pass) and fixed-schedule
// persistent ones for given sizes (fx_body: the schedule and strides are
// compile-time constants, the passes are inlined, work-groups loop over
// tiles).  The fixed-schedule per-axis instances (axis_fx_body) reuse the same
// inlined passes on the per-axis tiles of fft_kernels.hip.
//
// Tile geometry (element (r, c) of tile t of batch item b):
//   xy plane : rows r = y (ny), cols c = x (nx), tile t = z,
//              global index b*ngrid + z*nx*ny + r*nx + c
//   xz slab  : rows r = z (nz), cols c = x (nx), tile t = y,
//              global index b*ngrid + y*nx + r*nx*ny + c
#include "fbse_internal.h"
#include "fft_common.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace fbse {


// A thread stages ceil(E / R) butterflies of radix R per pass: a work-group
// of NT threads covers tiles of up to NT * E elements.  E is a compile-time
// kernel parameter; the autotuner picks among the instantiated values.  An
// instance runs radix R only when its staged values stay within E + 4, which
// bounds the registers a pass holds across its barrier.
constexpr bool tile_radix_fits(int e, int r) { return (e + r - 1) / r * r <= e + 4; }

template <typename T> struct TileArgs {
  const cplx_t<T> *src;
  cplx_t<T> *dst;
  const cplx_t<T> *twc, *twr;  // twiddles of the row (length nc) and column (length nr) transforms
  int nc, nr, ls;              // tile columns, rows, LDS row stride (elements)
  int nfc, nfr;
  unsigned long long fc, fr;   // radix schedules, 4 bits (radix - 1) per factor, first factor lowest
  long row_stride;             // global elements between tile rows
  long tile_stride;            // global elements between tiles of one batch item
  long item_stride;            // global elements between batch items (ngrid)
  int tiles_per_item;
  int col_split;               // tiles per slab along x (z passes; 1 elsewhere)
  long nbatch;                 // batch items (state x spin grids)
  // z passes: k2 = min(kx[x] + ky[y] + kz[z], kcap) when separable, else ksqr[x + nx*y + nx*ny*z]
  const T *kx, *ky, *kz, *ksqr;
  T kcap;
  long nxny;
  T pa, pb;                    // preconditioner g = pa / (1 + k2 * pb)
  // combine: V of spin s starts at vloc + s*ngrid (items run spin-fastest)
  const cplx_t<T> *psi_in;
  const T *vloc;
  int nspin;
  T a1, a0;
  // when a0_tab is set, a0 = a0_tab[*step_base + step_off] (graph replays)
  const double *a0_tab;
  const long *step_base;
  int step_off;
  // projector expansion in the combine pass (rec_off != nullptr): in rounds
  // k < pt_ranks, records rec_off[z*pt_ranks + k] .. rec_off[z*pt_ranks + k + 1]-1
  // of plane z, record e at point (y, x) = (rec_pa[2e] >> 16, rec_pa[2e] & 0xffff)
  // of atom rec_pa[2e+1] with values rec_val[8e ..] = proj, nl_proj, Re/Im Y_1m;
  // coefficients of this work-group's column pjd[(item*pj_natoms + atom)*6 + r]
  const int *rec_off, *rec_pa;
  const T *rec_val;
  const cplx_t<T> *pjd;
  int pj_natoms, pt_ranks;
  // combine: when set, the forward x and y transforms of the output plane are
  // also written here (same layout as dst)
  cplx_t<T> *fwd_dst;
};

// i mod p for 0 <= i < 2^22, p >= 1: float quotient, then one correction.
__device__ __forceinline__ int mod_small(int i, int p, float inv_p) {
  int q = (int)((float)i * inv_p);
  int r = i - q * p;
  if (r >= p) r -= p;
  if (r < 0) r += p;
  return r;
}

// The tile, its twiddle tables and the k^2 tables live in the work-group's
// dynamic LDS.  The radix passes below address it through this array with
// element offsets, so every access is an LDS instruction.
extern __shared__ __align__(16) unsigned char tile_lds[];

// One in-place Stockham pass of radix R over `nlines` lines of length n
// (m = n / R butterflies per line, p = product of the radices already done,
// tstep = n / (p R)).  Element j of line l sits at l*lstride + j*estride.  The
// work-group size is a multiple of nlines, so a thread keeps one line
// l0 = tid % nlines for the whole transform and takes butterflies i0, i0 + di,
// ... of it; consecutive threads are on consecutive lines: for column
// transforms (lstride 1) a wavefront touches consecutive LDS words, for row
// transforms (lstride = odd row pitch) words 2*pitch apart, which spread over
// all banks.  Twiddle w^(r kk tstep) is element r*kk*tstep of the length-n
// table at tw_off (r*kk*tstep < n), conjugated for the backward transform.
// Each radix pass is a separate (not inlined) function: one kernel holding
// every radix's inlined pass spills registers, while separate functions keep
// their own allocation at the cost of a call per pass.
template <typename T, int R, int KB, bool INV>
__device__ __attribute__((noinline)) void tile_radix_pass(int tw_off, int m, int p, int tstep, int estride,
                                                          int lstride, int l0, int i0, int di) {
  using C = cplx_t<T>;
  C *buf = reinterpret_cast<C *>(tile_lds);
  const C *tw = buf + tw_off;
  const float inv_p = p > 1 ? __builtin_amdgcn_rcpf((float)p) : 1.0f;  // mod_small tolerates the approximate reciprocal
  constexpr T sgn = INV ? (T)1 : (T)-1;
  const int lbase = l0 * lstride;
  const int mstride = m * estride, pstride = p * estride;
  C x[KB][R];
  int dsti[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int i = i0 + k * di;
    dsti[k] = -1;
    if (i < m) {
      // the first pass (p = 1) has kk = 0: no twiddles
      const int kk = p > 1 ? mod_small(i, p, inv_p) : 0;
      int o = lbase + i * estride;
#pragma unroll
      for (int r = 0; r < R; ++r, o += mstride) x[k][r] = buf[o];
      if (p > 1) {
        const int tj = kk * tstep;
        int t = 0;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          t += tj;
          const C w = tw[t];
          const C v = x[k][r];
          if (INV) {
            x[k][r].x = v.x * w.x + v.y * w.y;
            x[k][r].y = v.y * w.x - v.x * w.y;
          } else {
            x[k][r] = cmul(v, w);
          }
        }
      }
      dsti[k] = lbase + ((i - kk) * R + kk) * estride;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    if (dsti[k] >= 0) {
      Bfly<T, R>::run(x[k], sgn);
      int o = dsti[k];
#pragma unroll
      for (int q = 0; q < R; ++q, o += pstride) buf[o] = x[k][Bfly<T, R>::perm(q)];
    }
  }
  __syncthreads();
}

// 1D transforms of length n over nlines lines (nlines divides blockDim.x);
// the twiddle table (n entries) is at element tw_off of the LDS array.
template <typename T, int E, bool INV>
__device__ __forceinline__ void tile_fft(int tw_off, int n, int nf, unsigned long long f, int estride, int lstride,
                                         int nlines) {
  const int l0 = threadIdx.x % nlines, i0 = threadIdx.x / nlines, di = blockDim.x / nlines;
  int p = 1, rest = n;
  for (int s = 0; s < nf; ++s) {
    const int R = (int)((f >> (4 * s)) & 15u) + 1;
#define FBSE_PASS(RR)                                                                                    \
  if constexpr (tile_radix_fits(E, RR)) {                                                                \
    rest /= RR;                                                                                          \
    tile_radix_pass<T, RR, (E + RR - 1) / RR, INV>(tw_off, n / RR, p, rest, estride, lstride, l0, i0, di); \
  }
    switch (R) {
      case 2: FBSE_PASS(2); break;
      case 3: FBSE_PASS(3); break;
      case 4: FBSE_PASS(4); break;
      case 5: FBSE_PASS(5); break;
      case 6: FBSE_PASS(6); break;
      case 9: FBSE_PASS(9); break;
      case 10: FBSE_PASS(10); break;
      case 12: FBSE_PASS(12); break;
      case 15: FBSE_PASS(15); break;
      case 16: FBSE_PASS(16); break;
      default: FBSE_PASS(8); break;
    }
#undef FBSE_PASS
    p *= R;
  }
}

template <typename T, int KIND, int E>
__device__ __forceinline__ void tile_body(const TileArgs<T> &A) {
  using C = cplx_t<T>;
  const int nc = A.nc, nr = A.nr, ls = A.ls;
  const int NT = blockDim.x;
  C *buf = reinterpret_cast<C *>(tile_lds);
  const int twc_off = nr * ls, twr_off = twc_off + nc;
  C *twc = buf + twc_off;
  C *twr = buf + twr_off;
  for (int t = threadIdx.x; t < nc; t += NT) twc[t] = A.twc[t];
  for (int t = threadIdx.x; t < nr; t += NT) twr[t] = A.twr[t];

  // batch items vary fastest: work-groups that run together (and share an
  // XCD's L2: the dispatcher deals work-groups round-robin over the 8 XCDs)
  // transform the same plane of different states and read the same plane of
  // V_loc (or k^2) once per XCD
  const int wt = (int)(blockIdx.x / A.nbatch);
  const long item = blockIdx.x - (long)wt * A.nbatch;
  // z passes may split a slab into col_split tiles of nc columns (x ranges)
  const int w = A.col_split > 1 ? wt / A.col_split : wt;
  const int xs = wt - w * A.col_split;
  const long base = item * A.item_stride + (long)w * A.tile_stride + (long)xs * nc;
  // blockDim.x is a multiple of nc: thread -> column c0, rows r0, r0 + dr, ...
  const int c0 = threadIdx.x % nc, r0 = threadIdx.x / nc, dr = NT / nc;
  const cplx_t<T> *src = A.src + base + c0;

  // all of this thread's rows in flight at once (at most E: the launch has
  // blockDim.x * E >= nr * nc)
  {
    T vr[E], vi[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int r = r0 + u * dr;
      if (r < nr) {
        const C t = src[r * A.row_stride];
        vr[u] = t.x;
        vi[u] = t.y;
      }
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int r = r0 + u * dr;
      if (r < nr) buf[r * ls + c0] = C{vr[u], vi[u]};
    }
  }
  // z passes with a separable k^2: the x and z tables sit in LDS after the
  // twiddles, so the scaling step needs no global load inside the transform
  T *kxl = reinterpret_cast<T *>(twr + nr);
  T *kzl = kxl + nc;
  const bool ksep = (KIND == kTileKinZ || KIND == kTilePrecZ) && !A.ksqr;
  if (ksep) {
    for (int t = threadIdx.x; t < nc; t += NT) kxl[t] = A.kx[xs * nc + t];
    for (int t = threadIdx.x; t < nr; t += NT) kzl[t] = A.kz[t];
  }
  __syncthreads();

  if (KIND == kTileFwdXY) {
    tile_fft<T, E, false>(twc_off, nc, A.nfc, A.fc, 1, ls, nr);  // x rows
    tile_fft<T, E, false>(twr_off, nr, A.nfr, A.fr, ls, 1, nc);  // y columns
  } else if (KIND == kTileBwdXY || KIND == kTileBwdXYCombine) {
    tile_fft<T, E, true>(twr_off, nr, A.nfr, A.fr, ls, 1, nc);
    tile_fft<T, E, true>(twc_off, nc, A.nfc, A.fc, 1, ls, nr);
  } else {  // z columns: forward, scale by k^2 (or the preconditioner), backward
    tile_fft<T, E, false>(twr_off, nr, A.nfr, A.fr, ls, 1, nc);
    const T kxy = ksep ? kxl[c0] + A.ky[w] : (T)0;
    const T *ksq = ksep ? nullptr : A.ksqr + xs * nc + c0 + (long)w * nc * A.col_split;
    for (int r = r0; r < nr; r += dr) {
      T k2;
      if (ksep) {
        k2 = kxy + kzl[r];
        if (k2 > A.kcap) k2 = A.kcap;
      } else {
        k2 = ksq[(long)r * A.nxny];
      }
      if (KIND == kTilePrecZ) k2 = A.pa / ((T)1 + k2 * A.pb);
      C v = buf[r * ls + c0];
      v.x *= k2;
      v.y *= k2;
      buf[r * ls + c0] = v;
    }
    __syncthreads();
    tile_fft<T, E, true>(twr_off, nr, A.nfr, A.fr, ls, 1, nc);
  }

  cplx_t<T> *dst = A.dst + base + c0;
  if (KIND == kTileBwdXYCombine) {
    const T a0 = A.a0_tab ? (T)A.a0_tab[*A.step_base + A.step_off] : A.a0;
    // V is shared by the states: batch item b holds spin b % nspin
    const long vofs = (item % A.nspin) * A.item_stride + (long)w * A.tile_stride + c0;
    const T *vl = A.vloc + vofs;
    const cplx_t<T> *pin = A.psi_in + base + c0;
    const bool expand = A.rec_off != nullptr;
    constexpr int NB = 8;  // rows of psi_in and V in flight per thread
    for (int rb = r0; rb < nr; rb += dr * NB) {
      T pr[NB], pim[NB], vv[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int r = rb + u * dr;
        if (r < nr) {
          vv[u] = vl[r * A.row_stride];
          const C t = pin[r * A.row_stride];
          pr[u] = t.x;
          pim[u] = t.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int r = rb + u * dr;
        if (r < nr) {
          const C t = buf[r * ls + c0];
          C o;
          o.x = A.a1 * (t.x + vv[u] * pr[u]) + a0 * pr[u];
          o.y = A.a1 * (t.y + vv[u] * pim[u]) + a0 * pim[u];
          if (expand || A.fwd_dst) buf[r * ls + c0] = o;
          if (!expand) dst[r * A.row_stride] = o;
        }
      }
    }
    if (expand) {
      // projector expansion, one round per rank: the records of a round
      // touch distinct points, so they add in place without atomics, and
      // every point sums its atoms in index order
      __syncthreads();
      const cplx_t<T> *dcol = A.pjd + item * A.pj_natoms * 6;
      const int2 *pa = reinterpret_cast<const int2 *>(A.rec_pa);
      for (int k = 0; k < A.pt_ranks; ++k) {
        const int e0 = A.rec_off[w * A.pt_ranks + k], e1 = A.rec_off[w * A.pt_ranks + k + 1];
        for (int e = e0 + threadIdx.x; e < e1; e += 2 * NT) {
          // two records per iteration, loads of both in flight together
          const int eb = e + NT;
          const bool two = eb < e1;
          const int2 qa = pa[e];
          const int2 qb = two ? pa[eb] : qa;
          const C ta = record_term<T>(A.rec_val + (long)e * 8, dcol + (long)qa.y * 6);
          const C tb = record_term<T>(A.rec_val + (long)(two ? eb : e) * 8, dcol + (long)qb.y * 6);
          const int pos_a = (qa.x >> 16) * ls + (qa.x & 0xffff);
          C o = buf[pos_a];
          o.x += ta.x;
          o.y += ta.y;
          buf[pos_a] = o;
          if (two) {
            const int pos_b = (qb.x >> 16) * ls + (qb.x & 0xffff);
            C q = buf[pos_b];
            q.x += tb.x;
            q.y += tb.y;
            buf[pos_b] = q;
          }
        }
        __syncthreads();
      }
      for (int r = r0; r < nr; r += dr) dst[r * A.row_stride] = buf[r * ls + c0];
    }
    if (A.fwd_dst) {
      // the next Newton step's first pass on the plane just written
      __syncthreads();
      tile_fft<T, E, false>(twc_off, nc, A.nfc, A.fc, 1, ls, nr);  // x rows
      tile_fft<T, E, false>(twr_off, nr, A.nfr, A.fr, ls, 1, nc);  // y columns
      cplx_t<T> *fd = A.fwd_dst + base + c0;
      for (int r = r0; r < nr; r += dr) fd[r * A.row_stride] = buf[r * ls + c0];
    }
  } else {
    for (int r = r0; r < nr; r += dr) dst[r * A.row_stride] = buf[r * ls + c0];
  }
}

// Kernel instances: (precision, E, block-size limit, waves per SIMD).  Two
// FP32 work-groups of a 96 x 96 tile (2 x 76 KiB of LDS) share a CU; the
// waves-per-SIMD hint keeps the VGPRs low enough for both to be resident.
// An FP64 96 x 96 tile takes 152 KiB, one work-group per CU.
#define FBSE_TILE_KERNEL(NAME, T, E, MAXT, WAVES)                                                      \
  template <int KIND>                                                                                \
  __global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WAVES))) void NAME(TileArgs<T> A) { \
    tile_body<T, KIND, E>(A);                                                                        \
  }
FBSE_TILE_KERNEL(fft_tile_f32_e12, float, 12, 1024, 6)
FBSE_TILE_KERNEL(fft_tile_f32_e16, float, 16, 1024, 5)
FBSE_TILE_KERNEL(fft_tile_f32_e32, float, 32, 512, 2)
FBSE_TILE_KERNEL(fft_tile_f64_e12, double, 12, 1024, 1)
FBSE_TILE_KERNEL(fft_tile_f64_e16, double, 16, 768, 1)
#undef FBSE_TILE_KERNEL

struct TileVariant {
  int e;         // elements per thread
  int max_nt;    // block-size limit of the instance
};
static const TileVariant kTileF32[] = {{12, 1024}, {16, 1024}, {32, 512}};
static const TileVariant kTileF64[] = {{12, 1024}, {16, 768}};
template <typename T> static const TileVariant *tile_variants(int *n) {
  if (sizeof(T) == 4) {
    *n = 3;
    return kTileF32;
  }
  *n = 2;
  return kTileF64;
}

template <typename T, int KIND> static const void *tile_fn(int v) {
  if constexpr (sizeof(T) == 4) {
    return v == 0 ? (const void *)fft_tile_f32_e12<KIND>
                  : (v == 1 ? (const void *)fft_tile_f32_e16<KIND> : (const void *)fft_tile_f32_e32<KIND>);
  } else {
    return v == 0 ? (const void *)fft_tile_f64_e12<KIND> : (const void *)fft_tile_f64_e16<KIND>;
  }
}
template <typename T, int KIND>
static void tile_launch(int v, dim3 g, dim3 b, size_t lds, hipStream_t st, const TileArgs<T> &A) {
  if constexpr (sizeof(T) == 4) {
    if (v == 0)
      hipLaunchKernelGGL((fft_tile_f32_e12<KIND>), g, b, lds, st, A);
    else if (v == 1)
      hipLaunchKernelGGL((fft_tile_f32_e16<KIND>), g, b, lds, st, A);
    else
      hipLaunchKernelGGL((fft_tile_f32_e32<KIND>), g, b, lds, st, A);
  } else {
    if (v == 0)
      hipLaunchKernelGGL((fft_tile_f64_e12<KIND>), g, b, lds, st, A);
    else
      hipLaunchKernelGGL((fft_tile_f64_e16<KIND>), g, b, lds, st, A);
  }
}

// ---------------------------------------------------------------------------
// Fixed-schedule persistent instances: cubic N^3 grids with N = R1 * R2
// ---------------------------------------------------------------------------
// Each pass of the instances above is a function call with runtime line
// lengths, strides and radices, and a call starts by waiting for every
// outstanding memory operation of the wave.  For the grids below the radix
// schedule, the line length and every stride are compile-time constants: the
// passes are inlined with constant LDS offsets and index arithmetic, and each
// work-group loops over tiles (the launch holds as many work-groups as stay
// resident), so the twiddle and k^2 tables are loaded once per work-group.
// Rows of the next tile (PF) and of psi_in / V (PC) can be loaded ahead into
// registers.  Same arithmetic, in the same order, as the generic passes.

// One Stockham pass of radix R over NL lines of length N, P = product of the
// radices already done; element j of line l at l*LSTR + j*ES.
template <typename T, int N, int R, int P, int NL, int ES, int LSTR, int NT, bool INV>
__device__ __forceinline__ void fx_pass(int tw_off) {
  using C = cplx_t<T>;
  constexpr int M = N / R;       // butterflies per line
  constexpr int DI = NT / NL;    // threads per line
  constexpr int KB = (M + DI - 1) / DI;
  constexpr int TSTEP = N / (P * R);
  constexpr T sgn = INV ? (T)1 : (T)-1;
  static_assert(NT % NL == 0 && N % R == 0, "fixed pass shape");
  C *buf = reinterpret_cast<C *>(tile_lds);
  const C *tw = buf + tw_off;
  // opaque thread index: the LDS addresses below are recomputed in every
  // pass (a few integer operations) instead of being hoisted out of the tile
  // loop, where they would hold registers the butterflies need
  unsigned tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int l0 = (int)(tid % NL), i0 = (int)(tid / NL);  // unsigned: shifts for power-of-two NL
  const int lbase = l0 * LSTR;
  C x[KB][R];
  int dsti[KB], tjk[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int i = i0 + k * DI;
    dsti[k] = -1;
    if (KB * DI == M || i < M) {
      const int kk = P > 1 ? i % P : 0;
#pragma unroll
      for (int r = 0; r < R; ++r) x[k][r] = buf[lbase + (i + r * M) * ES];
      tjk[k] = kk * TSTEP;
      dsti[k] = lbase + ((i - kk) * R + kk) * ES;
    }
  }
  __syncthreads();
  // twiddles after the barrier (the table is never written by the passes):
  // only one butterfly's roots are live at a time
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    if (dsti[k] >= 0) {
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const C w = tw[r * tjk[k]];
          const C v = x[k][r];
          if (INV) {
            x[k][r].x = v.x * w.x + v.y * w.y;
            x[k][r].y = v.y * w.x - v.x * w.y;
          } else {
            x[k][r] = cmul(v, w);
          }
        }
      }
      Bfly<T, R>::run(x[k], sgn);
#pragma unroll
      for (int q = 0; q < R; ++q) buf[dsti[k] + q * P * ES] = x[k][Bfly<T, R>::perm(q)];
    }
  }
  __syncthreads();
}

template <typename T, int N, int R1, int R2, int NL, int ES, int LSTR, int NT, bool INV>
__device__ __forceinline__ void fx_fft(int tw_off) {
  fx_pass<T, N, R1, 1, NL, ES, LSTR, NT, INV>(tw_off);
  fx_pass<T, N, R2, R1, NL, ES, LSTR, NT, INV>(tw_off);
}

// PF = rows of the next tile loaded during the current tile's transforms
// (the rest are loaded when the tile starts), PC = rows of psi_in and V the
// combine pass loads before its backward transforms (the rest after them):
// registers held across the transforms, bounded by the occupancy target.
template <typename T, int KIND, int N, int R1, int R2, int NT, int PF, int PC>
__device__ __forceinline__ void fx_body(const TileArgs<T> &A) {
  using C = cplx_t<T>;
  constexpr int LS = N | 1;         // LDS row pitch (odd)
  constexpr int E = N * N / NT;     // rows per thread
  constexpr int DR = NT / N;        // row step
  static_assert(NT % N == 0 && E * DR == N && R1 * R2 == N, "fixed tile shape");
  constexpr bool XY = KIND == kTileFwdXY || KIND == kTileBwdXYCombine;
  constexpr long ROW = XY ? (long)N : (long)N * N;  // global elements between tile rows
  constexpr long TILE = XY ? (long)N * N : (long)N; // between tiles of one item
  constexpr long ITEM = (long)N * N * N;
  constexpr int twc_off = N * LS, twr_off = twc_off + N;
  C *buf = reinterpret_cast<C *>(tile_lds);
  C *twc = buf + twc_off;
  C *twr = buf + twr_off;
  T *kxl = reinterpret_cast<T *>(twr + N);
  T *kzl = kxl + N;
  // k^2 is separable for these instances (the host checks): x and z tables in LDS
  constexpr bool ksep = KIND == kTileKinZ;
  for (int t = threadIdx.x; t < N; t += NT) {
    twc[t] = A.twc[t];
    twr[t] = A.twr[t];
    if (ksep) {
      kxl[t] = A.kx[t];
      kzl[t] = A.kz[t];
    }
  }
  const int c0 = threadIdx.x % N, r0 = threadIdx.x / N;
  const unsigned nb = (unsigned)A.nbatch;
  const unsigned ntiles = nb * (unsigned)A.tiles_per_item;  // < 2^31 (checked by the host)
  const unsigned G = gridDim.x;
  // Global rows are addressed as (per-tile base pointer, uniform) + (32-bit
  // byte offset of the lane's row within the item, < N^3 * sizeof(C)).  The
  // empty asm at the top of each tile makes the lane offset opaque, so the
  // per-row offsets are recomputed per tile instead of being hoisted out of
  // the loop into registers the transforms need.
  unsigned voff = (unsigned)(c0 + r0 * ROW) * (unsigned)sizeof(C);
  constexpr unsigned RSTEP = (unsigned)(DR * ROW * sizeof(C));
  auto row_ptr = [&](const void *tb, int u) -> const C * {
    return reinterpret_cast<const C *>(reinterpret_cast<const char *>(tb) + (size_t)(voff + (unsigned)u * RSTEP));
  };
  auto row_ptr_w = [&](void *tb, int u) -> C * {
    return reinterpret_cast<C *>(reinterpret_cast<char *>(tb) + (size_t)(voff + (unsigned)u * RSTEP));
  };
  auto tile_off = [&](unsigned tt) -> long {
    const unsigned w = tt / nb, item = tt - w * nb;
    return (long)item * ITEM + (long)w * TILE;
  };
  static_assert(PF >= 0 && PF <= E && PC >= 0 && PC <= E, "prefetch depth");
  T vr[E], vi[E];
  auto issue = [&](unsigned tt, auto u0, auto u1) {
    const C *s = A.src + tile_off(tt);
#pragma unroll
    for (int u = decltype(u0)::value; u < decltype(u1)::value; ++u) {
      const C v = *row_ptr(s, u);
      vr[u] = v.x;
      vi[u] = v.y;
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using IPF = std::integral_constant<int, PF>;
  using IE = std::integral_constant<int, E>;
  unsigned t = blockIdx.x;
  if (t < ntiles) issue(t, I0{}, IPF{});
  for (; t < ntiles; t += G) {
    asm volatile("" : "+v"(voff));
    const unsigned w = t / nb, item = t - w * nb;
    const long tb = (long)item * ITEM + (long)w * TILE;
    issue(t, IPF{}, IE{});  // rows not prefetched
    __syncthreads();  // the previous tile's LDS reads are done (first tile: the tables are written)
#pragma unroll
    for (int u = 0; u < E; ++u) buf[(r0 + u * DR) * LS + c0] = C{vr[u], vi[u]};
    __syncthreads();
    const bool more = t + G < ntiles;
    if (KIND == kTileFwdXY) {
      if (more) issue(t + G, I0{}, IPF{});
      fx_fft<T, N, R1, R2, N, 1, LS, NT, false>(twc_off);  // x rows
      fx_fft<T, N, R1, R2, N, LS, 1, NT, false>(twr_off);  // y columns
      C *d = A.dst + tb;
#pragma unroll
      for (int u = 0; u < E; ++u) *row_ptr_w(d, u) = buf[(r0 + u * DR) * LS + c0];
    } else if (KIND == kTileKinZ) {
      if (more) issue(t + G, I0{}, IPF{});
      fx_fft<T, N, R1, R2, N, LS, 1, NT, false>(twr_off);  // z columns
      const T kxy = kxl[c0] + A.ky[w];
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int r = r0 + u * DR;
        T k2 = kxy + kzl[r];
        if (k2 > A.kcap) k2 = A.kcap;
        C v = buf[r * LS + c0];
        v.x *= k2;
        v.y *= k2;
        buf[r * LS + c0] = v;
      }
      __syncthreads();
      fx_fft<T, N, R1, R2, N, LS, 1, NT, true>(twr_off);
      C *d = A.dst + tb;
#pragma unroll
      for (int u = 0; u < E; ++u) *row_ptr_w(d, u) = buf[(r0 + u * DR) * LS + c0];
    } else {  // kTileBwdXYCombine
      const T a0 = A.a0_tab ? (T)A.a0_tab[*A.step_base + A.step_off] : A.a0;
      // psi_in and V rows of this tile, in flight during the backward transforms
      const T *vl = A.vloc + (long)(item % (unsigned)A.nspin) * ITEM + (long)w * TILE + c0;
      const C *pin = A.psi_in + tb;
      T pr[E], pim[E], vv[E];
      auto issue_in = [&](auto u0, auto u1) {
#pragma unroll
        for (int u = decltype(u0)::value; u < decltype(u1)::value; ++u) {
          vv[u] = vl[(long)(r0 + u * DR) * ROW];
          const C q = *row_ptr(pin, u);
          pr[u] = q.x;
          pim[u] = q.y;
        }
      };
      issue_in(I0{}, std::integral_constant<int, PC>{});
      fx_fft<T, N, R1, R2, N, LS, 1, NT, true>(twr_off);  // y columns
      fx_fft<T, N, R1, R2, N, 1, LS, NT, true>(twc_off);  // x rows
      issue_in(std::integral_constant<int, PC>{}, IE{});
      const bool expand = A.rec_off != nullptr;
      C *d = A.dst + tb;
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const int r = r0 + u * DR;
        const C q = buf[r * LS + c0];
        C o;
        o.x = A.a1 * (q.x + vv[u] * pr[u]) + a0 * pr[u];
        o.y = A.a1 * (q.y + vv[u] * pim[u]) + a0 * pim[u];
        if (expand || A.fwd_dst) buf[r * LS + c0] = o;
        if (!expand) *row_ptr_w(d, u) = o;
      }
      if (more) issue(t + G, I0{}, IPF{});
      if (expand) {
        // projector expansion in rounds per rank, as in tile_body; the
        // positions and atoms of a thread's next two records are loaded one
        // step ahead (across the round barrier too), so each step waits once
        // (records and coefficients) instead of twice (positions, then the
        // coefficients of their atoms)
        __syncthreads();
        const C *dcol = A.pjd + (long)item * A.pj_natoms * 6;
        const int2 *pa = reinterpret_cast<const int2 *>(A.rec_pa);
        const int *ro = A.rec_off + w * A.pt_ranks;
        const int nr = A.pt_ranks;
        int2 qa = int2{0, 0}, qb = int2{0, 0};
        int pf = -1;  // qa, qb hold the first step of round pf
        auto first_of = [&](int k, int2 &a, int2 &b) {
          const int s0 = ro[k] + (int)threadIdx.x, s1 = ro[k + 1];
          if (s0 < s1) a = pa[s0];
          if (s0 + NT < s1) b = pa[s0 + NT];
        };
        for (int k = 0; k < nr; ++k) {
          const int e1 = ro[k + 1];
          if (pf != k) first_of(k, qa, qb);
          for (int e = ro[k] + threadIdx.x; e < e1; e += 2 * NT) {
            const int eb = e + NT;
            const bool two = eb < e1;
            int2 na = qa, nb = qb;
            const int en = e + 2 * NT;
            if (en < e1) {
              na = pa[en];
              if (en + NT < e1) nb = pa[en + NT];
            } else if (k + 1 < nr) {
              first_of(k + 1, na, nb);
              pf = k + 1;
            }
            const C ta = record_term<T>(A.rec_val + (long)e * 8, dcol + (long)qa.y * 6);
            const C tb = record_term<T>(A.rec_val + (long)(two ? eb : e) * 8, dcol + (long)(two ? qb.y : qa.y) * 6);
            const int pos_a = (qa.x >> 16) * LS + (qa.x & 0xffff);
            C o = buf[pos_a];
            o.x += ta.x;
            o.y += ta.y;
            buf[pos_a] = o;
            if (two) {
              const int pos_b = (qb.x >> 16) * LS + (qb.x & 0xffff);
              C q = buf[pos_b];
              q.x += tb.x;
              q.y += tb.y;
              buf[pos_b] = q;
            }
            qa = na;
            qb = nb;
          }
          __syncthreads();
        }
#pragma unroll
        for (int u = 0; u < E; ++u) *row_ptr_w(d, u) = buf[(r0 + u * DR) * LS + c0];
      }
      if (A.fwd_dst) {
        __syncthreads();
        fx_fft<T, N, R1, R2, N, 1, LS, NT, false>(twc_off);  // x rows
        fx_fft<T, N, R1, R2, N, LS, 1, NT, false>(twr_off);  // y columns
        C *fd = A.fwd_dst + tb;
#pragma unroll
        for (int u = 0; u < E; ++u) *row_ptr_w(fd, u) = buf[(r0 + u * DR) * LS + c0];
      }
    }
  }
}

#define FBSE_FX_KERNEL(NAME, T, N, R1, R2, NT, WAVES, PF, PC)                                          \
  template <int KIND>                                                                                 \
  __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES))) void NAME(TileArgs<T> A) { \
    fx_body<T, KIND, N, R1, R2, NT, PF, PC>(A);                                                       \
  }
// No register prefetch (PF = PC = 0): on 128 FP32 grids of 96^3 every
// prefetch depth tried (6 or 12 rows of the next tile, 6 or 12 rows of
// psi_in/V) ran 5-55% slower than none, because the rows held across the
// transforms push the waves past the 80 VGPRs that two resident work-groups
// per CU allow.  Without it the fixed instances take the forward xy pass from
// 569 to 396 us, the z-slab pass from 615 to 423 us and the combine pass
// (with the next forward transform) from 978 to 770 us.
//
// Instances (N, radices, threads, waves per SIMD): threads = N * rows-step,
// chosen so that every thread stages the same number of rows and the radix
// passes keep at most two butterflies per thread; the waves hint lets two
// work-groups share a CU where the tile's LDS allows it (N <= 90 in FP32),
// one otherwise.  FP64 tiles above 96 do not fit the LDS.
#define FBSE_FX_F32(X)            \
  X(60, 10, 6, 600, 5)            \
  X(64, 8, 8, 512, 4)             \
  X(72, 9, 8, 648, 6)             \
  X(80, 10, 8, 640, 5)            \
  X(90, 10, 9, 540, 5)            \
  X(96, 12, 8, 768, 6)            \
  X(100, 10, 10, 1000, 4)         \
  X(108, 12, 9, 972, 4)           \
  X(120, 15, 8, 960, 4)
#define FBSE_FX_F64(X)            \
  X(64, 8, 8, 512, 4)             \
  X(96, 12, 8, 768, 3)
#define FBSE_FX_K32(N, R1, R2, NT, W) FBSE_FX_KERNEL(fft_fx_f32_n##N, float, N, R1, R2, NT, W, 0, 0)
#define FBSE_FX_K64(N, R1, R2, NT, W) FBSE_FX_KERNEL(fft_fx_f64_n##N, double, N, R1, R2, NT, W, 0, 0)
FBSE_FX_F32(FBSE_FX_K32)
FBSE_FX_F64(FBSE_FX_K64)
#undef FBSE_FX_K32
#undef FBSE_FX_K64
#undef FBSE_FX_KERNEL

// Grouped launches of the FP32 fixed instances: one launch runs a pass for
// up to kTileGroupMax contexts of one grid size (the dots of the batched
// uneven-dot filter), blockIdx.y selecting the context's arguments and
// blockIdx.x the persistent work-group of that context.  Each tile runs the
// arithmetic of the single-context instance, so results are identical.
template <typename T> struct TileGroup {
  TileArgs<T> a[kTileGroupMax];
};
static_assert(sizeof(TileGroup<float>) <= 4096, "grouped tile arguments exceed the kernel argument space");
#define FBSE_FX_GKERNEL(NAME, T, N, R1, R2, NT, WAVES)                                                    \
  template <int KIND>                                                                                    \
  __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES))) void NAME(TileGroup<T> G) { \
    fx_body<T, KIND, N, R1, R2, NT, 0, 0>(G.a[blockIdx.y]);                                              \
  }
#define FBSE_FX_G32(N, R1, R2, NT, W) FBSE_FX_GKERNEL(fft_fx_f32_n##N##_grp, float, N, R1, R2, NT, W)
FBSE_FX_F32(FBSE_FX_G32)
#undef FBSE_FX_G32
#undef FBSE_FX_GKERNEL

// ---------------------------------------------------------------------------
// Fixed-schedule per-axis passes (axis lengths 160 = 16*10 and 256 = 16*16)
// ---------------------------------------------------------------------------
// The per-axis design of fft_kernels.hip (tiles of L lines of one axis, five
// sweeps per H) with the compile-time machinery above: the lines of a tile
// sit in LDS with an odd pitch and are transformed in place by the inlined
// passes, and each work-group loops over tiles.  x tiles are L consecutive
// rows; y and z tiles are L consecutive inner indices (each axis point a
// contiguous 8*L- or 16*L-byte segment).
// SP (spin-paired combine, x rows): a tile holds rows y0 .. y0+L/2-1 of one
// plane of both spins of one state (lines l < L/2 spin up, the rest spin
// down), so the tile's projector records are read once for the two grids.
// strided (y / z) passes with the next tile's loads in flight (axis_fx_strided),
// FP32 only: the FP64 instances, two work-groups per CU, ran 4% slower with it
constexpr bool kAxisPipelined = true;

// y and z passes (strided lines; forward, backward, k^2 or preconditioner):
// a work-group's next tile is loaded into registers while the current tile
// is transformed, so that its HBM reads overlap the LDS sweeps.  Same tiles,
// same arithmetic, as the plain loop of axis_fx_body.
template <typename T, int KIND, int N, int R1, int R2, int L, int NT>
__device__ __forceinline__ void axis_fx_strided(const AxisArgs<T> &A) {
  using C = cplx_t<T>;
  constexpr int LS = N | 1, E = N * L / NT;
  static_assert(NT % L == 0 && E * NT == N * L && R1 * R2 == N, "fixed axis tile shape");
  constexpr int tw_off = L * LS;
  C *buf = reinterpret_cast<C *>(tile_lds);
  C *tw = buf + tw_off;
  T *kzl = reinterpret_cast<T *>(tw + N);
  T *kxyl = kzl + N;
  constexpr bool KIN = KIND == kPassKin || KIND == kPassPrec;
  constexpr bool INV = KIND == kPassBwd;
  for (int t = threadIdx.x; t < N; t += NT) {
    tw[t] = A.tw[t];
    if (KIN) kzl[t] = A.kz[t];
  }
  // element e = tid + u*NT of a tile: line e % L, axis point e / L; source
  // and result in the reference or the z-tile layout (AxisArgs::zt_*)
  unsigned tid = threadIdx.x;
  const bool zs = L == 16 && A.zt_src, zd = L == 16 && A.zt_dst;
  const unsigned inner = (unsigned)A.inner, zinner = (unsigned)A.zt_inner;
  struct Geo {
    long base, zbase, i0;
    int nl;
  };
  auto geo = [&](long t) {
    Geo g;
    const long o = t / A.tiles_per_outer;
    g.i0 = (t - o * A.tiles_per_outer) * L;
    g.base = o * A.outer_stride + g.i0;
    g.zbase = (zs || zd) ? (o / A.zt_q) * A.zt_grid + (o % A.zt_q) * 16 + (g.i0 / 16) * A.zt_nz16 : 0;
    const long rem = A.inner_count - g.i0;
    g.nl = rem < L ? (int)rem : L;
    return g;
  };
  C pre[E];
  auto load_tile = [&](const Geo &g) {
    const C *src = A.src + (zs ? g.zbase : g.base);
    const unsigned st = zs ? zinner : inner;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const unsigned e = tid + (unsigned)(u * NT);
      if ((int)(e % L) < g.nl) pre[u] = src[(e / L) * st + e % L];
    }
  };
  long t = blockIdx.x;
  Geo cur{};
  if (t < A.ntiles) {
    cur = geo(t);
    load_tile(cur);
  }
  for (; t < A.ntiles; t += gridDim.x) {
    asm volatile("" : "+v"(tid));  // per-tile recomputation of the lane's offsets
    const int itid = (int)tid;
    __syncthreads();  // the previous tile's LDS reads are done (first tile: tables written)
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const unsigned e = tid + (unsigned)(u * NT);
      if ((int)(e % L) < cur.nl) buf[(e % L) * LS + e / L] = pre[u];
    }
    if (KIN) {
      for (int l = itid; l < cur.nl; l += NT) {
        const long in = cur.i0 + l;
        const long jy = in / A.nx, jx = in - jy * A.nx;
        kxyl[l] = A.kx[jx] + A.ky[jy];
      }
    }
    // the next tile's loads are in flight during this tile's transforms
    const long tn = t + gridDim.x;
    Geo nxt{};
    if (tn < A.ntiles) {
      nxt = geo(tn);
      load_tile(nxt);
    }
    __syncthreads();
    // lines l >= nl of a last partial tile hold stale values: transformed, never stored
    fx_fft<T, N, R1, R2, L, 1, LS, NT, INV>(tw_off);
    if (KIN) {
#pragma unroll 4
      for (int u = 0; u < E; ++u) {
        const unsigned e = tid + (unsigned)(u * NT);
        const int l = (int)(e / N), p = (int)(e % N);
        T k2 = kxyl[l] + kzl[p];
        if (k2 > A.kcap) k2 = A.kcap;
        if (KIND == kPassPrec) k2 = A.pa / ((T)1 + k2 * A.pb);
        C v = buf[l * LS + p];
        v.x *= k2;
        v.y *= k2;
        buf[l * LS + p] = v;
      }
      __syncthreads();
      fx_fft<T, N, R1, R2, L, 1, LS, NT, true>(tw_off);
    }
    C *dst = A.dst + (zd ? cur.zbase : cur.base);
    const unsigned dt = zd ? zinner : inner;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const unsigned e = tid + (unsigned)(u * NT);
      if ((int)(e % L) < cur.nl) dst[(e / L) * dt + e % L] = buf[(e % L) * LS + e / L];
    }
    cur = nxt;
  }
}

template <typename T, int KIND, int N, int R1, int R2, int L, int NT, bool XAX, int EXP_GROUP, bool SP = false>
__device__ __forceinline__ void axis_fx_body(const AxisArgs<T> &A) {
  using C = cplx_t<T>;
  static_assert(!SP || (XAX && KIND == kPassBwdCombine && L % 2 == 0), "spin-paired tiles are combine tiles");
  if constexpr (!XAX && KIND != kPassBwdCombine && kAxisPipelined && sizeof(T) == 4) {
    axis_fx_strided<T, KIND, N, R1, R2, L, NT>(A);
    return;
  }
  constexpr int HL = SP ? L / 2 : L;  // rows of one grid in a tile
  constexpr int LS = N | 1, E = N * L / NT;
  constexpr int LB = E % 8 == 0 ? 8 : (E % 5 == 0 ? 5 : E);  // global rows in flight per batch
  static_assert(NT % L == 0 && E * NT == N * L && E % LB == 0 && R1 * R2 == N, "fixed axis tile shape");
  constexpr int tw_off = L * LS;
  C *buf = reinterpret_cast<C *>(tile_lds);
  C *tw = buf + tw_off;
  T *kzl = reinterpret_cast<T *>(tw + N);
  T *kxyl = kzl + N;
  constexpr bool KIN = KIND == kPassKin || KIND == kPassPrec;
  constexpr bool INV = KIND == kPassBwd || KIND == kPassBwdCombine;
  for (int t = threadIdx.x; t < N; t += NT) {
    tw[t] = A.tw[t];
    if (KIN) kzl[t] = A.kz[t];
  }
  // element e = tid + u*NT of a tile: line l, point p; global offset (within
  // the tile, elements) XAX ? e : p*inner + l, below 2^31 / sizeof(C)
  unsigned tid = threadIdx.x;
  for (long t = blockIdx.x; t < A.ntiles; t += gridDim.x) {
    asm volatile("" : "+v"(tid));  // per-tile recomputation of the lane's offsets
    // (unsigned: the element indices below divide by powers of two as shifts)
    const int itid = (int)tid;
    long base, i0 = 0, line0 = 0, zbase = 0;
    int nl;
    if (XAX) {
      long tt = t;
      // batch items of the tile order: grids, or states when spin-paired
      const long rows = (long)A.ny * A.nz, nbat = A.total_lines / rows / (SP ? 2 : 1);
      if (KIND == kPassBwdCombine && A.xcd_order) {
        // t -> XCD t mod 8, place s = t / 8 in that XCD's sequence of (row
        // block, batch item) pairs, row blocks xcd, xcd + 8, ... (the host
        // sets xcd_order only when the row blocks per item are a multiple of 8)
        const long s = t >> 3, q = s / nbat;
        tt = (s - q * nbat) * (rows / HL) + (t & 7) + 8 * q;
      }
      if (SP) {
        const long st = tt / (rows / HL);
        line0 = st * 2 * rows + (tt - st * (rows / HL)) * HL;  // spin-up grid's first row
        nl = L;
      } else {
        line0 = tt * L;
        const long rem = A.total_lines - line0;
        nl = rem < L ? (int)rem : L;
      }
      base = line0 * N;
    } else {
      const long o = t / A.tiles_per_outer;
      i0 = (t - o * A.tiles_per_outer) * L;
      base = o * A.outer_stride + i0;
      if (!XAX && L == 16 && (A.zt_src || A.zt_dst))
        zbase = (o / A.zt_q) * A.zt_grid + (o % A.zt_q) * 16 + (i0 / 16) * A.zt_nz16;
      const long rem = A.inner_count - i0;
      nl = rem < L ? (int)rem : L;
    }
    const unsigned inner = (unsigned)A.inner;
    // spin-paired: line l at (l / HL) grids and (l % HL) rows from the tile base
    const unsigned rowsN = (unsigned)A.ny * (unsigned)A.nz * (unsigned)N;
    auto xoff = [&](unsigned e) -> unsigned {
      return SP ? (unsigned)((e / N) / HL) * rowsN + (unsigned)(((e / N) % HL) * N + e % N) : (unsigned)e;
    };
    auto goff = [&](unsigned e) -> unsigned {
      return XAX ? xoff(e) : (unsigned)(e / L) * inner + (unsigned)(e % L);
    };
    // y / z lines: source and result in the reference or the z-tile layout
    const bool zs = !XAX && L == 16 && A.zt_src, zd = !XAX && L == 16 && A.zt_dst;
    const unsigned zinner = (unsigned)A.zt_inner;
    auto soff = [&](unsigned e) -> unsigned { return zs ? (unsigned)(e / L) * zinner + (unsigned)(e % L) : goff(e); };
    auto doff = [&](unsigned e) -> unsigned { return zd ? (unsigned)(e / L) * zinner + (unsigned)(e % L) : goff(e); };
    auto lpos = [&](unsigned e) -> int { return (int)(XAX ? (e / N) * LS + e % N : (e % L) * LS + e / L); };
    auto line = [&](unsigned e) -> int { return (int)(XAX ? e / N : e % L); };
    const C *src = A.src + (zs ? zbase : base);
    __syncthreads();  // the previous tile's LDS reads are done (first tile: tables written)
#pragma unroll
    for (int ub = 0; ub < E; ub += LB) {
      T vr[LB], vi[LB];
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        const unsigned e = tid + (unsigned)((ub + u) * NT);
        if (line(e) < nl) {
          const C v = src[soff(e)];
          vr[u] = v.x;
          vi[u] = v.y;
        }
      }
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        const unsigned e = tid + (unsigned)((ub + u) * NT);
        if (line(e) < nl) buf[lpos(e)] = C{vr[u], vi[u]};
      }
    }
    if (KIN) {
      for (int l = itid; l < nl; l += NT) {
        const long in = i0 + l;
        const long jy = in / A.nx, jx = in - jy * A.nx;
        kxyl[l] = A.kx[jx] + A.ky[jy];
      }
    }
    __syncthreads();
    // lines l >= nl of a last partial tile hold stale values: transformed, never stored
    fx_fft<T, N, R1, R2, L, 1, LS, NT, INV>(tw_off);
    if (KIN) {
      // partly unrolled: the scaling step's loads stay within the registers
      // the transforms leave free
#pragma unroll 4
      for (int u = 0; u < E; ++u) {
        const unsigned e = tid + (unsigned)(u * NT);
        const int l = (int)(e / N), p = (int)(e % N);
        T k2 = kxyl[l] + kzl[p];
        if (k2 > A.kcap) k2 = A.kcap;
        if (KIND == kPassPrec) k2 = A.pa / ((T)1 + k2 * A.pb);
        C v = buf[l * LS + p];
        v.x *= k2;
        v.y *= k2;
        buf[l * LS + p] = v;
      }
      __syncthreads();
      fx_fft<T, N, R1, R2, L, 1, LS, NT, true>(tw_off);
    }
    C *dst = A.dst + (zd ? zbase : base);
    if (KIND == kPassBwdCombine) {  // x axis: out = a1*(T psi + V psi_in) + a0*psi_in
      const T a0 = A.a0_tab ? (T)A.a0_tab[*A.step_base + A.step_off] : A.a0;
      const C *pin = A.psi_in + base;
      const bool expand = A.rec_yoff != nullptr;
      const bool keep = expand || A.fwd_dst != nullptr;
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const unsigned e = tid + (unsigned)(u * NT);
        const int l = (int)(e / N), p = (int)(e % N);
        if (l < nl) {
          const long row = (line0 + (SP ? (long)(l / HL) * A.ny * A.nz + l % HL : l)) % A.vloc_rows;
          const T vv = A.vloc[row * N + p];
          const C q = pin[xoff(e)];
          const C r = buf[l * LS + p];
          C o;
          o.x = A.a1 * (r.x + vv * q.x) + a0 * q.x;
          o.y = A.a1 * (r.y + vv * q.y) + a0 * q.y;
          if (keep) buf[l * LS + p] = o;
          if (!expand) dst[xoff(e)] = o;
        }
      }
      if (expand) {
        // projector expansion of the tile's rows (one plane z, rows y0 ..
        // y0+L-1 of batch item `item`) in rounds per rank: the records of one
        // rank touch distinct points, so every point adds its atoms in index
        // order, as the separate expansion pass does.  The rank segments of
        // the tile (seg_start[k], seg_len[k], in the k^2 tables of the LDS,
        // which the combine does not use) are read once.  Inside a round a
        // thread takes kG records at a time and issues all their loads before
        // it adds any term: a record's loads wait on its position and atom,
        // and a loop of one record per step would wait once per record.
        // (spin-paired: every record adds its term to the spin-up row with the
        // spin-up coefficients and to the spin-down row with the spin-down ones)
        constexpr int kG = SP ? (EXP_GROUP + 1) / 2 : EXP_GROUP;
        const long rows = (long)A.ny * A.nz;
        const long item = line0 / rows, zy = line0 - item * rows;
        const int z = (int)(zy / A.ny), y0 = (int)(zy - (long)z * A.ny);
        const C *dcol = A.pjd + item * A.pj_natoms * 6;
        const C *dcol2 = dcol + (long)A.pj_natoms * 6;  // SP: the spin-down grid (item + 1)
        const int2 *pa = reinterpret_cast<const int2 *>(A.rec_pa);
        const int nr = A.pt_ranks;
        int *seg_start = reinterpret_cast<int *>(kzl), *seg_len = seg_start + nr;
        if (itid < nr) {
          const int *ro = A.rec_yoff + ((long)z * nr + itid) * (A.ny + 1) + y0;
          const int s0 = ro[0];
          seg_start[itid] = s0;
          seg_len[itid] = ro[HL] - s0;
        }
        __syncthreads();  // also orders the combine's LDS writes before the rounds
        if constexpr (SP) {
          // A thread's first record of a round (position, atom, values) is
          // loaded while the round before is processed (across its barrier), so
          // that record waits once, for its atom's coefficients; the rest of
          // the round takes kG records at a time, their loads issued together.
          auto load_one = [&](int k, int i, int2 &q, T (&rv)[8]) {
            if (i < seg_len[k]) {
              const int e = seg_start[k] + i;
              q = pa[e];
              load_record<T>(A.rec_val + (long)e * 8, rv);
            } else {
              q = int2{-1, 0};
            }
          };
          auto add_one = [&](const int2 &q, const C &tv, const C &tv2) {
            if (q.x < 0) return;
            const int pos = ((q.x >> 16) - y0) * LS + (q.x & 0xffff);
            C o = buf[pos];
            o.x += tv.x;
            o.y += tv.y;
            buf[pos] = o;
            C o2 = buf[pos + HL * LS];
            o2.x += tv2.x;
            o2.y += tv2.y;
            buf[pos + HL * LS] = o2;
          };
          auto terms = [&](const int2 &q, const T (&rv)[8], C &tv, C &tv2) {
            if (q.x < 0) return;
            tv = record_term_r<T>(rv, dcol + (long)q.y * 6);
            tv2 = record_term_r<T>(rv, dcol2 + (long)q.y * 6);
          };
          // round k from the preloaded (cq, cv), which then take round k + 1's
          // first record (in flight during the round's LDS adds and barrier)
          auto round = [&](int k, int2 &cq, T (&cv)[8]) {
            {
              C tv{}, tv2{};
              terms(cq, cv, tv, tv2);
              const int2 q0 = cq;
              if (k + 1 < nr) load_one(k + 1, itid, cq, cv);
              add_one(q0, tv, tv2);
            }
            for (int b0 = NT; b0 < seg_len[k]; b0 += kG * NT) {
              int2 q[kG];
              T rv[kG][8];
              C tv[kG], tv2[kG];
#pragma unroll
              for (int g = 0; g < kG; ++g) load_one(k, b0 + g * NT + itid, q[g], rv[g]);
#pragma unroll
              for (int g = 0; g < kG; ++g) terms(q[g], rv[g], tv[g], tv2[g]);
#pragma unroll
              for (int g = 0; g < kG; ++g) add_one(q[g], tv[g], tv2[g]);
            }
            __syncthreads();
          };
          int2 qa;
          T va[8];
          if (nr > 0) load_one(0, itid, qa, va);
          for (int k = 0; k < nr; ++k) round(k, qa, va);
        } else {
          // inside a round a thread takes kG records at a time and issues all
          // their loads before it adds any term
          for (int k = 0; k < nr; ++k) {
            const int s0 = seg_start[k], n0 = seg_len[k];
            if (n0 == 0) continue;
            for (int b0 = 0; b0 < n0; b0 += kG * NT) {
              int2 q[kG];
              C tv[kG];
#pragma unroll
              for (int g = 0; g < kG; ++g) {
                const int i = b0 + g * NT + itid;
                q[g] = i < n0 ? pa[s0 + i] : int2{-1, 0};
              }
#pragma unroll
              for (int g = 0; g < kG; ++g)
                if (q[g].x >= 0)
                  tv[g] = record_term<T>(A.rec_val + (long)(s0 + b0 + g * NT + itid) * 8, dcol + (long)q[g].y * 6);
#pragma unroll
              for (int g = 0; g < kG; ++g) {
                if (q[g].x >= 0) {
                  const int pos = ((q[g].x >> 16) - y0) * LS + (q[g].x & 0xffff);
                  C o = buf[pos];
                  o.x += tv[g].x;
                  o.y += tv[g].y;
                  buf[pos] = o;
                }
              }
            }
            __syncthreads();
          }
        }
#pragma unroll 4
        for (int u = 0; u < E; ++u) {
          const unsigned e = tid + (unsigned)(u * NT);
          const int l = (int)(e / N), p = (int)(e % N);
          if (l < nl) dst[xoff(e)] = buf[l * LS + p];
        }
        if constexpr (SP) {
          if (A.pp_part) {
            // raw projections of the output rows for the next Newton step: a
            // lane takes kPpRecs records of one atom (contiguous loads of
            // their positions and values), the lanes of a segment (one atom's
            // lanes within the wavefront) are summed by a segmented scan (a
            // fixed order) over the tile's longest segment, and the segment's
            // last lane writes its slot
            constexpr int NR = kPpRecs;
            const long tau = (long)z * (A.ny / HL) + y0 / HL;
            const int l0 = A.pp_lane_off[tau], nln = A.pp_lane_off[tau + 1] - l0, span = A.pp_span[tau];
            C *part = A.pp_part + (item / 2) * A.pp_nslots * 12;
            const int wl = itid & 63;
            for (int r0 = 0; r0 < nln; r0 += NT) {
              const int li = r0 + itid;
              const bool live = li < nln;
              int lw = 1;  // lanes past the tile's: segments of their own, no slot
              int mq[NR];
              T rv[NR][8];
              if (live) {
                lw = A.pp_lanes[l0 + li];
                const int4 *m4 = reinterpret_cast<const int4 *>(A.pp_meta) + (long)(l0 + li) * (NR / 4);
#pragma unroll
                for (int h = 0; h < NR / 4; ++h) {
                  const int4 t = m4[h];
                  mq[4 * h] = t.x, mq[4 * h + 1] = t.y, mq[4 * h + 2] = t.z, mq[4 * h + 3] = t.w;
                }
#pragma unroll
                for (int g = 0; g < NR; ++g) load_record<T>(A.pp_val + ((long)(l0 + li) * NR + g) * 8, rv[g]);
              } else {
#pragma unroll
                for (int g = 0; g < NR; ++g) {
                  mq[g] = 0;
#pragma unroll
                  for (int q = 0; q < 8; ++q) rv[g][q] = (T)0;
                }
              }
              // first lane of this lane's segment: the highest head at or below it
              const unsigned long long heads = __ballot(lw & 1);
              const unsigned long long upto = wl == 63 ? heads : heads & ((2ull << wl) - 1);
              const int s0 = 63 - __clzll(upto);
#pragma unroll
              for (int spin = 0; spin < 2; ++spin) {
                // acc[2r], acc[2r+1]: r = SO m (0..2), NL m (3..5)
                T acc[12];
#pragma unroll
                for (int q = 0; q < 12; ++q) acc[q] = (T)0;
#pragma unroll
                for (int g = 0; g < NR; ++g) {
                  const int pos = (spin * HL + (mq[g] >> 16)) * LS + (mq[g] & 0xffff);
                  const C v = buf[pos];
                  const T wso = rv[g][0], wnl = rv[g][1];
#pragma unroll
                  for (int m = 0; m < 3; ++m) {
                    const T yx = rv[g][2 + 2 * m], yy = rv[g][3 + 2 * m];
                    const T ux = yx * v.x + yy * v.y, uy = yx * v.y - yy * v.x;  // conj(Y_m) psi
                    acc[2 * m] += wso * ux;
                    acc[2 * m + 1] += wso * uy;
                    acc[6 + 2 * m] += wnl * ux;
                    acc[7 + 2 * m] += wnl * uy;
                  }
                }
                for (int o = 1; o < span; o <<= 1) {
#pragma unroll
                  for (int q = 0; q < 12; ++q) {
                    const T f = __shfl_up(acc[q], o);
                    if (wl - o >= s0) acc[q] += f;
                  }
                }
                if (lw >> 1) {
                  C *pp = part + (long)((lw >> 1) - 1) * 12 + spin * 6;
#pragma unroll
                  for (int r = 0; r < 6; ++r) pp[r] = C{acc[2 * r], acc[2 * r + 1]};
                }
              }
            }
          }
        }
      }
      if (A.fwd_dst) {
        // the next Newton step's forward x transform of these rows
        __syncthreads();
        fx_fft<T, N, R1, R2, L, 1, LS, NT, false>(tw_off);
        C *fd = A.fwd_dst + base;
#pragma unroll 4
        for (int u = 0; u < E; ++u) {
          const unsigned e = tid + (unsigned)(u * NT);
          const int l = (int)(e / N), p = (int)(e % N);
          if (l < nl) fd[xoff(e)] = buf[l * LS + p];
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const unsigned e = tid + (unsigned)(u * NT);
        if (line(e) < nl) dst[doff(e)] = buf[lpos(e)];
      }
    }
  }
}

// G: expansion records a thread loads at a time in the fused combine (3 for
// FP32 at 256: the most that fits the 128 VGPRs of four waves per SIMD)
#define FBSE_AXIS_FX(NAME, T, N, R1, R2, L, NT, WAVES, G)                                                   \
  template <int KIND, bool XAX>                                                                          \
  __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES))) void NAME(AxisArgs<T> A) {  \
    axis_fx_body<T, KIND, N, R1, R2, L, NT, XAX, G>(A);                                                  \
  }
FBSE_AXIS_FX(axis_fx_f32_n256_l16, float, 256, 16, 16, 16, 256, 4, 3)
FBSE_AXIS_FX(axis_fx_f32_n256_l32, float, 256, 16, 16, 32, 512, 4, 3)
FBSE_AXIS_FX(axis_fx_f32_n160_l16, float, 160, 16, 10, 16, 256, 4, 4)
FBSE_AXIS_FX(axis_fx_f32_n160_l32, float, 160, 16, 10, 32, 512, 4, 4)
// 64 lines: the y and z tiles read 512-byte segments per axis point (FP32
// only; one work-group per CU)
FBSE_AXIS_FX(axis_fx_f32_n256_l64, float, 256, 16, 16, 64, 1024, 4, 3)
FBSE_AXIS_FX(axis_fx_f32_n160_l64, float, 160, 16, 10, 64, 1024, 4, 4)
FBSE_AXIS_FX(axis_fx_f64_n256_l16, double, 256, 16, 16, 16, 256, 2, 4)
FBSE_AXIS_FX(axis_fx_f64_n256_l32, double, 256, 16, 16, 32, 512, 2, 4)
FBSE_AXIS_FX(axis_fx_f64_n160_l16, double, 160, 16, 10, 16, 256, 2, 4)
FBSE_AXIS_FX(axis_fx_f64_n160_l32, double, 160, 16, 10, 32, 512, 2, 4)
#undef FBSE_AXIS_FX
// spin-paired combine tiles (16 lines: 8 rows of each spin)
#define FBSE_AXIS_FX_SP(NAME, T, N, R1, R2, NT, WAVES, G)                                                   \
  template <int KIND, bool XAX>                                                                          \
  __global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WAVES))) void NAME(AxisArgs<T> A) {  \
    axis_fx_body<T, KIND, N, R1, R2, 16, NT, XAX, G, true>(A);                                           \
  }
FBSE_AXIS_FX_SP(axis_fx_f32_n256_p16, float, 256, 16, 16, 256, 4, 3)
FBSE_AXIS_FX_SP(axis_fx_f32_n160_p16, float, 160, 16, 10, 256, 4, 4)
FBSE_AXIS_FX_SP(axis_fx_f64_n256_p16, double, 256, 16, 16, 256, 2, 4)
FBSE_AXIS_FX_SP(axis_fx_f64_n160_p16, double, 160, 16, 10, 256, 2, 4)
#undef FBSE_AXIS_FX_SP

bool axis_fx_length(int n) { return n == 160 || n == 256; }

template <typename T, int KIND, bool XAX> static const void *axis_fx_fn(int n, int lines, bool sp) {
  if constexpr (KIND == kPassBwdCombine && XAX) {
    if (sp) {
      if constexpr (sizeof(T) == 4)
        return n == 256 ? (const void *)axis_fx_f32_n256_p16<KIND, XAX> : (const void *)axis_fx_f32_n160_p16<KIND, XAX>;
      else
        return n == 256 ? (const void *)axis_fx_f64_n256_p16<KIND, XAX> : (const void *)axis_fx_f64_n160_p16<KIND, XAX>;
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (lines == 64)
      return n == 256 ? (const void *)axis_fx_f32_n256_l64<KIND, XAX> : (const void *)axis_fx_f32_n160_l64<KIND, XAX>;
    if (n == 256)
      return lines == 16 ? (const void *)axis_fx_f32_n256_l16<KIND, XAX> : (const void *)axis_fx_f32_n256_l32<KIND, XAX>;
    return lines == 16 ? (const void *)axis_fx_f32_n160_l16<KIND, XAX> : (const void *)axis_fx_f32_n160_l32<KIND, XAX>;
  } else {
    if (n == 256)
      return lines == 16 ? (const void *)axis_fx_f64_n256_l16<KIND, XAX> : (const void *)axis_fx_f64_n256_l32<KIND, XAX>;
    return lines == 16 ? (const void *)axis_fx_f64_n160_l16<KIND, XAX> : (const void *)axis_fx_f64_n160_l32<KIND, XAX>;
  }
}

template <typename T, int KIND, bool XAX>
static int axis_fx_launch(fbse_ctx *c, const AxisArgs<T> &A, int n, int lines) {
  // device, precision, length, lines, kind, axis -> resident work-groups
  static int cache[64][2][2][4][5][2] = {};
  const bool sp = KIND == kPassBwdCombine && XAX && A.spin_pair && lines == 16;
  const int d = c->device, pi = sizeof(T) == 8 ? 0 : 1, ni = n == 256 ? 1 : 0;
  const int li = sp ? 3 : (lines == 64 ? 2 : (lines == 32 ? 1 : 0));
  const int nt = lines * 16;  // 256, 512 or 1024 threads
  const size_t lds = (size_t)(lines * (n | 1) + n) * sizeof(cplx_t<T>) + (size_t)(n + lines) * sizeof(T);
  const void *fn = axis_fx_fn<T, KIND, XAX>(n, lines, sp);
  if (d < 0 || d >= 64) return set_err(c, FBSE_ERR_UNSUPPORTED, "device %d", d);
  int &res = cache[d][pi][ni][li][KIND][XAX ? 1 : 0];
  if (res == 0) {
    int per_cu = 0, ncu = 0;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, lds) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess) {
      (void)hipGetLastError();
      per_cu = 0;
    }
    res = per_cu > 0 && ncu > 0 ? per_cu * ncu : -1;
  }
  if (res < 0) return set_err(c, FBSE_ERR_UNSUPPORTED, "fixed axis instance %d/%d not resident", n, lines);
  if (A.ntiles <= 0) return FBSE_OK;
  const dim3 g((unsigned)std::min<long>(res, A.ntiles)), b((unsigned)nt);
  if constexpr (KIND == kPassBwdCombine && XAX) {
    if (sp) {
      if constexpr (sizeof(T) == 4) {
        if (n == 256) hipLaunchKernelGGL((axis_fx_f32_n256_p16<KIND, XAX>), g, b, lds, c->stream, A);
        else hipLaunchKernelGGL((axis_fx_f32_n160_p16<KIND, XAX>), g, b, lds, c->stream, A);
      } else {
        if (n == 256) hipLaunchKernelGGL((axis_fx_f64_n256_p16<KIND, XAX>), g, b, lds, c->stream, A);
        else hipLaunchKernelGGL((axis_fx_f64_n160_p16<KIND, XAX>), g, b, lds, c->stream, A);
      }
      FBSE_HIP(c, hipGetLastError());
      return FBSE_OK;
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (n == 256 && lines == 16) hipLaunchKernelGGL((axis_fx_f32_n256_l16<KIND, XAX>), g, b, lds, c->stream, A);
    else if (n == 256 && lines == 32) hipLaunchKernelGGL((axis_fx_f32_n256_l32<KIND, XAX>), g, b, lds, c->stream, A);
    else if (n == 256) hipLaunchKernelGGL((axis_fx_f32_n256_l64<KIND, XAX>), g, b, lds, c->stream, A);
    else if (lines == 16) hipLaunchKernelGGL((axis_fx_f32_n160_l16<KIND, XAX>), g, b, lds, c->stream, A);
    else if (lines == 32) hipLaunchKernelGGL((axis_fx_f32_n160_l32<KIND, XAX>), g, b, lds, c->stream, A);
    else hipLaunchKernelGGL((axis_fx_f32_n160_l64<KIND, XAX>), g, b, lds, c->stream, A);
  } else {
    if (n == 256 && lines == 16) hipLaunchKernelGGL((axis_fx_f64_n256_l16<KIND, XAX>), g, b, lds, c->stream, A);
    else if (n == 256) hipLaunchKernelGGL((axis_fx_f64_n256_l32<KIND, XAX>), g, b, lds, c->stream, A);
    else if (lines == 16) hipLaunchKernelGGL((axis_fx_f64_n160_l16<KIND, XAX>), g, b, lds, c->stream, A);
    else hipLaunchKernelGGL((axis_fx_f64_n160_l32<KIND, XAX>), g, b, lds, c->stream, A);
  }
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}

template <typename T> int axis_fx_pass(fbse_ctx *c, int kind, const AxisArgs<T> &A, int lines) {
  const bool kin = kind == kPassKin || kind == kPassPrec;
  const bool lines_ok = lines == 16 || lines == 32 || (lines == 64 && sizeof(T) == 4);
  if (!axis_fx_length(A.n) || !lines_ok || A.lines != lines || (A.ksqr && kin) ||
      (kin && A.xaxis) || (kind == kPassBwdCombine && !A.xaxis))
    return set_err(c, FBSE_ERR_UNSUPPORTED, "no fixed axis instance for length %d, %d lines", A.n, lines);
  // instances: forward and backward on either layout, k^2 passes on z (strided),
  // the combine on x (rows)
  switch (kind) {
    case kPassFwd:
      return A.xaxis ? axis_fx_launch<T, kPassFwd, true>(c, A, A.n, lines)
                     : axis_fx_launch<T, kPassFwd, false>(c, A, A.n, lines);
    case kPassBwd:
      return A.xaxis ? axis_fx_launch<T, kPassBwd, true>(c, A, A.n, lines)
                     : axis_fx_launch<T, kPassBwd, false>(c, A, A.n, lines);
    case kPassKin: return axis_fx_launch<T, kPassKin, false>(c, A, A.n, lines);
    case kPassPrec: return axis_fx_launch<T, kPassPrec, false>(c, A, A.n, lines);
    default: return axis_fx_launch<T, kPassBwdCombine, true>(c, A, A.n, lines);
  }
}
template int axis_fx_pass<double>(fbse_ctx *, int, const AxisArgs<double> &, int);
template int axis_fx_pass<float>(fbse_ctx *, int, const AxisArgs<float> &, int);

constexpr int kFxMaxN = 120;  // largest fixed-instance size (sizes are even)

// the fixed instance for a grid (0 when there is none)
template <typename T> static bool fx_has(int n) {
#define FBSE_FX_HAS(N, R1, R2, NT, W) \
  if (n == N) return true;
  if constexpr (sizeof(T) == 4) {
    FBSE_FX_F32(FBSE_FX_HAS)
  } else {
    FBSE_FX_F64(FBSE_FX_HAS)
  }
#undef FBSE_FX_HAS
  return false;
}
template <typename T> static int fixed_n(const fbse_ctx *c) {
  if (c->nx != c->ny || c->nx != c->nz || !c->ksqr_sep) return 0;
  return fx_has<T>((int)c->nx) ? (int)c->nx : 0;
}
template <typename T, int KIND> static const void *fx_fn(int n) {
#define FBSE_FX_FN32(N, R1, R2, NT, W) \
  if (n == N) return (const void *)fft_fx_f32_n##N<KIND>;
#define FBSE_FX_FN64(N, R1, R2, NT, W) \
  if (n == N) return (const void *)fft_fx_f64_n##N<KIND>;
  if constexpr (sizeof(T) == 4) {
    FBSE_FX_F32(FBSE_FX_FN32)
  } else {
    FBSE_FX_F64(FBSE_FX_FN64)
  }
#undef FBSE_FX_FN32
#undef FBSE_FX_FN64
  return nullptr;
}
template <typename T> static int fx_threads(int n) {
#define FBSE_FX_NT(N, R1, R2, NT, W) \
  if (n == N) return NT;
  if constexpr (sizeof(T) == 4) {
    FBSE_FX_F32(FBSE_FX_NT)
  } else {
    FBSE_FX_F64(FBSE_FX_NT)
  }
#undef FBSE_FX_NT
  return 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool radices_ok(const AxisPlan &pl) {
  for (int i = 0; i < pl.nf; ++i) {
    const int r = pl.f[i];
    if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return false;
  }
  return pl.nf > 0 && pl.nf <= 16;
}

// Radix schedule of the plane/slab passes for a line of length n (a product
// of 2, 3 and 5) in an instance staging e values per thread: the fewest
// passes over the radices below that fit e (the composite ones run as one
// register butterfly), ties to the smaller largest radix,
// larger radices first.  Packed as 4 bits (radix - 1) per pass into *f;
// returns the number of passes.
static int tile_schedule(int n, int e, unsigned long long *f) {
  static const int rad[] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
  std::vector<int> passes(n + 1, 1 << 20), maxr(n + 1, 0), pick(n + 1, 0);
  passes[1] = 0;
  for (int m = 2; m <= n; ++m) {
    if (n % m) continue;
    for (int r : rad) {
      if (m % r || passes[m / r] >= (1 << 20) || !tile_radix_fits(e, r)) continue;
      const int pp = passes[m / r] + 1, mr = std::max(maxr[m / r], r);
      if (pp < passes[m] || (pp == passes[m] && mr < maxr[m])) {
        passes[m] = pp;
        maxr[m] = mr;
        pick[m] = r;
      }
    }
  }
  std::vector<int> fs;
  for (int m = n; m > 1 && pick[m]; m /= pick[m]) fs.push_back(pick[m]);
  std::sort(fs.begin(), fs.end(), [](int a, int b) { return a > b; });
  *f = 0;
  for (size_t i = 0; i < fs.size() && i < 16; ++i) *f |= (unsigned long long)(fs[i] - 1) << (4 * i);
  return (int)fs.size();
}

// LDS row pitch: odd, so that row transforms (threads on consecutive rows)
// hit distinct banks.
static int tile_pitch(int nc) { return nc | 1; }

// tile, twiddles of both axes, then the real k^2 tables of x and z
template <typename T> static size_t tile_lds_bytes(int nr, int nc) {
  return ((size_t)nr * tile_pitch(nc) + nc + nr) * sizeof(cplx_t<T>) + (size_t)(nc + nr) * sizeof(T);
}

static int gcd_int(int a, int b) { return b ? gcd_int(b, a % b) : a; }

// Work-group size: a multiple of both line counts (nr for row transforms, nc
// for column transforms and the load/store mapping) with at least
// ceil(nr*nc / e) threads; 0 when none fits within max_nt.
static int tile_threads(int nr, int nc, int e, int max_nt) {
  const long lcm = (long)nr / gcd_int(nr, nc) * nc;
  const long need = ((long)nr * nc + e - 1) / e;
  long nt = (need + lcm - 1) / lcm * lcm;
  if (nt < 64) nt = (64 + lcm - 1) / lcm * lcm;
  return nt <= max_nt ? (int)nt : 0;
}

// first instance (in list order) that fits both tile shapes, -1 if none
template <typename T> static int first_tile_variant(const fbse_ctx *c) {
  int nv;
  const TileVariant *tv = tile_variants<T>(&nv);
  for (int v = 0; v < nv; ++v)
    if (tile_threads((int)c->ny, (int)c->nx, tv[v].e, tv[v].max_nt) &&
        tile_threads((int)c->nz, (int)c->nx, tv[v].e, tv[v].max_nt))
      return v;
  return -1;
}

template <typename T> bool tiles_supported(const fbse_ctx *c) {
  if (!radices_ok(c->plan[0]) || !radices_ok(c->plan[1]) || !radices_ok(c->plan[2])) return false;
  const int nx = (int)c->nx, ny = (int)c->ny, nz = (int)c->nz;
  const size_t cap = 160 * 1024;
  if (tile_lds_bytes<T>(ny, nx) > cap || tile_lds_bytes<T>(nz, nx) > cap) return false;
  if (first_tile_variant<T>(c) < 0) return false;
  return true;
}
template bool tiles_supported<double>(const fbse_ctx *);
template bool tiles_supported<float>(const fbse_ctx *);

template <typename T> bool tiles_fixed(const fbse_ctx *c) {
  if (!fixed_n<T>(c)) return false;
  int nv;
  tile_variants<T>(&nv);
  if (c->force_variant == 1 && c->force_lines > 0) return c->force_lines == nv + 1;
  return c->tile_variant[sizeof(T) == 8 ? 0 : 1] == nv;
}
template bool tiles_fixed<double>(const fbse_ctx *);
template bool tiles_fixed<float>(const fbse_ctx *);

template <typename T> static const cplx_t<T> *twiddles(const fbse_ctx *c, int axis);
template <> const double2 *twiddles<double>(const fbse_ctx *c, int axis) { return c->d_tw[axis]; }
template <> const float2 *twiddles<float>(const fbse_ctx *c, int axis) { return c->d_tw_f[axis]; }

// Dynamic LDS above 64 KiB has to be enabled per kernel and device.
template <typename T, int KIND> static void set_lds_limit(int device, int v) {
  static bool done[64][3] = {};
  if (device < 0 || device >= 64 || !done[device][v]) {
    (void)hipFuncSetAttribute(tile_fn<T, KIND>(v), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (device >= 0 && device < 64) done[device][v] = true;
  }
}

template <typename T, int KIND>
static void fx_launch(int n, dim3 g, dim3 b, size_t lds, hipStream_t st, const TileArgs<T> &A) {
#define FBSE_FX_L32(N, R1, R2, NT, W) \
  if (n == N) hipLaunchKernelGGL((fft_fx_f32_n##N<KIND>), g, b, lds, st, A);
#define FBSE_FX_L64(N, R1, R2, NT, W) \
  if (n == N) hipLaunchKernelGGL((fft_fx_f64_n##N<KIND>), g, b, lds, st, A);
  if constexpr (sizeof(T) == 4) {
    FBSE_FX_F32(FBSE_FX_L32)
  } else {
    FBSE_FX_F64(FBSE_FX_L64)
  }
#undef FBSE_FX_L32
#undef FBSE_FX_L64
}

// Work-groups of a persistent fixed-instance launch: the number resident at
// once (occupancy x CUs, cached per device, precision, size and pass), at
// most one per tile; 0 when the instance cannot run.
template <typename T, int KIND> static int fx_grid(fbse_ctx *c, int n, long ntiles) {
  static int cache[64][2][kFxMaxN / 2 + 1][5] = {};  // sizes are even
  const int d = c->device, pi = sizeof(T) == 8 ? 0 : 1, ni = n / 2;
  if (d < 0 || d >= 64 || n % 2 || n > kFxMaxN) return 0;
  int &res = cache[d][pi][ni][KIND];
  if (res == 0) {
    const void *fn = fx_fn<T, KIND>(n);
    if (!fn) return 0;
    const size_t lds = tile_lds_bytes<T>(n, n);
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, fx_threads<T>(n), lds) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    res = per_cu > 0 && ncu > 0 ? per_cu * ncu : -1;
  }
  if (res < 0) return 0;
  return (int)std::min<long>(res, ntiles);
}

template <typename T>
static int tile_pass_v(fbse_ctx *c, int v, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
                       const cplx_t<T> *psi_in, T a1, T a0, const cplx_t<T> *pjd = nullptr,
                       cplx_t<T> *fwd_dst = nullptr, int zsplit = 1, TileArgs<T> *fill = nullptr) {
  TileArgs<T> A{};
  A.src = src;
  A.dst = dst;
  const bool xy = kind == kTileFwdXY || kind == kTileBwdXY || kind == kTileBwdXYCombine;
  // z passes transform columns only: a slab may be cut into zsplit x ranges
  A.col_split = xy ? 1 : zsplit;
  if (A.col_split < 1 || c->nx % A.col_split)
    return set_err(c, FBSE_ERR_ARG, "slab split %d of %ld columns", A.col_split, c->nx);
  A.nc = (int)(c->nx / A.col_split);
  A.ls = tile_pitch(A.nc);
  A.item_stride = c->ngrid;
  const AxisPlan &pc = c->plan[0], &pr = c->plan[xy ? 1 : 2];
  A.nr = pr.n;
  A.twc = twiddles<T>(c, 0);
  A.twr = twiddles<T>(c, xy ? 1 : 2);
  if (xy) {
    A.row_stride = c->nx;
    A.tile_stride = c->nx * c->ny;
    A.tiles_per_item = (int)c->nz;
  } else {
    A.row_stride = c->nx * c->ny;
    A.tile_stride = c->nx;
    A.tiles_per_item = (int)c->ny * A.col_split;
  }
  A.nxny = c->nx * c->ny;
  const bool f32 = sizeof(T) == 4;
  if (c->ksqr_sep) {
    A.kx = f32 ? (const T *)c->d_kx_f : (const T *)c->d_kx;
    A.ky = f32 ? (const T *)c->d_ky_f : (const T *)c->d_ky;
    A.kz = f32 ? (const T *)c->d_kz_f : (const T *)c->d_kz;
    A.ksqr = nullptr;
  } else {
    A.ksqr = f32 ? (const T *)c->d_ksqr_f : (const T *)c->d_ksqr;
  }
  A.kcap = (T)c->kcap;
  A.pa = a1;  // the preconditioner pass takes (pa, pb) in the (a1, a0) slots
  A.pb = a0;
  A.psi_in = psi_in;
  A.vloc = f32 ? (const T *)c->d_vloc_f : (const T *)c->d_vloc;
  A.nspin = (int)c->nspin;
  A.a1 = a1;
  A.a0 = a0;
  if (kind == kTileBwdXYCombine && c->dyn_a0_tab) {
    A.a0_tab = c->dyn_a0_tab;
    A.step_base = c->dyn_step_base;
    A.step_off = c->dyn_step_off;
  }
  if (kind == kTileBwdXYCombine && pjd) {
    A.rec_off = c->d_rec_off;
    A.rec_pa = c->d_rec_pa;
    if constexpr (sizeof(T) == 8) A.rec_val = c->d_rec_val;
    else A.rec_val = c->d_rec_val_f;
    A.pt_ranks = c->pt_ranks;
    A.pjd = pjd;
    A.pj_natoms = c->pj.n_atoms;
  }
  if (kind == kTileBwdXYCombine) A.fwd_dst = fwd_dst;
  const long ntiles = nbatch * A.tiles_per_item;
  A.nbatch = nbatch;
  if (ntiles > 0x7fffffffL) return set_err(c, FBSE_ERR_UNSUPPORTED, "%ld FFT tiles in one launch", ntiles);
  int nv;
  const TileVariant *tv = tile_variants<T>(&nv);
  if (v == nv) {  // the fixed-schedule persistent instance
    const int n = fixed_n<T>(c);
    if (!n || A.col_split != 1 || kind == kTileBwdXY || kind == kTilePrecZ)
      return set_err(c, FBSE_ERR_UNSUPPORTED, "no fixed FFT instance for this pass");
    if (fill) {  // arguments for a grouped launch
      *fill = A;
      return FBSE_OK;
    }
    if (ntiles <= 0) return FBSE_OK;
    int g;
    switch (kind) {
      case kTileFwdXY: g = fx_grid<T, kTileFwdXY>(c, n, ntiles); break;
      case kTileKinZ: g = fx_grid<T, kTileKinZ>(c, n, ntiles); break;
      default: g = fx_grid<T, kTileBwdXYCombine>(c, n, ntiles); break;
    }
    if (g <= 0) return set_err(c, FBSE_ERR_UNSUPPORTED, "fixed FFT instance not resident");
    const size_t lds = tile_lds_bytes<T>(n, n);
    hipEvent_t ev = nullptr;
    const int pk = kind == kTileBwdXYCombine ? 1 : 0;
    prof_begin(c, pk, &ev);
    const dim3 gd((unsigned)g), bd((unsigned)fx_threads<T>(n));
    switch (kind) {
      case kTileFwdXY: fx_launch<T, kTileFwdXY>(n, gd, bd, lds, c->stream, A); break;
      case kTileKinZ: fx_launch<T, kTileKinZ>(n, gd, bd, lds, c->stream, A); break;
      default: fx_launch<T, kTileBwdXYCombine>(n, gd, bd, lds, c->stream, A); break;
    }
    FBSE_HIP(c, hipGetLastError());
    prof_end(c, pk, ev);
    return FBSE_OK;
  }
  A.nfc = tile_schedule(pc.n, tv[v].e, &A.fc);
  A.nfr = tile_schedule(pr.n, tv[v].e, &A.fr);
  const int nt = tile_threads(A.nr, A.nc, tv[v].e, tv[v].max_nt);
  if (nt == 0) return set_err(c, FBSE_ERR_UNSUPPORTED, "tile %d x %d does not fit FFT instance %d", A.nr, A.nc, v);
  const size_t lds = tile_lds_bytes<T>(A.nr, A.nc);
  hipEvent_t ev = nullptr;
  const int pk = kind == kTileBwdXYCombine ? 1 : 0;
  prof_begin(c, pk, &ev);
  const dim3 g((unsigned)ntiles), b((unsigned)nt);
  switch (kind) {
    case kTileFwdXY:
      set_lds_limit<T, kTileFwdXY>(c->device, v);
      tile_launch<T, kTileFwdXY>(v, g, b, lds, c->stream, A);
      break;
    case kTileBwdXY:
      set_lds_limit<T, kTileBwdXY>(c->device, v);
      tile_launch<T, kTileBwdXY>(v, g, b, lds, c->stream, A);
      break;
    case kTileBwdXYCombine:
      set_lds_limit<T, kTileBwdXYCombine>(c->device, v);
      tile_launch<T, kTileBwdXYCombine>(v, g, b, lds, c->stream, A);
      break;
    case kTileKinZ:
      set_lds_limit<T, kTileKinZ>(c->device, v);
      tile_launch<T, kTileKinZ>(v, g, b, lds, c->stream, A);
      break;
    default:
      set_lds_limit<T, kTilePrecZ>(c->device, v);
      tile_launch<T, kTilePrecZ>(v, g, b, lds, c->stream, A);
      break;
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, pk, ev);
  return FBSE_OK;
}

// Columns per z-slab tile: a slab cut into S x ranges needs 1/S of the LDS,
// so more work-groups share a CU and hide each other's loads.  Forced by
// fbse_set_slab_split, else timed once per precision on the first batch of at
// least 1M grid points (z pass into a scratch buffer, best of 3) over the
// splits that keep at least 16 columns and fit the instance; 1 before that.
template <typename T> static bool slab_split_ok(const fbse_ctx *c, int v, int s) {
  if (s < 1 || c->nx % s || c->nx / s < 16) return s == 1;
  int nv;
  const TileVariant *tv = tile_variants<T>(&nv);
  return v >= 0 && v < nv && tile_threads((int)c->nz, (int)(c->nx / s), tv[v].e, tv[v].max_nt) > 0;
}
template <typename T> static int slab_split(fbse_ctx *c, int v, const cplx_t<T> *src, long nbatch) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  if (c->force_zsplit > 0) return slab_split_ok<T>(c, v, c->force_zsplit) ? c->force_zsplit : 1;
  if (c->zsplit[pi] > 0) return c->zsplit[pi];
  if (!(c->autotune && nbatch * c->ngrid >= (1L << 20))) return 1;
  int best_s = 1;
  cplx_t<T> *scratch = nullptr;
  const size_t bytes = (size_t)nbatch * c->ngrid * sizeof(cplx_t<T>);
  if (hipMalloc(&scratch, bytes) == hipSuccess) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int saved_profile = c->profile;
    c->profile = 0;
    double best = 1e30;
    for (int s = 1; s <= 4; ++s) {
      if (!slab_split_ok<T>(c, v, s)) continue;
      double tmin = 1e30;
      for (int rep = 0; rep < 3; ++rep) {
        (void)hipEventRecord(e0, c->stream);
        int rc = tile_pass_v<T>(c, v, kTileKinZ, src, scratch, nbatch, nullptr, 0, 0, nullptr, nullptr, s);
        (void)hipEventRecord(e1, c->stream);
        if (rc) {
          c->err.clear();
          tmin = 1e30;
          break;
        }
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        tmin = std::min(tmin, (double)ms);
      }
      if (tmin < best) {
        best = tmin;
        best_s = s;
      }
    }
    c->profile = saved_profile;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(scratch);
  } else {
    (void)hipGetLastError();
  }
  c->zsplit[pi] = best_s;
  return best_s;
}

// Instance choice: the autotuner times every instance that fits on the
// first batch of at least 1M grid points (forward xy pass into a scratch
// buffer, best of 3) and keeps the fastest per precision; smaller batches
// before that use the first instance that fits.  Instance index nv (one past
// the generic ones) is the fixed-schedule persistent instance of the grid,
// when it has one; it takes whole z slabs and has no preconditioner passes
// (those run on the first generic instance that fits).
template <typename T>
static int tile_pass(fbse_ctx *c, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
                     const cplx_t<T> *psi_in, T a1, T a0, const cplx_t<T> *pjd = nullptr,
                     cplx_t<T> *fwd_dst = nullptr) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  int nv;
  const TileVariant *tv = tile_variants<T>(&nv);
  const bool has_fx = fixed_n<T>(c) > 0;
  const bool fx_kind = kind == kTileFwdXY || kind == kTileKinZ || kind == kTileBwdXYCombine;
  const bool forced = c->force_variant == 1 && c->force_lines > 0;
  if (forced && (c->force_lines > nv + 1 || (c->force_lines == nv + 1 && !has_fx)))
    return set_err(c, FBSE_ERR_ARG, "FFT tile instance %d of %d", c->force_lines, nv + (has_fx ? 1 : 0));
  if (!forced && c->tile_variant[pi] < 0) {
    int v0 = first_tile_variant<T>(c);
    if (!(c->autotune && nbatch * c->ngrid >= (1L << 20))) {
      if (kind == kTileKinZ || kind == kTilePrecZ)
        return tile_pass_v<T>(c, v0, kind, src, dst, nbatch, psi_in, a1, a0, pjd, fwd_dst,
                              slab_split<T>(c, v0, src, nbatch));
      return tile_pass_v<T>(c, v0, kind, src, dst, nbatch, psi_in, a1, a0, pjd, fwd_dst);
    }
    cplx_t<T> *scratch = nullptr;
    const size_t bytes = (size_t)nbatch * c->ngrid * sizeof(cplx_t<T>);
    int best_v = v0;
    if (hipMalloc(&scratch, bytes) == hipSuccess) {
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      const int saved_profile = c->profile;
      c->profile = 0;
      double best = 1e30;
      for (int v = 0; v <= nv; ++v) {
        if (v == nv ? !has_fx
                    : (!tile_threads((int)c->ny, (int)c->nx, tv[v].e, tv[v].max_nt) ||
                       !tile_threads((int)c->nz, (int)c->nx, tv[v].e, tv[v].max_nt)))
          continue;
        double tmin = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
          (void)hipEventRecord(e0, c->stream);
          int rc = tile_pass_v<T>(c, v, kTileFwdXY, src, scratch, nbatch, nullptr, 0, 0);
          (void)hipEventRecord(e1, c->stream);
          if (rc) {
            c->err.clear();
            tmin = 1e30;
            break;
          }
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          tmin = std::min(tmin, (double)ms);
        }
        if (tmin < best) {
          best = tmin;
          best_v = v;
        }
      }
      c->profile = saved_profile;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(scratch);
    } else {
      (void)hipGetLastError();
    }
    c->tile_variant[pi] = best_v;
  }
  int v = forced ? c->force_lines - 1 : c->tile_variant[pi];
  if (v == nv && !fx_kind) v = first_tile_variant<T>(c);
  if (v < 0 || v > nv) return set_err(c, FBSE_ERR_ARG, "FFT tile instance %d of %d", v + 1, nv);
  if (kind == kTileKinZ || kind == kTilePrecZ) {
    const int zs = v == nv ? 1 : slab_split<T>(c, v, src, nbatch);
    return tile_pass_v<T>(c, v, kind, src, dst, nbatch, psi_in, a1, a0, pjd, fwd_dst, zs);
  }
  return tile_pass_v<T>(c, v, kind, src, dst, nbatch, psi_in, a1, a0, pjd, fwd_dst);
}

// Work-groups per context of a grouped fixed-instance launch: the resident
// work-groups of the instance shared out over the contexts.
template <int KIND> static int fx_group_grid(fbse_ctx *c, int n, int n_ctx) {
  static int cache[64][kFxMaxN / 2 + 1][5] = {};
  const int d = c->device, ni = n / 2;
  if (d < 0 || d >= 64 || n % 2 || n > kFxMaxN) return 0;
  int &res = cache[d][ni][KIND];
  if (res == 0) {
    const void *fn = nullptr;
#define FBSE_FX_GFN(N, R1, R2, NT, W) \
  if (n == N) fn = (const void *)fft_fx_f32_n##N##_grp<KIND>;
    FBSE_FX_F32(FBSE_FX_GFN)
#undef FBSE_FX_GFN
    int per_cu = 0, ncu = 0;
    if (!fn || hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, fx_threads<float>(n), tile_lds_bytes<float>(n, n)) !=
            hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess) {
      (void)hipGetLastError();
      per_cu = 0;
    }
    res = per_cu > 0 && ncu > 0 ? per_cu * ncu : -1;
  }
  return res < 0 ? 0 : std::max(1, res / n_ctx);
}

bool tile_groupable(const fbse_ctx *c) {
  int nv;
  tile_variants<float>(&nv);
  const bool forced = c->force_variant == 1 && c->force_lines > 0;
  return c->force_variant != 0 && tiles_supported<float>(c) && fixed_n<float>(c) > 0 &&
         (!forced || c->force_lines == nv + 1);
}

int tile_group_pass(int n_ctx, fbse_ctx *const *cs, int kind, const float2 *const *src, float2 *const *dst,
                    const long *nbatch, const float2 *const *psi_in, const float *a1, const float *a0,
                    const float2 *const *pjd, float2 *const *fwd_dst) {
  fbse_ctx *c0 = cs[0];
  if (n_ctx < 1 || n_ctx > kTileGroupMax) return set_err(c0, FBSE_ERR_ARG, "%d contexts in one grouped pass", n_ctx);
  if (!(kind == kTileFwdXY || kind == kTileKinZ || kind == kTileBwdXYCombine))
    return set_err(c0, FBSE_ERR_UNSUPPORTED, "no grouped instance of pass %d", kind);
  const int n = fixed_n<float>(c0);
  int nv;
  tile_variants<float>(&nv);
  TileGroup<float> G{};
  long most = 0;
  int rc;
  for (int i = 0; i < n_ctx; ++i) {
    if (fixed_n<float>(cs[i]) != n || n == 0)
      return set_err(c0, FBSE_ERR_UNSUPPORTED, "grouped pass: context %d has no fixed instance of size %d", i, n);
    if ((rc = tile_pass_v<float>(cs[i], nv, kind, src[i], dst[i], nbatch[i], psi_in ? psi_in[i] : nullptr,
                                 a1 ? a1[i] : 0.f, a0 ? a0[i] : 0.f, pjd ? pjd[i] : nullptr,
                                 fwd_dst ? fwd_dst[i] : nullptr, 1, &G.a[i]))) {
      if (cs[i] != c0) set_err(c0, rc, "grouped pass, context %d: %s", i, cs[i]->err.c_str());
      return rc;
    }
    most = std::max(most, nbatch[i] * (long)G.a[i].tiles_per_item);
  }
  if (most <= 0) return FBSE_OK;
  int g;
  switch (kind) {
    case kTileFwdXY: g = fx_group_grid<kTileFwdXY>(c0, n, n_ctx); break;
    case kTileKinZ: g = fx_group_grid<kTileKinZ>(c0, n, n_ctx); break;
    default: g = fx_group_grid<kTileBwdXYCombine>(c0, n, n_ctx); break;
  }
  if (g <= 0) return set_err(c0, FBSE_ERR_UNSUPPORTED, "grouped FFT instance %d not resident", n);
  const dim3 gd((unsigned)std::min<long>(g, most), (unsigned)n_ctx), bd((unsigned)fx_threads<float>(n));
  const size_t lds = tile_lds_bytes<float>(n, n);
  hipEvent_t ev = nullptr;
  const int pk = kind == kTileBwdXYCombine ? 1 : 0;
  prof_begin(c0, pk, &ev);
#define FBSE_FX_GL(N, R1, R2, NT, W)                                                                   \
  if (n == N) {                                                                                       \
    switch (kind) {                                                                                   \
      case kTileFwdXY: hipLaunchKernelGGL((fft_fx_f32_n##N##_grp<kTileFwdXY>), gd, bd, lds, c0->stream, G); break; \
      case kTileKinZ: hipLaunchKernelGGL((fft_fx_f32_n##N##_grp<kTileKinZ>), gd, bd, lds, c0->stream, G); break;   \
      default:                                                                                        \
        hipLaunchKernelGGL((fft_fx_f32_n##N##_grp<kTileBwdXYCombine>), gd, bd, lds, c0->stream, G);   \
        break;                                                                                        \
    }                                                                                                 \
  }
  FBSE_FX_F32(FBSE_FX_GL)
#undef FBSE_FX_GL
  FBSE_HIP(c0, hipGetLastError());
  prof_end(c0, pk, ev);
  return FBSE_OK;
}

// out = a1*(T + V) psi + a0*psi over nstates*nspin grids (projectors are
// added by the caller); work holds one batch of grids.
template <typename T>
int kinetic_local_tiles(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1,
                        T a0, const cplx_t<T> *pjd, bool fwd_ready, bool fwd_next, hipEvent_t join) {
  const long nb = nstates * c->nspin;
  int rc;
  if (!fwd_ready && (rc = tile_pass<T>(c, kTileFwdXY, psi, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = tile_pass<T>(c, kTileKinZ, work, work, nb, nullptr, 0, 0))) return rc;
  if (join) FBSE_HIP(c, hipStreamWaitEvent(c->stream, join, 0));
  return tile_pass<T>(c, kTileBwdXYCombine, work, out, nb, psi, a1, a0, pjd, fwd_next ? work : nullptr);
}
template int kinetic_local_tiles<double>(fbse_ctx *, const double2 *, double2 *, double2 *, long, double, double,
                                         const double2 *, bool, bool, hipEvent_t);
template int kinetic_local_tiles<float>(fbse_ctx *, const float2 *, float2 *, float2 *, long, float, float,
                                        const float2 *, bool, bool, hipEvent_t);

template <typename T>
int precondition_tiles(fbse_ctx *c, const cplx_t<T> *r, cplx_t<T> *out, cplx_t<T> *work, long nstates, T pa,
                       T pb) {
  const long nb = nstates * c->nspin;
  int rc;
  if ((rc = tile_pass<T>(c, kTileFwdXY, r, work, nb, nullptr, 0, 0))) return rc;
  if ((rc = tile_pass<T>(c, kTilePrecZ, work, work, nb, nullptr, pa, pb))) return rc;
  return tile_pass<T>(c, kTileBwdXY, work, out, nb, nullptr, 0, 0);
}
template int precondition_tiles<double>(fbse_ctx *, const double2 *, double2 *, double2 *, long, double, double);

}  // namespace fbse
