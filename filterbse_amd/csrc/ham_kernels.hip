// ham_kernels.hip -- projector (SO / NL) application, windowed Newton
// accumulation and per-state reductions for gfx950.
#include "fbse_internal.h"

#include <algorithm>
#include "fft_common.h"

namespace fbse {

template <typename C> __device__ __forceinline__ C cm(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// conj(a) * b
template <typename C> __device__ __forceinline__ C cjm(C a, C b) {
  C r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.x * b.y - a.y * b.x;
  return r;
}


// ---------------------------------------------------------------------------
// Projectors (phamiltonian.c:140-444).  For every atom and every column (one
// (state, spin) grid) the 30 projections
//   SO rows r = p*3+m    : c = dv sum_g proj_p(g) conj(Y_m(g)) psi(g)
//   NL rows r = 15+p*3+m : c = dv sum_g NL_p(g)   conj(Y_m(g)) psi(g)
// are mixed (L.S across the two spins of a state for SO, the projector sign
// for NL) and expanded back onto the sphere:
//   out(g) += sum_{p,m} Y_m(g) (proj_p(g) d_so[p][m] + NL_p(g) d_nl[p][m]).
// Atoms are processed colour by colour (one launch per colour); atoms of one
// colour share no grid point, so the expansion is a plain read-modify-write
// with a deterministic summation order (the reference scatters with omp
// atomics, phamiltonian.c:295-303,427-435).
// Work-group = one atom x kPjCols columns; thread = (column, point lane): it
// walks every kPjLanes-th point of the sphere holding all 30 projections of
// its column in registers; the lanes are then summed through LDS.
// ---------------------------------------------------------------------------
constexpr int kPjCols = kProjColChunk;
constexpr int kPjLanes = 16;  // kPjCols * kPjLanes = 256 threads

template <typename T, int MAXP, bool EXPAND>
__device__ __forceinline__ void proj_color_body(const ProjArgs<T> &P) {
  using C = cplx_t<T>;
  __shared__ C s_d[kProjRows][kPjCols];
  // expansion launches: atoms along x; projection-only launch: column chunks
  // along x, so that the work-groups one XCD runs together (the dispatcher
  // deals them round-robin over the 8 XCDs) take neighbouring atoms of the
  // same columns and share the cache lines of overlapping spheres
  const int atom = P.atoms[EXPAND ? blockIdx.x : blockIdx.y];
  const int col0 = (EXPAND ? blockIdx.y : blockIdx.x) * kPjCols;
  const int lane = threadIdx.x % kPjLanes;   // point lane (fastest: consecutive points)
  const int cc = threadIdx.x / kPjLanes;     // column within the chunk
  const int col = col0 + cc;
  const bool live = col < P.ncols;
  const int off = P.a_off[atom], cnt = P.a_cnt[atom];
  const int st = live ? col / P.nspin : 0, sp = live ? col - st * P.nspin : 0;
  const C *psi = P.psi_in + (long)st * P.nsg + (long)sp * P.ngrid;
  const int nso = P.so ? P.n_so : 0, nnl = P.nl ? P.n_nl : 0;

  // acc[p*3+m] (SO) and acc[3*MAXP + p*3+m] (NL) for p < MAXP radial projectors
  constexpr int NR = 6 * MAXP;
  C acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r].x = acc[r].y = 0;
  if (live) {
    // points per lane in flight: the index loads, then the gathers (eight
    // FP32 gathers per lane fit the registers that the occupancy leaves)
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    for (int g0 = lane; g0 < cnt; g0 += U * kPjLanes) {
      int gi[U];
      C vv[U];
#pragma unroll
      for (int q = 0; q < U; ++q) gi[q] = g0 + q * kPjLanes < cnt ? P.jxyz[off + g0 + q * kPjLanes] : -1;
#pragma unroll
      for (int q = 0; q < U; ++q)
        if (gi[q] >= 0) vv[q] = psi[gi[q]];
#pragma unroll
      for (int q = 0; q < U; ++q) {
      if (gi[q] < 0) continue;
      const int idx = off + g0 + q * kPjLanes;
      const C v = vv[q];
      C u[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const C y = P.y1[(long)m * P.npts + idx];
        u[m].x = y.x * v.x + y.y * v.y;  // conj(Y) psi
        u[m].y = y.x * v.y - y.y * v.x;
      }
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        if (p < nso) {
          const T w = P.proj[(long)p * P.npts + idx];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            acc[p * 3 + m].x += w * u[m].x;
            acc[p * 3 + m].y += w * u[m].y;
          }
        }
        if (p < nnl) {
          const T w = P.nlproj[(long)p * P.npts + idx];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            acc[3 * MAXP + p * 3 + m].x += w * u[m].x;
            acc[3 * MAXP + p * 3 + m].y += w * u[m].y;
          }
        }
      }
      }
    }
  }
  // sum the point lanes: the kPjLanes lanes of a column are consecutive lanes
  // of one wavefront, reduced with butterfly shuffles
#pragma unroll
  for (int r = 0; r < NR; ++r) {
#pragma unroll
    for (int o = kPjLanes / 2; o > 0; o >>= 1) {
      acc[r].x += __shfl_xor(acc[r].x, o, kPjLanes);
      acc[r].y += __shfl_xor(acc[r].y, o, kPjLanes);
    }
  }
  // raw projections into s_d (30 rows x columns; rows of absent projectors 0)
  for (int r = lane; r < kProjRows; r += kPjLanes) s_d[r][cc] = C{};
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 3 * MAXP; ++r)
    if (lane == r % kPjLanes) {
      s_d[r][cc] = acc[r];
      s_d[15 + r][cc] = acc[3 * MAXP + r];
    }
  __syncthreads();
  // mix: SO with L.S (needs both spins of the state), NL with the sign; x dv*a1
  C dmix[2];
  if (lane < kProjRows / 2) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int r = half * 15 + lane;
      C d;
      d.x = d.y = 0;
      if (r < 15) {
        const int p = r / 3, m = r % 3;
        if (p < nso && P.nspin == 2) {
          const int spo = (col0 + cc) % 2;
          const int cbase = cc - spo;
          const int j = spo * 3 + m;
          for (int jp = 0; jp < 6; ++jp) {
            const C ls = P.ls[jp * 6 + j];
            if (ls.x == (T)0 && ls.y == (T)0) continue;
            const C c = s_d[p * 3 + jp % 3][cbase + jp / 3];
            d.x += ls.x * c.x - ls.y * c.y;
            d.y += ls.x * c.y + ls.y * c.x;
          }
        }
      } else {
        const int p = (r - 15) / 3;
        if (p < nnl) {
          const T sg = (T)P.sign[atom * 5 + p];
          d = s_d[r][cc];
          d.x *= sg;
          d.y *= sg;
        }
      }
      d.x *= P.scale;
      d.y *= P.scale;
      dmix[half] = d;
    }
  }
  __syncthreads();
  if (lane < kProjRows / 2) {
    s_d[lane][cc] = dmix[0];
    s_d[15 + lane][cc] = dmix[1];
  }
  __syncthreads();
  if (!live) return;
  if (!EXPAND) {  // projection only: hand the mixed coefficients to the plane pass
    if (lane < 6 * MAXP) {
      const int rr = lane < 3 * MAXP ? lane : 15 + lane - 3 * MAXP;
      P.dout[((long)col * P.n_atoms + atom) * (6 * MAXP) + lane] = s_d[rr][cc];
    }
    return;
  }
  // expansion: this thread's column, every kPjLanes-th point
  C d[NR];
#pragma unroll
  for (int r = 0; r < 3 * MAXP; ++r) {
    d[r] = s_d[r][cc];
    d[3 * MAXP + r] = s_d[15 + r][cc];
  }
  C *out = P.psi_out + (long)st * P.nsg + (long)sp * P.ngrid;
  for (int g = lane; g < cnt; g += kPjLanes) {
    const int idx = off + g;
    C s3[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) s3[m].x = s3[m].y = 0;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      if (p < nso) {
        const T w = P.proj[(long)p * P.npts + idx];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          s3[m].x += w * d[p * 3 + m].x;
          s3[m].y += w * d[p * 3 + m].y;
        }
      }
      if (p < nnl) {
        const T w = P.nlproj[(long)p * P.npts + idx];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          s3[m].x += w * d[3 * MAXP + p * 3 + m].x;
          s3[m].y += w * d[3 * MAXP + p * 3 + m].y;
        }
      }
    }
    C a;
    a.x = a.y = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const C y = P.y1[(long)m * P.npts + idx];
      a.x += y.x * s3[m].x - y.y * s3[m].y;
      a.y += y.x * s3[m].y + y.y * s3[m].x;
    }
    const int j = P.jxyz[idx];
    C o = out[j];
    o.x += a.x;
    o.y += a.y;
    out[j] = o;
  }
}

template <typename T, int MAXP, bool EXPAND = true>
__global__ __launch_bounds__(256) void proj_color_kernel(ProjArgs<T> P) {
  proj_color_body<T, MAXP, EXPAND>(P);
}

// Projection-only launches of several contexts at once (the grouped Newton
// steps of the batched uneven dots): blockIdx.z selects the context, whose
// work-groups past its column chunks or atoms leave at once.
template <typename T> struct ProjGroup {
  ProjArgs<T> a[kTileGroupMax];
};
template <typename T>
__global__ __launch_bounds__(256) void proj_group_kernel(ProjGroup<T> G) {
  const ProjArgs<T> &P = G.a[blockIdx.z];
  if ((int)blockIdx.x * kPjCols >= P.ncols || (int)blockIdx.y >= P.n_atoms) return;
  proj_color_body<T, 1, false>(P);
}

template <typename T>
int proj_apply(fbse_ctx *c, const cplx_t<T> *psi_in, cplx_t<T> *psi_out, long nstates, double a1) {
  ProjArgs<T> P{};
  const ProjTables<T> *pt;
  if constexpr (sizeof(T) == 8) pt = &c->pj;
  else pt = &c->pj_f;
  P.n_atoms = pt->n_atoms;
  P.npts = pt->npts;
  P.so = pt->so;
  P.nl = pt->nl;
  P.n_so = pt->n_so;
  P.n_nl = pt->n_nl;
  P.a_off = pt->a_off;
  P.a_cnt = pt->a_cnt;
  P.jxyz = pt->jxyz;
  P.proj = pt->proj;
  P.nlproj = pt->nlproj;
  P.y1 = pt->y1;
  P.sign = pt->sign;
  P.ls = pt->ls;
  P.psi_in = psi_in;
  P.psi_out = psi_out;
  P.ngrid = c->ngrid;
  P.nsg = c->nsg;
  P.nspin = (int)c->nspin;
  P.ncols = (int)(nstates * c->nspin);
  P.scale = (T)(c->dv * a1);
  if (P.n_atoms == 0 || P.ncols == 0) return FBSE_OK;
  hipEvent_t ev = nullptr;
  prof_begin(c, 2, &ev);
  const int nchunk = (P.ncols + kPjCols - 1) / kPjCols;
  const int ncol = (int)pt->color_off.size() - 1;
  for (int k = 0; k < ncol; ++k) {
    const int na = pt->color_off[k + 1] - pt->color_off[k];
    if (na == 0) continue;
    P.atoms = pt->color_atoms + pt->color_off[k];
    const int nso = P.so ? P.n_so : 0, nnl = P.nl ? P.n_nl : 0;
    if (nso <= 1 && nnl <= 1)
      hipLaunchKernelGGL((proj_color_kernel<T, 1>), dim3(na, nchunk), dim3(256), 0, c->stream, P);
    else
      hipLaunchKernelGGL((proj_color_kernel<T, 5>), dim3(na, nchunk), dim3(256), 0, c->stream, P);
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, 2, ev);
  return FBSE_OK;
}
bool proj_fusable(const fbse_ctx *c) {
  if (!c->has_proj || !c->fuse_proj || !c->d_rec_off) return false;
  const int nso = c->pj.so ? c->pj.n_so : 0, nnl = c->pj.nl ? c->pj.n_nl : 0;
  return nso <= 1 && nnl <= 1 && c->pj.n_atoms <= 65535;  // the projection launch puts atoms along grid y
}

// arguments of a projection-only launch (mixed coefficients into d)
template <typename T>
static void proj_only_args(fbse_ctx *c, const cplx_t<T> *psi_in, long nstates, double a1, cplx_t<T> *d,
                           ProjArgs<T> &P) {
  P = ProjArgs<T>{};
  const ProjTables<T> *pt;
  if constexpr (sizeof(T) == 8) pt = &c->pj;
  else pt = &c->pj_f;
  P.n_atoms = pt->n_atoms;
  P.npts = pt->npts;
  P.so = pt->so;
  P.nl = pt->nl;
  P.n_so = pt->n_so;
  P.n_nl = pt->n_nl;
  P.a_off = pt->a_off;
  P.a_cnt = pt->a_cnt;
  P.jxyz = pt->jxyz;
  P.proj = pt->proj;
  P.nlproj = pt->nlproj;
  P.y1 = pt->y1;
  P.sign = pt->sign;
  P.ls = pt->ls;
  P.psi_in = psi_in;
  P.psi_out = nullptr;
  P.dout = d;
  P.ngrid = c->ngrid;
  P.nsg = c->nsg;
  P.nspin = (int)c->nspin;
  P.ncols = (int)(nstates * c->nspin);
  P.scale = (T)(c->dv * a1);
  P.atoms = pt->spatial_atoms;  // every atom (any order works: nothing is written to the grid)
}

template <typename T>
int proj_project(fbse_ctx *c, const cplx_t<T> *psi_in, long nstates, double a1, cplx_t<T> *d) {
  if (!proj_fusable(c)) return set_err(c, FBSE_ERR_UNSUPPORTED, "projection-only pass needs one radial projector");
  ProjArgs<T> P;
  proj_only_args<T>(c, psi_in, nstates, a1, d, P);
  if (P.n_atoms == 0 || P.ncols == 0) return FBSE_OK;
  hipEvent_t ev = nullptr;
  prof_begin(c, 2, &ev);
  const int nchunk = (P.ncols + kPjCols - 1) / kPjCols;
  if (P.n_atoms > 65535) return set_err(c, FBSE_ERR_UNSUPPORTED, "%d atoms in one projection launch", P.n_atoms);
  hipLaunchKernelGGL((proj_color_kernel<T, 1, false>), dim3(nchunk, P.n_atoms), dim3(256), 0, c->stream, P);
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, 2, ev);
  return FBSE_OK;
}

int proj_project_group(int n, fbse_ctx *const *cs, const float2 *const *psi, const long *nstates, const float *a1,
                       float2 *const *d) {
  fbse_ctx *c0 = cs[0];
  if (n < 1 || n > kTileGroupMax) return set_err(c0, FBSE_ERR_ARG, "%d contexts in one grouped projection", n);
  ProjGroup<float> G{};
  int most_chunks = 0, most_atoms = 0;
  for (int i = 0; i < n; ++i) {
    if (!proj_fusable(cs[i]))
      return set_err(c0, FBSE_ERR_UNSUPPORTED, "grouped projection: context %d needs one radial projector", i);
    proj_only_args<float>(cs[i], psi[i], nstates[i], (double)a1[i], d[i], G.a[i]);
    if (G.a[i].n_atoms > 65535) return set_err(c0, FBSE_ERR_UNSUPPORTED, "%d atoms in one projection launch", G.a[i].n_atoms);
    most_chunks = std::max(most_chunks, (G.a[i].ncols + kPjCols - 1) / kPjCols);
    most_atoms = std::max(most_atoms, G.a[i].n_atoms);
  }
  if (most_chunks == 0 || most_atoms == 0) return FBSE_OK;
  hipEvent_t ev = nullptr;
  prof_begin(c0, 2, &ev);
  hipLaunchKernelGGL(proj_group_kernel<float>, dim3(most_chunks, most_atoms, n), dim3(256), 0, c0->stream, G);
  FBSE_HIP(c0, hipGetLastError());
  prof_end(c0, 2, ev);
  return FBSE_OK;
}

// Expansion of the rank-k records of every plane into out (one batch item
// per blockIdx.z, one plane per blockIdx.y): the arithmetic of the plane
// pass's rounds; the records of one rank touch distinct points, and ranks run
// as separate launches in order, so every point sums its atoms in index order.
template <typename T>
__global__ __launch_bounds__(256) void proj_expand_rank_kernel(const int *rec_off, const int *rec_pa,
                                                               const T *rec_val, int ranks, int k,
                                                               const cplx_t<T> *d, int natoms, cplx_t<T> *out,
                                                               long item_stride, long nx, long nxny, int nitems) {
  // one thread per record of rank k in plane z, every batch item in turn:
  // the record is read once, not once per item
  const int z = blockIdx.y;
  const int e = rec_off[z * ranks + k] + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= rec_off[z * ranks + k + 1]) return;
  const int2 q = reinterpret_cast<const int2 *>(rec_pa)[e];
  T r[8];
  load_record<T>(rec_val + (long)e * 8, r);
  cplx_t<T> *o = out + z * nxny + (long)(q.x >> 16) * nx + (q.x & 0xffff);
  const cplx_t<T> *dd = d + (long)q.y * 6;
#pragma unroll 4
  for (int item = 0; item < nitems; ++item) {
    const cplx_t<T> t = record_term_r<T>(r, dd + (long)item * natoms * 6);
    cplx_t<T> v = o[item * item_stride];
    v.x += t.x;
    v.y += t.y;
    o[item * item_stride] = v;
  }
}

template <typename T>
int proj_expand(fbse_ctx *c, const cplx_t<T> *d, cplx_t<T> *out, long nstates) {
  if (!proj_fusable(c)) return set_err(c, FBSE_ERR_UNSUPPORTED, "expansion pass needs one radial projector");
  const long nb = nstates * c->nspin;
  if (nb > 65535) return set_err(c, FBSE_ERR_UNSUPPORTED, "%ld grids in one expansion launch", nb);
  const T *val;
  if constexpr (sizeof(T) == 8) val = c->d_rec_val;
  else val = c->d_rec_val_f;
  hipEvent_t ev = nullptr;
  prof_begin(c, 2, &ev);
  for (int k = 0; k < c->pt_ranks; ++k) {
    const int seg = c->rec_seg_max[k];
    if (seg == 0) continue;
    const dim3 g((unsigned)((seg + 255) / 256), (unsigned)c->nz);
    hipLaunchKernelGGL((proj_expand_rank_kernel<T>), g, dim3(256), 0, c->stream, c->d_rec_off, c->d_rec_pa, val,
                       c->pt_ranks, k, d, c->pj.n_atoms, out, c->ngrid, c->nx, c->nx * c->ny, (int)nb);
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, 2, ev);
  return FBSE_OK;
}
template int proj_expand<double>(fbse_ctx *, const double2 *, double2 *, long);
template int proj_expand<float>(fbse_ctx *, const float2 *, float2 *, long);

template int proj_project<double>(fbse_ctx *, const double2 *, long, double, double2 *);
template int proj_project<float>(fbse_ctx *, const float2 *, long, double, float2 *);

// The projections of a state batch from the per-tile sums that the
// spin-paired combine wrote (AxisArgs::pp_*): every atom's 12 raw sums (both
// spins, SO and NL rows) added over its slots in tile order, then mixed as in
// the projection pass (L.S across the two spins for SO, the projector sign
// for NL, times dv * a1) into the coefficients d (column x atom x 6).
template <typename T> struct PpReduceArgs {
  const cplx_t<T> *part;
  const int *atom_off;
  long nslots;
  int n_atoms, so, nl;
  const cplx_t<T> *ls;
  const int *sign;
  T scale;
  cplx_t<T> *d;
};
constexpr int kPpAtoms = 16;  // atoms per work-group, 12 sums each

template <typename T>
__global__ __launch_bounds__(kPpAtoms * 12) void proj_reduce_kernel(PpReduceArgs<T> R) {
  using C = cplx_t<T>;
  __shared__ C s[kPpAtoms][12];
  const int al = threadIdx.x / 12, r = threadIdx.x % 12;
  const int atom = blockIdx.x * kPpAtoms + al;
  const long st = blockIdx.y;
  C acc{(T)0, (T)0};
  if (atom < R.n_atoms) {
    const C *p = R.part + st * R.nslots * 12;
    for (int g = R.atom_off[atom]; g < R.atom_off[atom + 1]; ++g) {
      const C v = p[(long)g * 12 + r];
      acc.x += v.x;
      acc.y += v.y;
    }
  }
  s[al][r] = acc;
  __syncthreads();
  if (atom >= R.n_atoms) return;
  const int spin = r / 6, q = r % 6;
  C d{(T)0, (T)0};
  if (q < 3) {
    if (R.so) {
      const int j = spin * 3 + q;
      for (int jp = 0; jp < 6; ++jp) {
        const C l = R.ls[jp * 6 + j];
        const C c = s[al][(jp / 3) * 6 + jp % 3];
        d.x += l.x * c.x - l.y * c.y;
        d.y += l.x * c.y + l.y * c.x;
      }
    }
  } else if (R.nl) {
    const T sg = (T)R.sign[atom * 5];
    d = s[al][r];
    d.x *= sg;
    d.y *= sg;
  }
  d.x *= R.scale;
  d.y *= R.scale;
  R.d[((st * 2 + spin) * R.n_atoms + atom) * 6 + q] = d;
}

template <typename T> int proj_reduce(fbse_ctx *c, long nstates, double a1, cplx_t<T> *d) {
  const ProjTables<T> *pt;
  if constexpr (sizeof(T) == 8) pt = &c->pj;
  else pt = &c->pj_f;
  const size_t need = (size_t)nstates * c->pp_nslots * 12 * sizeof(cplx_t<T>);
  if (!c->d_pp_atom_off || c->nspin != 2 || !c->d_pp_part || c->pp_part_bytes < need)
    return set_err(c, FBSE_ERR_UNSUPPORTED, "no per-tile projection sums for %ld states", nstates);
  if (nstates > 65535) return set_err(c, FBSE_ERR_UNSUPPORTED, "%ld states in one projection reduction", nstates);
  PpReduceArgs<T> R{};
  R.part = (const cplx_t<T> *)c->d_pp_part;
  R.atom_off = c->d_pp_atom_off;
  R.nslots = c->pp_nslots;
  R.n_atoms = pt->n_atoms;
  R.so = pt->so && pt->n_so > 0;
  R.nl = pt->nl && pt->n_nl > 0;
  R.ls = pt->ls;
  R.sign = pt->sign;
  R.scale = (T)(c->dv * a1);
  R.d = d;
  if (R.n_atoms == 0 || nstates == 0) return FBSE_OK;
  hipEvent_t ev = nullptr;
  prof_begin(c, 2, &ev);
  const dim3 g((unsigned)((R.n_atoms + kPpAtoms - 1) / kPpAtoms), (unsigned)nstates);
  hipLaunchKernelGGL((proj_reduce_kernel<T>), g, dim3(kPpAtoms * 12), 0, c->stream, R);
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, 2, ev);
  return FBSE_OK;
}
template int proj_reduce<double>(fbse_ctx *, long, double, double2 *);
template int proj_reduce<float>(fbse_ctx *, long, double, float2 *);

template int proj_apply<double>(fbse_ctx *, const double2 *, double2 *, long, double);
template int proj_apply<float>(fbse_ctx *, const float2 *, float2 *, long, double);

// ---------------------------------------------------------------------------
// Layout conversions.  Real states are packed two per complex grid (state 2q
// in .x, 2q+1 in .y): T, V_loc and V_NL are real operators, so H acting on the
// packed grid acts on each real state separately (an odd count pads a zero).
// ---------------------------------------------------------------------------
__global__ void pack_real_kernel(const double *src, long src_stride, long nstates, long nsg, double2 *dst) {
  const long q = blockIdx.y;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nsg; g += (long)gridDim.x * blockDim.x) {
    double2 v;
    v.x = src[(2 * q) * src_stride + g];
    v.y = (2 * q + 1 < nstates) ? src[(2 * q + 1) * src_stride + g] : 0.0;
    dst[q * nsg + g] = v;
  }
}
__global__ void unpack_real_kernel(const double2 *src, long nstates, long nsg, double *dst, long dst_stride) {
  const long q = blockIdx.y;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nsg; g += (long)gridDim.x * blockDim.x) {
    const double2 v = src[q * nsg + g];
    dst[(2 * q) * dst_stride + g] = v.x;
    if (2 * q + 1 < nstates) dst[(2 * q + 1) * dst_stride + g] = v.y;
  }
}
static dim3 grid1d(long n, long y) {
  long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  return dim3((unsigned)b, (unsigned)y);
}
int pack_real(fbse_ctx *c, const double *src, long src_stride, long nstates, double2 *dst) {
  const long np = (nstates + 1) / 2;
  if (np == 0) return FBSE_OK;
  hipLaunchKernelGGL(pack_real_kernel, grid1d(c->nsg, np), dim3(256), 0, c->stream, src, src_stride, nstates, c->nsg,
                     dst);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}
int unpack_real(fbse_ctx *c, const double2 *src, long nstates, double *dst, long dst_stride) {
  const long np = (nstates + 1) / 2;
  if (np == 0) return FBSE_OK;
  hipLaunchKernelGGL(unpack_real_kernel, grid1d(c->nsg, np), dim3(256), 0, c->stream, src, nstates, c->nsg, dst,
                     dst_stride);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}
__global__ void to_f32_kernel(const double2 *src, float2 *dst, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}
int to_f32(fbse_ctx *c, const double2 *src, float2 *dst, long n) {
  hipLaunchKernelGGL(to_f32_kernel, grid1d(n, 1), dim3(256), 0, c->stream, src, dst, n);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}
// count blocks of n elements, block q read from src + q*src_stride, written
// contiguously (the start states of a psi_rank block into an FP32 ring slot)
__global__ void to_f32_strided_kernel(const double2 *src, long src_stride, float2 *dst, long n) {
  const long q = blockIdx.y;
  const double2 *s = src + q * src_stride;
  float2 *d = dst + q * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double2 v = s[i];
    d[i] = make_float2((float)v.x, (float)v.y);
  }
}
int to_f32_strided(fbse_ctx *c, const double2 *src, long src_stride, float2 *dst, long n, long count) {
  for (long q0 = 0; q0 < count; q0 += 65535) {
    const long nq = count - q0 < 65535 ? count - q0 : 65535;
    hipLaunchKernelGGL(to_f32_strided_kernel, grid1d(n, nq), dim3(256), 0, c->stream, src + q0 * src_stride,
                       src_stride, dst + q0 * n, n);
    FBSE_HIP(c, hipGetLastError());
  }
  return FBSE_OK;
}
// zero fill by kernel (graph-capture safe, unlike hipMemsetAsync)
__global__ void fill_zero_kernel(uint4 *dst, size_t n16, unsigned char *tail, size_t ntail) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}
int fill_zero(fbse_ctx *c, void *dst, size_t bytes) {
  const size_t n16 = bytes / 16, ntail = bytes % 16;
  hipLaunchKernelGGL(fill_zero_kernel, grid1d((long)(n16 ? n16 : 1), 1), dim3(256), 0, c->stream,
                     reinterpret_cast<uint4 *>(dst), n16, reinterpret_cast<unsigned char *>(dst) + n16 * 16, ntail);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}

// ---------------------------------------------------------------------------
// Per-state reductions (deterministic: fixed block partition, then an ordered
// second stage).  elems = doubles per state.
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 64;

template <int OP>
__global__ void state_red_kernel(const double *a, const double *b, const double *eps, long elems,
                                 double *partial) {
  const long s = blockIdx.y;
  const double *pa = a + s * elems;
  const double *pb = b ? b + s * elems : nullptr;
  double acc = 0.0;
  const double e = eps ? eps[s] : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < elems; i += (long)gridDim.x * blockDim.x) {
    if (OP == 0) acc += pa[i] * pa[i];
    else if (OP == 1) acc += pa[i] * pb[i];
    else {
      const double r = pb[i] - e * pa[i];
      acc += r * r;
    }
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[s * gridDim.x + blockIdx.x] = sh[0];
}
// second stage for many partials per state: one work-group per state, a
// fixed strided split over 256 threads, then a fixed tree
__global__ void red_final_block_kernel(const double *partial, int nb, double *out) {
  const long s = blockIdx.x;
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) t += partial[s * nb + b];
  __shared__ double sh[256];
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = sh[0];
}
__global__ void red_final_kernel(const double *partial, int nb, long nstates, double *out) {
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < nstates; s += (long)gridDim.x * blockDim.x) {
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += partial[s * nb + b];
    out[s] = t;
  }
}
template <int OP>
static int state_red(fbse_ctx *c, const double *a, const double *b, const double *eps, long nstates, long elems,
                     double *d_out) {
  if (nstates == 0) return FBSE_OK;
  // blocks per state: kRedBlocks, more for few large states (one state of a
  // 256^3 spinor grid is 67M doubles: 64 blocks left most CUs idle), at
  // least 4096 doubles per block; the partition depends only on (nstates,
  // elems), so results are reproducible
  long nblk = kRedBlocks;
  if (nstates < 32) {
    nblk = std::min<long>(2048 / nstates, elems / 4096);
    if (nblk < kRedBlocks) nblk = kRedBlocks;
  }
  int rc = ensure_red(c, (size_t)std::min<long>(nstates, 65535) * nblk * sizeof(double));
  if (rc) return rc;
  for (long s0 = 0; s0 < nstates; s0 += 65535) {
    const long ns = nstates - s0 < 65535 ? nstates - s0 : 65535;
    hipLaunchKernelGGL(state_red_kernel<OP>, dim3((unsigned)nblk, ns), dim3(256), 0, c->stream, a + s0 * elems,
                       b ? b + s0 * elems : nullptr, eps ? eps + s0 : nullptr, elems, c->d_red);
    FBSE_HIP(c, hipGetLastError());
    if (nblk > kRedBlocks)
      hipLaunchKernelGGL(red_final_block_kernel, dim3((unsigned)ns), dim3(256), 0, c->stream, c->d_red, (int)nblk,
                         d_out + s0);
    else
      hipLaunchKernelGGL(red_final_kernel, dim3((ns + 255) / 256), dim3(256), 0, c->stream, c->d_red, (int)nblk, ns,
                         d_out + s0);
    FBSE_HIP(c, hipGetLastError());
  }
  return FBSE_OK;
}
int state_norms(fbse_ctx *c, const double *psi, long nstates, long elems, double *d_out) {
  return state_red<0>(c, psi, nullptr, nullptr, nstates, elems, d_out);
}
int state_dots(fbse_ctx *c, const double *a, const double *b, long nstates, long elems, double *d_out) {
  return state_red<1>(c, a, b, nullptr, nstates, elems, d_out);
}
int state_resid(fbse_ctx *c, const double *a, const double *ha, const double *eps, long nstates, long elems,
                double *d_out) {
  return state_red<2>(c, a, ha, eps, nstates, elems, d_out);
}
__global__ void state_scale_kernel(double *psi, long elems, const double *fac) {
  const long s = blockIdx.y;
  const double f = fac[s];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < elems; i += (long)gridDim.x * blockDim.x)
    psi[s * elems + i] *= f;
}
int state_scale(fbse_ctx *c, double *psi, long nstates, long elems, const double *d_fac) {
  for (long s0 = 0; s0 < nstates; s0 += 65535) {
    const long ns = nstates - s0 < 65535 ? nstates - s0 : 65535;
    hipLaunchKernelGGL(state_scale_kernel, grid1d(elems, ns), dim3(256), 0, c->stream, psi + s0 * elems, elems,
                       d_fac + s0);
    FBSE_HIP(c, hipGetLastError());
  }
  return FBSE_OK;
}
__global__ void axpby_kernel(const double *x, double *y, long n, double a, double b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
__global__ void sub_scaled_kernel(const double *x, double *y, long elems, const double *theta) {
  const long s = blockIdx.y;
  const double t = theta[s];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < elems; i += (long)gridDim.x * blockDim.x)
    y[s * elems + i] -= t * x[s * elems + i];
}
int states_sub_scaled(fbse_ctx *c, const double *x, double *y, long nstates, long elems, const double *d_theta) {
  for (long s0 = 0; s0 < nstates; s0 += 65535) {
    const long ns = nstates - s0 < 65535 ? nstates - s0 : 65535;
    hipLaunchKernelGGL(sub_scaled_kernel, grid1d(elems, ns), dim3(256), 0, c->stream, x + s0 * elems,
                       y + s0 * elems, elems, d_theta + s0);
    FBSE_HIP(c, hipGetLastError());
  }
  return FBSE_OK;
}
int axpby_states(fbse_ctx *c, const double *x, double *y, long n, double a, double b) {
  hipLaunchKernelGGL(axpby_kernel, grid1d(n, 1), dim3(256), 0, c->stream, x, y, n, a, b);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}

// ---------------------------------------------------------------------------
// Windowed Newton accumulation: for every start b, target t and grid element,
//   acc[b,t] += sum_{j_lo <= j < j_hi} an[t,j] * psi_j
// (filter.c:284-312 term 0 and :329-359, blocked over a window of steps so the
// ms accumulators are read and written once per window instead of per step).
// ring slot of step j is j % ring_slots; slot_stride = elements per slot.
// packed = 1: real states packed in pairs, only an.re is used (an.im is
// identically 0 for the Gaussian filter, coeff.c:89-105).
// ---------------------------------------------------------------------------
constexpr int kMaxWin = kMaxWindow + 1;  // ring slots one accumulation reads
constexpr int kMaxMs = 16;

constexpr int kAccGroup = 4;  // ring slots loaded per group in accumulate_kernel

// The coefficients are wave-uniform (same step j and target t in every lane):
// read through a restrict-qualified pointer they become scalar loads and
// enter the FMAs as SGPR operands, instead of one LDS broadcast per (t, j)
// whose 16-byte-per-lane return would cost more than the FMAs it feeds.  They
// are stored step-major (anT[j*ms + t], the transpose of the reference's
// an[t*ncheby + j]): one step's targets are contiguous (128 bytes for 16 real
// coefficients, fetched by wide scalar loads), so a window of K steps keeps
// K*ms*8 bytes in the scalar cache instead of one cache line per (t, j).
// REAL_A: every an.im is 0 (the Gaussian filter, coeff.c:89-105; always for
// packed real states): real parts only, 2 FMAs per (t, j) instead of 4.
template <typename T, bool REAL_A>
__global__ __launch_bounds__(256) void accumulate_kernel(const cplx_t<T> *__restrict__ ring, int ring_slots,
                                                         long slot_stride, long j_lo, int nj,
                                                         const double *__restrict__ an, int ms, int t0, int mc,
                                                         double *__restrict__ psi_rank, long n_starts, long n_items,
                                                         int packed, long nsg, const long *__restrict__ jlo_dev,
                                                         int store) {
  const long item = blockIdx.y;  // complex: start state; packed: start pair
  if (jlo_dev) j_lo = *jlo_dev;
  const long slot0 = j_lo % ring_slots;
  // REAL_A: anT[(j_lo + j)*ms + t]; else anT[((j_lo + j)*ms + t)*2 + {0,1}]
  // this launch accumulates targets t0 .. t0 + mc - 1 (mc <= kMaxMs) of the ms
  const double *a_win = an + (REAL_A ? 1 : 2) * (j_lo * ms + t0);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nsg; g += (long)gridDim.x * blockDim.x) {
    const cplx_t<T> *src = ring + item * nsg + g;
    double2 acc[kMaxMs];
#pragma unroll
    for (int t = 0; t < kMaxMs; ++t) acc[t] = make_double2(0.0, 0.0);
    // The window's slots follow each other modulo the ring.  Loads run one
    // group of kAccGroup slots ahead of the arithmetic.
    cplx_t<T> cur[kAccGroup], nxt[kAccGroup];
    long slot = slot0;
#pragma unroll
    for (int u = 0; u < kAccGroup; ++u) {
      if (u < nj) cur[u] = src[slot * slot_stride];
      if (++slot == ring_slots) slot = 0;
    }
    for (int j0 = 0; j0 < nj; j0 += kAccGroup) {
      if (j0 + kAccGroup < nj) {
#pragma unroll
        for (int u = 0; u < kAccGroup; ++u) {
          if (j0 + kAccGroup + u < nj) nxt[u] = src[slot * slot_stride];
          if (++slot == ring_slots) slot = 0;
        }
      }
#pragma unroll
      for (int u = 0; u < kAccGroup; ++u) {
        if (j0 + u < nj) {
          const double vx = (double)cur[u].x, vy = (double)cur[u].y;
          const double *aj = a_win + (REAL_A ? 1 : 2) * (long)(j0 + u) * ms;
#pragma unroll
          for (int t = 0; t < kMaxMs; ++t) {
            if (t < mc) {
              const double ar = REAL_A ? aj[t] : aj[2 * t];
              if (REAL_A) {
                acc[t].x += ar * vx;
                acc[t].y += ar * vy;
              } else {
                const double ai = aj[2 * t + 1];
                acc[t].x += ar * vx - ai * vy;
                acc[t].y += ar * vy + ai * vx;
              }
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kAccGroup; ++u) cur[u] = nxt[u];
    }
    if (!packed) {
      const long stlen = 2 * nsg;
#pragma unroll
      for (int t = 0; t < kMaxMs; ++t) {
        if (t < mc) {
          double2 *dst = reinterpret_cast<double2 *>(psi_rank + (item * ms + t0 + t) * stlen + 2 * g);
          if (store) {
            *dst = acc[t];
          } else {
            double2 o = *dst;
            o.x += acc[t].x;
            o.y += acc[t].y;
            *dst = o;
          }
        }
      }
    } else {
      const long stlen = nsg;
#pragma unroll
      for (int t = 0; t < kMaxMs; ++t) {
        if (t < mc) {
          double *d0 = psi_rank + ((2 * item) * ms + t0 + t) * stlen + g;
          *d0 = store ? acc[t].x : *d0 + acc[t].x;
          if (2 * item + 1 < n_starts) {
            double *d1 = psi_rank + ((2 * item + 1) * ms + t0 + t) * stlen + g;
            *d1 = store ? acc[t].y : *d1 + acc[t].y;
          }
        }
      }
    }
  }
}

template <typename T>
int accumulate_window(fbse_ctx *c, const cplx_t<T> *ring, int ring_slots, long slot_stride, long j_lo, long j_hi,
                      const double *d_an, long ms, long ncheby, double *psi_rank, long n_starts, long n_items,
                      int packed, int real_coef, const long *jlo_dev, bool store) {
  const int nj = (int)(j_hi - j_lo);
  if (nj <= 0) return FBSE_OK;
  if (nj > kMaxWin) return set_err(c, FBSE_ERR_UNSUPPORTED, "window %d > %d", nj, kMaxWin);
  hipEvent_t ev = nullptr;
  prof_begin(c, 3, &ev);
  long blocks = (c->nsg + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  const dim3 grid((unsigned)blocks, (unsigned)n_items);
  // up to kMaxMs targets per launch (their accumulators live in registers);
  // more targets per start take further launches over the same window
  for (long t0 = 0; t0 < ms; t0 += kMaxMs) {
    const int mc = (int)std::min<long>(kMaxMs, ms - t0);
    if (real_coef || packed)
      hipLaunchKernelGGL((accumulate_kernel<T, true>), grid, dim3(256), 0, c->stream, ring, ring_slots, slot_stride,
                         j_lo, nj, d_an, (int)ms, (int)t0, mc, psi_rank, n_starts, n_items, packed, c->nsg, jlo_dev,
                         store ? 1 : 0);
    else
      hipLaunchKernelGGL((accumulate_kernel<T, false>), grid, dim3(256), 0, c->stream, ring, ring_slots,
                         slot_stride, j_lo, nj, d_an, (int)ms, (int)t0, mc, psi_rank, n_starts, n_items, packed,
                         c->nsg, jlo_dev, store ? 1 : 0);
  }
  FBSE_HIP(c, hipGetLastError());
  prof_end(c, 3, ev);
  return FBSE_OK;
}
template int accumulate_window<double>(fbse_ctx *, const double2 *, int, long, long, long, const double *, long,
                                       long, double *, long, long, int, int, const long *, bool);
template int accumulate_window<float>(fbse_ctx *, const float2 *, int, long, long, long, const double *, long, long,
                                      double *, long, long, int, int, const long *, bool);

__global__ void step_advance_kernel(long *step, long by) {
  if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += by;
}
int step_advance(fbse_ctx *c, long *step, long by) {
  hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, c->stream, step, by);
  FBSE_HIP(c, hipGetLastError());
  return FBSE_OK;
}

}  // namespace fbse
