// fbse_internal.h -- shared definitions of the gfx950 filter-diagonalization
// library: context, device-side tables and kernel argument blocks.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fbse_gpu.h"

namespace fbse {

// deepest Newton accumulation window K (ring of K+1 slots)
constexpr int kMaxWindow = 127;

template <typename T> struct cplx_of;
template <> struct cplx_of<double> { using type = double2; };
template <> struct cplx_of<float> { using type = float2; };
template <typename T> using cplx_t = typename cplx_of<T>::type;

constexpr int kMaxFactors = 16;
constexpr int kFftThreads = 256;
constexpr int kProjRows = 30;     // 15 SO rows (5 radial x 3 m) + 15 NL rows
constexpr int kProjCols = 16;     // columns (state x spin) per projector work-group
constexpr int kProjPts = 64;      // grid points staged in LDS per projector tile

// One 1D transform length with its radix schedule and tile shape.
struct AxisPlan {
  int n = 0;
  int nf = 0;
  int f[kMaxFactors] = {};
  int lines = 0;       // lines per work-group tile
};

// Argument block of the batched axis FFT kernel (fft_kernels.hip).
template <typename T> struct AxisArgs {
  const cplx_t<T> *src;
  cplx_t<T> *dst;
  const cplx_t<T> *tw;     // exp(-2 pi i t / n), t < n
  int n, nf;
  int f[kMaxFactors];
  int lines;               // lines per tile
  int xaxis;               // 1: lines are contiguous rows
  long inner;              // element stride between axis points (y: nx, z: nx*ny)
  long inner_count;        // distinct inner indices per outer block
  long outer_stride;       // elements between outer blocks (n*inner)
  long tiles_per_outer;
  long total_lines;        // x axis: rows in the batch
  // kinetic scaling (z pass): k2 = min(kx[jx] + ky[jy] + kz[jz], kcap) when the
  // table is separable, else ksqr[jz*nx*ny + inner]
  const T *kx, *ky, *kz, *ksqr;
  T kcap;
  int nx;
  // combine (x backward pass): dst = a1*(T psi + V psi_in) + a0*psi_in
  const cplx_t<T> *psi_in;
  const T *vloc;
  long vloc_rows;          // rows per V period (nspin*nz*ny)
  T a1, a0;
  // when a0_tab is set, a0 = a0_tab[*step_base + step_off] (graph replays)
  const double *a0_tab;
  const long *step_base;
  int step_off;
  // preconditioner (z pass): g = pa / (1 + k2 * pb)
  T pa, pb;
  long ntiles;             // tiles of the launch (fixed-schedule instances loop over them)
  // combine of the fixed instances with the projector expansion of the tile's
  // rows (rec_yoff != nullptr; ny a multiple of the tile's lines): in rounds
  // k < pt_ranks, the records of plane z, rank k and rows y0 .. y0+L-1 are
  // rec_yoff[(z*pt_ranks + k)*(ny+1) + y0] .. rec_yoff[.. + y0 + L] - 1 (record
  // layout of fbse_ctx::d_rec_pa / d_rec_val), the coefficients of batch item
  // i pjd[(i*pj_natoms + atom)*6 + r]
  const int *rec_yoff, *rec_pa;
  const T *rec_val;
  const cplx_t<T> *pjd;
  int pj_natoms, pt_ranks, ny, nz;
  // combine: when set, the forward x transform of the output rows is also
  // written here (the first sweep of the next Newton step)
  cplx_t<T> *fwd_dst;
  // combine: tile order in which the tiles of one row block (all batch items)
  // follow each other on one XCD (work-group b runs on XCD b mod 8), so the
  // row block's records are fetched into one L2 once
  int xcd_order;
  // combine of spinors (16-line fixed instances): a tile holds 8 rows of both
  // spins of one state (batch items in spin pairs, ny a multiple of 8)
  int spin_pair;
  // y and z passes of the 16-line fixed instances (not x): source / result
  // in the z-tile layout, where the 16 x values of one x block of all z of
  // one (item, y) are one contiguous block of nz * 16 elements: element
  // (item, z, y, x) at item * zt_grid + (y * nx / 16 + x / 16) * zt_nz16 +
  // z * 16 + x % 16.  A tile of outer index o and first inner index i0 starts
  // at (o / zt_q) * zt_grid + (o % zt_q) * 16 + (i0 / 16) * zt_nz16 (zt_q = nz
  // for y lines, 1 for z lines), axis points zt_inner apart (nx * nz for y
  // lines, 16 for z lines).
  int zt_src, zt_dst;
  long zt_q, zt_inner, zt_grid, zt_nz16;
  // spin-paired combine (pp_part != nullptr): also the raw projections of its
  // output rows.  Tile tau = z * (ny / 8) + y0 / 8 has lanes pp_lane_off[tau]
  // .. pp_lane_off[tau + 1] - 1; lane i holds the kPpRecs records
  // pp_meta[kPpRecs * i + u] = ((y - y0) << 16) | x, values
  // pp_val[(kPpRecs * i + u) * 8 ..] (the record layout of rec_val; padding
  // records are zero) of one atom, pp_lanes[i] = (slot + 1) << 1 | head: head
  // marks the first lane of a segment (an atom's lanes within one wavefront),
  // slot + 1 > 0 its last lane, whose segment sums go to
  // pp_part[(state * pp_nslots + slot) * 12 + spin * 6 + (SO m, NL m)];
  // pp_span[tau]: the tile's longest segment (lanes, a power of two)
  const int *pp_lane_off, *pp_lanes, *pp_span, *pp_meta;
  const T *pp_val;
  cplx_t<T> *pp_part;
  long pp_nslots;
};

// grids (state, spin columns) one work-group of the projection pass takes
constexpr int kProjColChunk = 16;
// projector records per lane of the projections inside the spin-paired combine
constexpr int kPpRecs = 4;

enum PassKind { kPassFwd = 0, kPassBwd = 1, kPassKin = 2, kPassBwdCombine = 3, kPassPrec = 4 };

// Device projector tables (compact per-atom storage).
template <typename T> struct ProjTables {
  int n_atoms = 0, npts = 0;
  int so = 0, nl = 0, n_so = 0, n_nl = 0;
  int *a_off = nullptr, *a_cnt = nullptr;
  int *jxyz = nullptr;
  T *proj = nullptr;        // [5][npts]
  T *nlproj = nullptr;      // [5][npts]
  cplx_t<T> *y1 = nullptr;  // [3][npts]
  int *sign = nullptr;      // [n_atoms*5]
  cplx_t<T> *ls = nullptr;  // 36
  // atoms grouped by colour: atoms of one colour share no grid point
  int *color_atoms = nullptr;          // atom indices, colour by colour
  std::vector<int> color_off;          // host: start of each colour in color_atoms
  // every atom in the z-y-x order of its sphere's middle point: work-groups
  // that run together then gather overlapping spheres (shared cache lines)
  int *spatial_atoms = nullptr;
};

template <typename T> struct ProjArgs {
  int n_atoms, npts, so, nl, n_so, n_nl;
  const int *atoms;         // atoms of the colour processed by this launch
  const int *a_off, *a_cnt, *jxyz;
  const T *proj, *nlproj;
  const cplx_t<T> *y1;
  const int *sign;
  const cplx_t<T> *ls;
  const cplx_t<T> *psi_in;  // columns: state*nsg + spin*ngrid
  cplx_t<T> *psi_out;
  cplx_t<T> *dout;          // projection only: mixed coefficients [column][atom][6*MAXP]
  long ngrid, nsg;
  int nspin, ncols;
  T scale;                  // dv * a1
};

struct Timer {
  long launches = 0;
  double ms = 0.0;
};

}  // namespace fbse

struct fbse_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // second stream of the Newton step: the projection pass runs there beside
  // the z-slab pass (fork/join events; captured into the window graphs too)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int concurrent_proj = 1;
  int capturing = 0;            // the filter is capturing a Newton window into a hipGraph
  rocblas_handle blas = nullptr;
  std::string err;

  long nx = 0, ny = 0, nz = 0, ngrid = 0, nspin = 1, nsg = 0;
  int is_complex = 0;
  double dv = 0.0;
  int ksqr_sep = 0;
  double kcap = 0.0;

  // local part of H (FP64 and FP32 copies)
  double *d_vloc = nullptr;
  float *d_vloc_f = nullptr;
  double *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_ksqr = nullptr;
  float *d_kx_f = nullptr, *d_ky_f = nullptr, *d_kz_f = nullptr, *d_ksqr_f = nullptr;

  // twiddles per axis
  fbse::AxisPlan plan[3];
  fbse::AxisPlan aplan[3];      // per-axis passes' radix schedule (composite radices allowed)
  double2 *d_tw[3] = {nullptr, nullptr, nullptr};
  float2 *d_tw_f[3] = {nullptr, nullptr, nullptr};

  // projectors
  fbse::ProjTables<double> pj;
  fbse::ProjTables<float> pj_f;
  int has_proj = 0;
  // projector entries of the xy planes for the plane passes: the entries of
  // plane z with rank k (the k-th atom, in atom order, whose sphere holds the
  // point) are records rec_off[z*pt_ranks + k] .. rec_off[z*pt_ranks + k + 1]-1,
  // ordered by point; record e holds rec_pa[2e] = (y << 16) | x and
  // rec_pa[2e+1] = atom, values rec_val[8e ..] = proj, nl_proj, Re/Im Y_1m
  // (m = 0..2), FP64 and FP32 copies.  Within one rank no two records share a
  // point.
  int *d_rec_off = nullptr, *d_rec_pa = nullptr;
  double *d_rec_val = nullptr;
  float *d_rec_val_f = nullptr;
  int pt_ranks = 0;
  // per (plane z, rank k, row y): the first record of bucket (z, k) whose row
  // is >= y, at d_rec_yoff[(z*pt_ranks + k)*(ny+1) + y] (y = ny: the bucket's
  // end); the per-axis combine finds a tile's records with it
  int *d_rec_yoff = nullptr;
  std::vector<int> rec_seg_max;  // per rank: the most records one plane holds
  void *d_pjd = nullptr;        // mixed projection coefficients (column x atom x 6), grown on demand
  size_t pjd_bytes = 0;
  int fuse_proj = 1;            // expansion inside the plane passes when the tables allow it
  // per-axis passes on axes of length 160 / 256: the combine adds the
  // projector expansion and writes the next step's forward x transform
  // (fbse_set_axis_fusion); ax_pjd / ax_fwd carry one combine's operands
  int axis_fuse = 1;
  const void *ax_pjd = nullptr;
  void *ax_fwd = nullptr;
  // projections inside the spin-paired combine (fbse_set_combine_projection, off by default):
  // per tile of 8 rows of one plane its records grouped by atom (AxisArgs::pp_*),
  // the atoms' slots atom-major (atom a: slots pp_atom_off[a] .. pp_atom_off[a+1]-1);
  // the combine writes the per-slot sums into d_pp_part, and the next Newton
  // step reduces them per atom instead of running the projection pass when
  // pp_ready and its input is pp_for (the last such combine's output)
  int pp_on = 0;
  int *d_pp_lane_off = nullptr, *d_pp_lanes = nullptr, *d_pp_span = nullptr, *d_pp_meta = nullptr;
  int *d_pp_atom_off = nullptr;
  double *d_pp_val = nullptr;
  float *d_pp_val_f = nullptr;
  long pp_nslots = 0;
  void *d_pp_part = nullptr;
  size_t pp_part_bytes = 0;
  int pp_chain = 0;             // set while a filter run chains its Newton steps
  int pp_ready = 0;
  const void *pp_for = nullptr;
  long pp_for_n = 0;
  void *ax_pp = nullptr;        // one combine's partial-sum buffer
  // the Newton step's y and z passes in the z-tile layout (AxisArgs::zt_*):
  // the next per-axis pass reads (ax_zt & 1) / writes (ax_zt & 2) that layout
  int ax_zt = 0;

  // scratch (grown on demand)
  void *d_scratch = nullptr;
  size_t scratch_bytes = 0;
  void *d_scratch2 = nullptr;
  size_t scratch2_bytes = 0;
  void *d_gram = nullptr;       // partial Grams of the split-K products
  size_t gram_bytes = 0;
  double *d_red = nullptr;      // reductions
  size_t red_bytes = 0;
  double *h_red = nullptr;      // pinned host mirror
  size_t h_red_bytes = 0;
  // state-file streaming (io.hip): two pinned staging buffers and their events
  void *h_io[2] = {nullptr, nullptr};
  size_t io_bytes = 0;
  size_t io_chunk = (size_t)64 << 20;
  hipEvent_t io_ev[2] = {nullptr, nullptr};

  int window = 0;               // forced window depth (0 = auto)

  // FFT autotuner of the per-axis passes: per precision (0 FP64, 1 FP32),
  // batch size class (0 below 32M elements, 1 above), axis and pass class
  // (0 plain, 1 z kinetic or preconditioner, 2 x combine) the tile lines;
  // 0 = not tuned yet.
  int tune_lines[2][2][3][3] = {};
  int tile_variant[2] = {-1, -1};  // plane/slab kernel instance per precision (-1 = not tuned)
  int zsplit[2] = {0, 0};       // z-slab tiles per slab along x, per precision (0 = not tuned)
  int force_zsplit = 0;         // fbse_set_slab_split (0 = autotune)

  // Newton-window graphs: the Newton shift a0 of step (base + off) is read
  // from a device table while a window is captured (dyn_a0_tab != nullptr)
  int use_graph = 1;
  const double *dyn_a0_tab = nullptr;
  const long *dyn_step_base = nullptr;
  int dyn_step_off = 0;
  long *d_step = nullptr;       // device step counter of the replayed window
  long graph_replays = 0;       // replays in the last filter
  int last_window = 0;          // accumulation window depth K of the last filter
  long last_runs = 0;           // consecutive sub-batch runs of the start states in the last filter
  long run_starts = 0;          // start states per filter run (0: chosen by the filter)
  int last_group = 0;           // contexts that shared grouped FFT launches in the last fbse_filter_many (1: own)
  double tune_ms[2][2][3][3] = {};
  int autotune = 1;
  int no_alloc = 0;             // inside fbse_apply_h_dev_noalloc: buffers may not grow
  int force_variant = -1;       // fbse_set_fft_kernel: -1 auto, 0 per-axis passes, 1 plane/slab passes
  int force_lines = 0;

  // profiling
  int profile = 0;
  fbse::Timer timers[5];
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
};

namespace fbse {

// error helpers
int set_err(fbse_ctx *c, int code, const char *fmt, ...);
// releases the pinned staging buffers of io.hip
void io_release(fbse_ctx *c);
#define FBSE_HIP(c, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::fbse::set_err((c), FBSE_ERR_HIP, "%s failed: %s (%s:%d)", #call,            \
                             hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)
#define FBSE_BLAS(c, call)                                                                 \
  do {                                                                                     \
    rocblas_status s_ = (call);                                                            \
    if (s_ != rocblas_status_success)                                                      \
      return ::fbse::set_err((c), FBSE_ERR_SOLVER, "%s failed: status %d (%s:%d)", #call,  \
                             (int)s_, __FILE__, __LINE__);                                 \
  } while (0)

int ensure_scratch(fbse_ctx *c, size_t bytes);
int ensure_scratch2(fbse_ctx *c, size_t bytes);
int ensure_red(fbse_ctx *c, size_t bytes);

// profiling brackets (no-ops unless profiling is enabled)
void prof_begin(fbse_ctx *c, int kind, hipEvent_t *ev);
void prof_end(fbse_ctx *c, int kind, hipEvent_t ev);
void prof_collect(fbse_ctx *c);

// FFT launchers (fft_kernels.hip).  buf holds nbatch = (states*nspin) grids.
template <typename T>
int fft_pass(fbse_ctx *c, int axis, int kind, const cplx_t<T> *src, cplx_t<T> *dst, long nbatch,
             const cplx_t<T> *psi_in, T a1, T a0);
// Plane/slab FFT passes (fft_tile_kernels.hip): available when every axis
// factors into radices 2, 3, 4, 5, 8 and an xy plane and an xz slab fit the
// 160 KiB LDS of a work-group.  out = a1*(T + V_loc) psi + a0*psi.
template <typename T> bool tiles_supported(const fbse_ctx *c);
// the plane/slab passes of precision T run the fixed-schedule instance
template <typename T> bool tiles_fixed(const fbse_ctx *c);
// Fixed-schedule per-axis passes (fft_tile_kernels.hip): axis lengths with an
// instance, and the launch of one (lines = 16 or 32; A.ntiles tiles)
bool axis_fx_length(int n);
template <typename T> int axis_fx_pass(fbse_ctx *c, int kind, const AxisArgs<T> &A, int lines);
// Grouped launches of the FP32 fixed-schedule plane/slab passes
// (fft_tile_kernels.hip): one launch runs pass `kind` (kTileFwdXY, kTileKinZ,
// kTileBwdXYCombine) for n_ctx <= kTileGroupMax contexts of one grid size,
// each with its own operands, on cs[0]'s stream.  tile_groupable: the context
// runs the plane/slab passes and has a fixed FP32 instance (not forced off).
constexpr int kTileGroupMax = 8;
enum TileKind { kTileFwdXY = 0, kTileBwdXY = 1, kTileBwdXYCombine = 2, kTileKinZ = 3, kTilePrecZ = 4 };
bool tile_groupable(const fbse_ctx *c);
int tile_group_pass(int n_ctx, fbse_ctx *const *cs, int kind, const float2 *const *src, float2 *const *dst,
                    const long *nbatch, const float2 *const *psi_in, const float *a1, const float *a0,
                    const float2 *const *pjd, float2 *const *fwd_dst);
// pjd != nullptr: the combine pass also adds the projector expansion with the
// coefficients of proj_project
// fwd_next: the combine pass also leaves the forward xy transform of out in
// work (the first pass of the next Newton step); fwd_ready: work already
// holds the forward xy transform of psi, so the first pass is skipped.
// join: when set, the stream waits for this event before the combine pass
// (the projection pass that fills pjd ran on another stream)
template <typename T>
int kinetic_local_tiles(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1,
                        T a0, const cplx_t<T> *pjd = nullptr, bool fwd_ready = false, bool fwd_next = false,
                        hipEvent_t join = nullptr);
template <typename T>
int precondition_tiles(fbse_ctx *c, const cplx_t<T> *r, cplx_t<T> *out, cplx_t<T> *work, long nstates, T pa, T pb);

// FFT path in use for a precision (f32 = 1: the mixed filter's FP32 passes):
// 1 plane/slab passes, 0 per-axis passes; *instance = plane/slab kernel
// instance (1-based, 0 = per-axis or not tuned yet).
int fft_path(const fbse_ctx *c, int f32, int *instance);

// Full H (or Newton step) for nstates complex states: out = a1*H psi + a0*psi.
// work must hold nstates*nsg complex values.
template <typename T>
int apply_h_cplx(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1,
                 T a0);
// One step of the Newton recurrence, out = a1*H psi + a0*psi, for a chain of
// steps out_j = step(out_{j-1}) that keep `work` between calls: with the
// plane/slab passes each step also leaves the forward xy transform of out in
// work for the next one.  first: psi does not come from the previous step.
template <typename T>
int apply_h_chain(fbse_ctx *c, const cplx_t<T> *psi, cplx_t<T> *out, cplx_t<T> *work, long nstates, T a1, T a0,
                  bool first);
// The same step for n <= kTileGroupMax contexts of one grid size on one
// stream (fft_kernels.hip): projections per context, FFT passes as grouped
// launches (tile_groupable contexts, FP32, projectors absent or expanded in
// the plane passes).
// projection-only launch for n <= kTileGroupMax contexts (FP32, one radial
// projector each) on cs[0]'s stream (ham_kernels.hip)
int proj_project_group(int n, fbse_ctx *const *cs, const float2 *const *psi, const long *nstates, const float *a1,
                       float2 *const *d);
int apply_h_group(int n, fbse_ctx *const *cs, const float2 *const *psi, float2 *const *out, float2 *const *work,
                  const long *nstates, const float *a1, const float *a0, bool first);

// n contiguous states (reference layout, FP64): out = F^-1 [ F[r] / (N (1 + 0.5|k|^2 / e_p)) ]
// (capi.hip; out may equal r)
int precondition_dev(fbse_ctx *c, const double *d_r, double *d_out, long n, double e_p);

// kinetic-energy preconditioner out = F^-1 [ F[r] / (N (1 + 0.5|k|^2 / e_p)) ]
template <typename T>
int precondition_cplx(fbse_ctx *c, const cplx_t<T> *r, cplx_t<T> *out, cplx_t<T> *work, long nstates, double e_p);

// expansion from the rank records (ham_kernels.hip, one launch per rank):
// out += the expansion of the coefficients d of proj_project
template <typename T>
int proj_expand(fbse_ctx *c, const cplx_t<T> *d, cplx_t<T> *out, long nstates);

// projector launcher (ham_kernels.hip): psi_out += scale * Vproj psi_in
template <typename T>
int proj_apply(fbse_ctx *c, const cplx_t<T> *psi_in, cplx_t<T> *psi_out, long nstates, double a1);
// projections only: d[(col*n_atoms + atom)*6 + r] = the L.S / sign mixed, dv*a1
// scaled coefficients (r < 3 spin-orbit m, r >= 3 nonlocal m) of one radial
// projector of each kind; returns FBSE_ERR_UNSUPPORTED when the tables have more
template <typename T>
int proj_project(fbse_ctx *c, const cplx_t<T> *psi_in, long nstates, double a1, cplx_t<T> *d);
// true when the plane passes can add the projector expansion (one radial
// projector of each kind, entries by point built)
bool proj_fusable(const fbse_ctx *c);
// d (column x atom x 6) from the per-tile projection sums of the last
// spin-paired combine (fbse_ctx::d_pp_part) for nstates states
template <typename T> int proj_reduce(fbse_ctx *c, long nstates, double a1, cplx_t<T> *d);

// element-wise helpers (ham_kernels.hip)
int pack_real(fbse_ctx *c, const double *src, long src_stride, long nstates, double2 *dst);
int unpack_real(fbse_ctx *c, const double2 *src, long nstates, double *dst, long dst_stride);
int to_f32(fbse_ctx *c, const double2 *src, float2 *dst, long n);
int to_f32_strided(fbse_ctx *c, const double2 *src, long src_stride, float2 *dst, long n, long count);
int fill_zero(fbse_ctx *c, void *dst, size_t bytes);
// per-state sums over nsg complex elements (or real layout): out[s] = sum |a|^2,
// sum Re(conj(a) b), sum |b - e a|^2 ...
int state_norms(fbse_ctx *c, const double *psi, long nstates, long elems_per_state, double *d_out);
int state_dots(fbse_ctx *c, const double *a, const double *b, long nstates, long elems_per_state, double *d_out);
int state_resid(fbse_ctx *c, const double *a, const double *ha, const double *eps, long nstates,
                long elems_per_state, double *d_out);
int state_scale(fbse_ctx *c, double *psi, long nstates, long elems_per_state, const double *d_fac);
int axpby_states(fbse_ctx *c, const double *x, double *y, long n, double a, double b);  // y = a*x + b*y
// y_s -= theta_s * x_s for nstates states of elems doubles (theta on the device)
int states_sub_scaled(fbse_ctx *c, const double *x, double *y, long nstates, long elems, const double *d_theta);

// FP64 matrix-core products (gemm_kernels.hip).  Columns of m elements
// (complex double2 or real double), column-major.
// Gram: part[p*pstride + i + j*ldp] = sum over the rows of slab p of
// conj(A[r, i]) B[r, j]; upper (n1 == n2): only the 64 x 64 tiles on or above
// the diagonal are formed.  Returns 0 or -1 (launch error / shape limits).
long gram_tile_count(long n1, long n2, bool upper);
int launch_gram_mfma(hipStream_t s, bool cplx, const double *A, long lda, const double *B, long ldb, long m, long n1,
                     long n2, bool upper, long slab, long parts, double *part, long ldp, long pstride);
// Rotation: Out[r, j] (+)= sum_k A[r, k] V[k, j], r < nrows, k < nin, j < nout
// (accumulate: added to Out).
int launch_rot_mfma(hipStream_t s, bool cplx, const double *A, long lda, long nrows, long nin, const double *V,
                    long ldv, long nout, double *Out, long ldo, bool accumulate = false);

// window accumulation (ham_kernels.hip)
// jlo_dev != nullptr: the window starts at step *jlo_dev (read on the device,
// for graph replays); j_hi - j_lo still gives the window length.  store: the
// window's sums are written instead of added (the first window of a filter,
// so psi_rank needs no clearing beforehand).
template <typename T>
int accumulate_window(fbse_ctx *c, const cplx_t<T> *ring, int ring_slots, long slot_stride, long j_lo,
                      long j_hi, const double *d_an, long ms, long ncheby, double *psi_rank, long n_starts,
                      long n_items, int packed, int real_coef, const long *jlo_dev = nullptr, bool store = false);
// *step += by (one thread; graph-safe counter update)
int step_advance(fbse_ctx *c, long *step, long by);

}  // namespace fbse
