/*
 * fbse_gpu.h -- C ABI of the MI355X (gfx950) filter-diagonalization library.
 *
 * Plain C, POD only: no FFTW, MKL, MPI or HIP types cross this boundary, so
 * FilterBSE's host C (which includes fd.h -> mpi.h/mkl.h/fftw3.h) can call it
 * without either side seeing the other's headers.  Every host array stays
 * caller-owned; device buffers are owned by the context.  All functions return
 * FBSE_OK (0) or a negative FBSE_ERR_* code; fbse_last_error() gives the text.
 * A context owns one device and one HIP stream and is not re-entrant across
 * threads (the reference calls these from each MPI rank's master thread).
 *
 * Layouts are the reference's (filter_mpi/):
 *   grid index   jxyz = jx + nx*(jy + ny*jz)                (init.c:592-595)
 *   spinor       [up: ngrid | down: ngrid]                    (phamiltonian.c:236)
 *   one state    stlen = nspin*ngrid*(is_complex+1) doubles, complex values
 *                interleaved re,im                            (filter.c:259-278)
 *   state block  psi_rank[(jns*ms + jms)*stlen + ...]         (filter.c:253,286)
 *   coefficients an[2*(ncheby*jms + jc)] re,im                (filter.c:287,331)
 *   Leja nodes   zn[jc-1] used at Newton step jc             (filter.c:320)
 *   projectors   per-atom tables of stride n_gridpts          (init.c:1016)
 * "*_dev" entry points take device pointers (buffers already resident in HBM)
 * in the same layouts; the others take host pointers.
 */
#ifndef FBSE_GPU_H
#define FBSE_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define FBSE_OK 0
#define FBSE_ERR_ARG (-1)
#define FBSE_ERR_HIP (-2)
#define FBSE_ERR_NORM_ZERO (-3)   /* normalize_all: "Norm is 0! exiting!" (norm.c:150-154) */
#define FBSE_ERR_NORM_NAN (-4)    /* normalize_all: "Norm is NAN! exiting!" (norm.c:155-159) */
#define FBSE_ERR_SOLVER (-5)      /* LAPACK info != 0 analogue (ortho.c:56-62, Hmat.c:202-216) */
#define FBSE_ERR_NOMEM (-6)
#define FBSE_ERR_UNSUPPORTED (-7)
#define FBSE_ERR_IO (-8)         /* a state file could not be opened, read or written (fbse_io.h) */

#define FBSE_PREC_FP64 0          /* FFT, recurrence and accumulation in FP64 */
#define FBSE_PREC_MIXED 1         /* FP32 complex FFT/recurrence, FP64 coefficients,
                                     FP64 accumulators, FP64 ortho/RR */

typedef struct fbse_ctx fbse_ctx;

/* Grid and local part of H.  Flat copies of ist->nx/ny/nz/nspin, flag->isComplex,
 * par->dv, ksqr (build_grid_ksqr, init.c:318-349: 0.5|k|^2/N, capped at
 * KE_max/N) and pot_local (nspin*ngrid, build_local_pot, init.c:666). */
typedef struct {
  long nx, ny, nz;
  long nspin;
  int is_complex;
  double dv;
  const double *ksqr;
  const double *pot_local;
} fbse_grid_desc;

/* Spin-orbit / nonlocal projector tables: the nlc_st array (fd.h:382-398)
 * split into structure-of-arrays by the caller.  Index idx = atom*n_gridpts + g
 * for g < n_gpt[atom] (init_NL_projectors, init.c:988-1068). */
typedef struct {
  int so, nl;               /* flag->SO, flag->NL */
  long n_atoms;             /* ist->n_NL_atoms */
  long n_gridpts;           /* ist->n_NL_gridpts (table stride) */
  const long *n_gpt;        /* nl[atom] */
  const long *jxyz;         /* nlc[idx].jxyz */
  const double *y1;         /* nlc[idx].y1[m] re,im: 6 doubles per idx */
  const double *proj;       /* nlc[idx].proj[p]: 5 doubles per idx */
  const double *nl_proj;    /* nlc[idx].NL_proj[p]: 5 doubles per idx */
  const int *nl_sign;       /* nlc[atom*n_gridpts].NL_proj_sign[p]: 5 ints per atom */
  int n_so_proj;            /* radial SO projectors (5 in phamiltonian.c:225) */
  int n_nl_proj;            /* ist->nproj */
  const double *LS;         /* L.S, 36 complex (def_LS, hamiltonian.c:312-442) */
} fbse_proj_desc;

/* Replaces the FFTW plan set-up and table copies of run_filter_cycles
 * (filter.c:5-91) and of every H consumer.  proj may be NULL (no SO/NL).
 * device < 0 selects the current device. */
int fbse_create(const fbse_grid_desc *grid, const fbse_proj_desc *proj, int device, fbse_ctx **out);
void fbse_destroy(fbse_ctx *ctx);
const char *fbse_last_error(const fbse_ctx *ctx);

/* Use the caller's HIP stream (a hipStream_t passed as void*); NULL restores the
 * context's own stream. */
int fbse_set_stream(fbse_ctx *ctx, void *stream);
int fbse_synchronize(fbse_ctx *ctx);

/* H applied to nstates states: replaces p_hamiltonian (phamiltonian.c:10-60) /
 * hamiltonian (hamiltonian.c:9-57) as used by energy (energy.c:6-57),
 * energy_all (energy.c:61-165), diag_H (Hmat.c:141) and calc_sigma_E
 * (energy.c:474,500).  Real layouts return the real part, as those callers read it. */
int fbse_apply_h(fbse_ctx *ctx, const double *psi, double *hpsi, long nstates);
int fbse_apply_h_dev(fbse_ctx *ctx, const double *d_psi, double *d_hpsi, long nstates);

/* fbse_apply_h_dev that allocates nothing (no hipMalloc / hipFree, no
 * allocating library routine): for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  No counterpart in the reference.  Its work space is one static device
 * array of 640 MiB per GPU, held from library load to process end; calls
 * needing more (nstates x nspin*ngrid x 16 bytes of FFT work, real states
 * packed two per grid, plus the projection coefficients) return FBSE_ERR_ARG
 * without running.  Autotuning is off for the call (a context that already
 * ran fbse_apply_h_dev reuses its tuned FFT kernels).  All calls on one GPU
 * share the array: calls that overlap on one GPU (several threads, or
 * contexts on different streams) race and must be serialised by the caller. */
int fbse_apply_h_dev_noalloc(fbse_ctx *ctx, const double *d_psi, double *d_hpsi, long nstates);

/* The Newton filter over all start states of psi_rank: replaces the loop
 * filter.c:102-116 (filter_cycle, filter.c:187-371, with p_hnorm filter.c:375-410).
 * On entry start state jns sits at psi_rank[jns*ms*stlen]; on exit
 * psi_rank[(jns*ms+jms)*stlen] holds sum_jc an[jms,jc] psi_jc.  Not normalised
 * (run_filter_cycles calls normalize_all afterwards). */
int fbse_filter(fbse_ctx *ctx, double *psi_rank, long n_starts, long ms, long ncheby, const double *an,
                const double *zn, double Vmin, double dE_1, int precision);
int fbse_filter_dev(fbse_ctx *ctx, double *d_psi_rank, long n_starts, long ms, long ncheby, const double *an,
                    const double *zn, double Vmin, double dE_1, int precision);

/* Filters of several dots (one context each, e.g. of different grid sizes) in
 * one call.  In mixed precision, dots of one grid size that run the fixed
 * FP32 plane/slab instance (cubic 60 ... 120 grids; projectors absent or with
 * one radial projector) form groups of up to 8 whose FFT passes are single
 * grouped launches over all of the group's dots (one launch per pass and
 * Newton step; the group's windows are one hipGraph) on the first dot's
 * stream; every other dot's recurrence runs on its own context's stream.  The
 * host issues the groups' and dots' steps and graph replays interleaved so
 * their kernels overlap on the GPU, and the window depth is sized for all
 * rings together.  Per dot d the arguments mean what they mean for
 * fbse_filter_dev; ms and M are common.  Results equal n_dots separate
 * fbse_filter_dev calls that run the same FFT instance (grouped dots always
 * run the fixed one).  No counterpart in the reference (one dot per run);
 * this is the batched uneven-dot workload. */
int fbse_filter_many(int n_dots, fbse_ctx *const *ctx, double *const *d_psi_rank, const long *n_starts, long ms,
                     long M, const double *const *an, const double *const *zn, const double *Vmin,
                     const double *dE_1, int precision);

/* normalize_all (norm.c:82-180): each state scaled to sum |psi|^2 dv = 1.
 * Returns FBSE_ERR_NORM_ZERO / FBSE_ERR_NORM_NAN where the reference exits. */
int fbse_normalize_all(fbse_ctx *ctx, double *psitot, long n_states);
int fbse_normalize_all_dev(fbse_ctx *ctx, double *d_psitot, long n_states);

/* energy_all (energy.c:61-165): ene[i] = <psi_i|H|psi_i> dv (host array out). */
int fbse_energy_all(fbse_ctx *ctx, const double *psitot, long n_states, double *ene);
int fbse_energy_all_dev(fbse_ctx *ctx, const double *d_psitot, long n_states, double *ene);

/* ortho_real / ortho_cplx (ortho.c:5-185): the first *rank columns of psitot
 * are overwritten with an orthonormal basis of the left singular space whose
 * singular values satisfy s_i/s_0 >= svd_eps (SVDEPS = 1e-10, fd.h:535), then
 * normalised as mod_ortho.c:84 does.  sing_vals (n_cols, may be NULL)
 * receives the singular values (descending) in units of the unweighted
 * column norms. */
int fbse_ortho(fbse_ctx *ctx, double *psitot, long n_cols, double svd_eps, long *rank, double *sing_vals);
int fbse_ortho_dev(fbse_ctx *ctx, double *d_psitot, long n_cols, double svd_eps, long *rank, double *sing_vals);

/* diag_H (Hmat.c:5-297) followed by normalize_all (mod_diag.c:433): builds
 * H_ij = <psi_i|H|psi_j> dv, diagonalises it (ascending eigenvalues into
 * eval) and rotates psitot in place into the eigenvectors. */
int fbse_rayleigh_ritz(fbse_ctx *ctx, double *psitot, long n, double *eval);
int fbse_rayleigh_ritz_dev(fbse_ctx *ctx, double *d_psitot, long n, double *eval);

/* calc_sigma_E (energy.c:383-539): sigma_E[i] = sqrt|<H^2> - <H>^2| and
 * eval_aux[i] = <H> (may be NULL), evaluated as ||H psi - <H> psi|| dv^(1/2)
 * (the same quantity for Hermitian H, without the cancellation). */
int fbse_sigma(fbse_ctx *ctx, const double *psitot, long n, double *sigma_E, double *eval_aux);
int fbse_sigma_dev(fbse_ctx *ctx, const double *d_psitot, long n, double *sigma_E, double *eval_aux);

/* Subspace products on the FP64 matrix cores (device state blocks, column i
 * of a block at d_a + i*stlen).  fbse_overlap_dev: s[i + j*n1] (host, n1 x n2,
 * complex interleaved for complex states) = dv * sum conj(a_i) b_j, the
 * quantity dotp / dotpreal return (Hmat.c:301-338) and diag_H assembles into
 * H_ij (Hmat.c:141-164).  fbse_rotate_dev: in place, column j < nout of the
 * block becomes sum_k a_k v[k + j*nin] (host v, nin x nout, nout <= nin), the
 * change of basis of diag_H (Hmat.c:230-281). */
int fbse_overlap_dev(fbse_ctx *ctx, const double *d_a, long n1, const double *d_b, long n2, double *s);
int fbse_rotate_dev(fbse_ctx *ctx, double *d_a, long nin, const double *v, long nout);

/* One guarded FP64 refinement sweep after a mixed-precision filter (no
 * counterpart in the reference, whose filter is FP64 throughout).  Input: the
 * n_ritz orthonormal Ritz vectors of a preceding fbse_rayleigh_ritz with their
 * eval and sigma (fbse_sigma).  The pairs with eval in [e_lo, e_hi] and
 * sigma < sigma_max (lowest sigma first, as many as the device memory holds
 * two blocks of) get residuals r = H u - e u, preconditioned in k-space
 * (r / (1 + 0.5|k|^2 / e_p); e_p <= 0 selects 1 Ha) and orthogonalised
 * against all n_ritz Ritz vectors.  Over that extended basis each selected
 * pair gets its refined vector, the unit vector with the least
 * ||H v - e v|| (degenerate pairs together), which replaces the pair only
 * when its own residual is lower and it duplicates no other state.  On return
 * psitot holds n_ritz pairs in the same slots, eval_out / sigma_out (n_ritz-
 * long host arrays) their values, *n_replaced how many changed: no pair is
 * dropped and none gets a larger residual. */
int fbse_refine(fbse_ctx *ctx, double *psitot, long n_ritz, const double *eval, const double *sigma,
                double sigma_max, double e_lo, double e_hi, double e_p, double *eval_out, double *sigma_out,
                long *n_replaced);
int fbse_refine_dev(fbse_ctx *ctx, double *d_psitot, long n_ritz, const double *eval, const double *sigma,
                    double sigma_max, double e_lo, double e_hi, double e_p, double *eval_out, double *sigma_out,
                    long *n_replaced);

/* get_energy_range iterative branch (energy.c:222-300) and range inflation
 * (energy.c:343-369).  psi0_min / psi0_max are the two host-drawn start vectors
 * (init_psi with the fixed seed, energy.c:207,228-231,270-273), one state each
 * in the reference layout.  out[0..5] = Emin, Emax, par->Vmin, par->dE,
 * iterations(Emin), iterations(Emax).  nl_lambda_min / vloc_min drive the
 * NL-only trimmed inflation (pass 0.0 for the blanket one). */
int fbse_energy_range(fbse_ctx *ctx, const double *psi0_min, const double *psi0_max, long max_iter,
                      double nl_lambda_min, double vloc_min, double *out);

/* Per-kernel GPU time recorded with HIP events on the context's stream while
 * profiling is on (for roofline reporting).  kind: 0 = FFT passes (all axes),
 * 1 = x-backward+Newton combine, 2 = projectors, 3 = window accumulation,
 * 4 = all kernels of the filter.  Returns launches and total milliseconds. */
int fbse_profile(fbse_ctx *ctx, int enable);
int fbse_profile_read(fbse_ctx *ctx, int kind, long *launches, double *total_ms);

/* Window depth K of the blocked Newton accumulation, FFT tile widths etc. are
 * chosen by the built-in autotuner at the first call; this forces K
 * (1..127; 0 = auto: as deep as 75% of the free HBM allows, within 120 GB
 * of ring and at most 127). */
int fbse_set_window(fbse_ctx *ctx, int window);

/* FFT pass selection for the kinetic operator: variant -1 = automatic
 * (default: plane/slab passes when available, else per-axis passes),
 * 0 = five per-axis passes (any length), 1 = three plane/slab passes (every
 * axis length a product of 2, 3, 4, 5, 8 and an nx x max(ny, nz) tile within
 * 160 KiB of LDS; otherwise the per-axis passes run).  lines > 0 fixes,
 * instead of autotuning, the tile height of the per-axis passes (variant 0)
 * or the plane/slab kernel instance, 1-based (variant 1: FP32 instances
 * 1..3 stage 12, 16, 32 elements per thread, FP64 instances 1..2 stage 12, 16;
 * on cubic grids with a separable k^2 table of a size with a fixed instance
 * (FP32: 60, 64, 72, 80, 90, 96, 100, 108, 120; FP64: 64, 96) one more
 * instance, FP32 4 / FP64 3, runs the fixed-schedule persistent passes).
 * Variant 0 with lines = 1016, 1032 or 1064 (FP32 only) forces the
 * fixed-schedule per-axis instances (16, 32 or 64 lines per tile) on axes of
 * length 160 or 256 (the generic kernel with that many lines elsewhere); the
 * autotuner considers all of them. */
int fbse_set_fft_kernel(fbse_ctx *ctx, int variant, int lines);

/* Plane/slab passes: the z-slab pass cuts each xz slab into `split` tiles of
 * nx/split columns (smaller LDS tiles, more work-groups per CU).  0 = autotune
 * (the default), otherwise forced when it divides nx and leaves at least 16
 * columns (else 1).  Results do not depend on it. */
int fbse_set_slab_split(fbse_ctx *ctx, int split);

/* Newton-window graphs (default on): the filter captures one full
 * accumulation window (R Newton steps, their accumulation and a device step
 * counter update) into a hipGraph and replays it for every further full
 * window; 0 launches every kernel from the host instead.  Results are
 * identical.  No counterpart in the reference. */
int fbse_set_graph(fbse_ctx *ctx, int enable);

/* Per-axis passes on an x axis of length 160 or 256 with ny a multiple of 16
 * (default on): the Newton filter's x combine pass (backward x transform,
 * V_loc, Newton shift) also adds the projector expansion of its rows and
 * writes the forward x transform of its output, which the next Newton step
 * starts from; 0 runs the expansion and the forward x transform as passes of
 * their own.  enable = 1 (default): tiles of 16 rows of one grid, for spinors
 * 8 rows of both spins of one state (the tile's projector records are read
 * once for the two grids); 16, 32 or 64 (FP32): tiles of that many
 * consecutive rows of one grid.  Results are identical.  No counterpart in
 * the reference (its kinetic and projector terms are separate loops,
 * hamiltonian.c:62-99, phamiltonian.c:140-444). */
int fbse_set_axis_fusion(fbse_ctx *ctx, int enable);

/* Projections inside the spin-paired combine (default off): in the Newton
 * filter's chain of steps over spinors with the combine's 8-row tiles of both
 * spins (fbse_set_axis_fusion 1, ny a multiple of 8, one radial SO and NL
 * projector), the combine also sums the projections of its output rows per
 * tile and atom, and the next step adds those sums per atom in a fixed order
 * instead of running the projection pass, which gathers the sphere points
 * from HBM again; 0 (default) runs the projection pass every step: it moves
 * 4 GB less per 8-state launch at 256^3 but the combine grows by more than
 * the pass it replaces.  Results agree to the rounding of the changed
 * summation order.  No counterpart in the
 * reference (its projections are a separate loop over the atoms' sphere
 * points, phamiltonian.c:140-444). */
int fbse_set_combine_projection(fbse_ctx *ctx, int enable);

/* Releases the context's work buffers (the filter's ring of Newton terms,
 * FFT work space); they are allocated again on demand.  For callers that keep
 * many contexts alive (one per dot) and filter them in turns. */
int fbse_trim(fbse_ctx *ctx);

/* Graph replays of the last filter call (0 when every window ran eagerly). */
long fbse_graph_replays(const fbse_ctx *ctx);
/* Accumulation window depth K of the last filter call (fbse_set_window, or
 * the depth the free memory allowed); for reporting. */
int fbse_last_window(const fbse_ctx *ctx);
/* Consecutive runs the last fbse_filter call split its start states into (1:
 * one batch).  A large batch runs as sub-batches of the start states when
 * their deeper rings (fewer items per ring slot) lower the accumulation's
 * traffic by more than the extra runs cost; fbse_set_window keeps the batch
 * whole.  For reporting.  No counterpart in the reference (its ranks filter
 * one start each, filter.c:102-116). */
long fbse_last_runs(const fbse_ctx *ctx);
/* Start states per filter run: 0 (default) chooses as above; n > 0 runs the
 * start states in consecutive sub-batches of n (real states: n rounded up to
 * an even count), whatever the window depth.  Results equal one batch's at the
 * same window depth.  No counterpart in the reference. */
int fbse_set_run_starts(fbse_ctx *ctx, long n);
/* Dots whose FFT passes ran as shared grouped launches with this context in
 * its last fbse_filter_many call (1: its own launches; 0: no such call yet).
 * For reporting.  No counterpart in the reference (one dot per run). */
int fbse_last_group(const fbse_ctx *ctx);
/* Device memory the context holds between calls: out[0..3] = the filter ring /
 * work buffer, the second work buffer, the partial Grams, the rocBLAS
 * workspace (bytes).  For reporting; fbse_trim releases the first three. */
int fbse_held_bytes(fbse_ctx *ctx, size_t *out);

/* The FFT path the context uses for a precision: *path = 1 plane/slab
 * passes, 0 per-axis passes; *instance (may be NULL) = the plane/slab kernel
 * instance, 1-based, as picked by the autotuner or forced (0 when the
 * per-axis passes run or the autotuner has not run yet).  No counterpart in
 * the reference (its FFTW plans are opaque); for reporting. */
int fbse_fft_choice(fbse_ctx *ctx, int precision, int *path, int *instance);

/* Device memory helpers for callers without a GPU runtime of their own. */
int fbse_dev_alloc(fbse_ctx *ctx, size_t bytes, void **d_ptr);
int fbse_dev_free(fbse_ctx *ctx, void *d_ptr);
int fbse_memcpy_h2d(fbse_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int fbse_memcpy_d2h(fbse_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

/* ---- input side on the GPU (host arrays in and out; device = -1: current) ----
 * Results are bit-identical to the reference's arithmetic (same IEEE operations
 * in the same order, no fused multiply-add). */

/* build_local_pot (init.c:587-669): pot_local[jxyz] = sum over all atoms, in
 * index order, of interpolate(|r - R_a|) in the table of atom type atype[a]:
 * pot_r / pot_v [n_types x max_len] rows, file_lens[t] rows used, pot_dr[t]
 * the row spacing; 0 beyond a table.  nspin = 2 repeats the grid (init.c:666).
 * Grid point (jx, jy, jz) is (gx[jx], gy[jy], gz[jz]), jxyz = jx + nx (jy + ny jz). */
int fbse_build_local_pot(int device, long nspin, const double *gx, const double *gy, const double *gz, long nx,
                         long ny, long nz, const double *R, const int *atype, long natoms, const double *pot_r,
                         const double *pot_v, const long *file_lens, const double *pot_dr, long n_types,
                         long max_len, double *pot_local);

/* The per-atom count of grid points with |r - R_a|^2 < R_NLcut2 and its
 * maximum, n_NL_gridpts (init.c:351-394).  Returns the count or a negative
 * error code. */
long fbse_count_nl_gridpts(int device, const double *R, long natoms, const double *gx, const double *gy,
                           const double *gz, long nx, long ny, long nz, double R_NLcut2);

/* init_NL_projectors (init.c:988-1068) in structure-of-arrays form: for atom a
 * its nl[a] points in grid order at slots a*stride + k: jxyz, y1 (3 complex,
 * the Y_1m with the EPSDX offsets of init.c:1028-1038), proj[5 per slot]
 * (radial table so_tab[p*1024 + i] interpolated at |r|, times sqrt(SO_par[a]),
 * init.c:1046) and nl_proj[5 per slot] (nl_tab[(a*nproj + p)*1024 + i]); the
 * radial grid is i*sqrt(R_NLcut2)/1024.  sign[a*5 + p] = nl_sgn[a*5 + p].
 * Slots beyond nl[a] and projectors beyond nproj are set to 0.  Returns
 * FBSE_ERR_ARG when an atom has more than stride points. */
int fbse_init_nl_projectors(int device, long *nl, long *jxyz, double *y1, double *proj, double *nl_proj, int *sign,
                            const double *R, long natoms, const double *gx, const double *gy, const double *gz,
                            long nx, long ny, long nz, double R_NLcut2, long stride, long nproj,
                            const double *so_tab, const double *nl_tab, const int *nl_sgn, const double *SO_par);

#ifdef __cplusplus
}
#endif

#endif /* FBSE_GPU_H */
