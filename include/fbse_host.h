/*
 * fbse_host.h -- host-side C of the filter-diagonalization path that feeds the
 * GPU library: the parts FilterBSE computes on the host before the filter
 * (start states, grid and k^2 tables, Leja nodes, Newton coefficients, energy
 * targets, L.S).  Same argument meaning as the reference functions named in
 * each comment (filter_mpi/), with flat arrays instead of the fd.h structs.
 * A host that keeps its own copies of these (the reference itself) only needs
 * fbse_gpu.h; this library serves callers that do not.
 */
#ifndef FBSE_HOST_H
#define FBSE_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* ran_nrc (rand.c:22-62): Numerical Recipes ran2, process-wide state as in the
 * reference; a non-positive *seed re-initialises it. */
double fbse_ran_nrc(long *seed);

/* init_psi (init.c:1256-1305) for one spin component: uniform [-1,1) values in
 * grid order (imaginary part drawn after the real part when is_complex),
 * normalised to sum |psi|^2 dv = 1.  psi is ngrid complex (re,im pairs). */
void fbse_init_psi(double *psi, long *seed, long nx, long ny, long nz, int is_complex, double dv);

/* init_filter_states random branch (init.c:1207-1253): start jns at
 * psi_rank[jns*ms*stlen], spin components drawn in order. */
void fbse_init_filter_states(double *psi_rank, long n_starts, long ms, long *seed, long nx, long ny, long nz,
                             long nspin, int is_complex, double dv);

/* init_grid_params (init.c:5-255, non-periodic): n[3] in = requested, out =
 * final (auto-grown, even); d[3] spacings; out mins[3], maxs[3], dk[3]. */
void fbse_init_grid_params(const double *R, long natoms, long *n, const double *d, double *mins, double *maxs,
                           double *dk);

/* build_grid_ksqr (init.c:286-349): grid axes and ksqr = min(0.5|k|^2/N, KE_max/N). */
void fbse_build_grid_ksqr(double *ksqr, double *gx, double *gy, double *gz, const long *n, const double *d,
                          const double *mins, const double *dk, double KE_max);

/* samp_points_ashkenazy (coeff.c:227-293) on [-2,2]: the ncheby Leja nodes
 * (imaginary parts are identically 0).  Returns the capacity estimate fk. */
double fbse_samp_points_ashkenazy(double *zn, long ncheby, int nthreads);

/* gen_newton_coeff (coeff.c:58-115): an[2*(ie*ncheby+j)] re,im for the
 * Gaussian filters sqrt(dt/pi) exp(-dt (E - target)^2). */
void fbse_gen_newton_coeff(double *an, const double *zn, const double *targets, long ms, long ncheby, double Vmin,
                           double dE, double dt, int nthreads);

/* set_ene_targets (init.c:401-478, setTargets branch). */
void fbse_set_ene_targets(double *targets, long nVB, long nCB, double VBmin, double VBmax, double CBmin,
                          double CBmax);

/* def_LS (hamiltonian.c:312-442): L.S in the flattened (s,m) basis, LS[jp*6+j]. */
void fbse_def_LS(double *LS);

#ifdef __cplusplus
}
#endif

#endif /* FBSE_HOST_H */
