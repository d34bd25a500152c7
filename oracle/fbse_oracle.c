/*
 * fbse_oracle.c -- CPU restatement of FilterBSE's filter-diagonalization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke() check and bench.py's cpu_baseline leg may load it, nothing in the
 * shipped library (filterbse_amd/) links or calls it.  Every function states
 * the reference file:line it restates (paths relative to the FilterBSE tree,
 * filter_mpi/ unless noted).  Arithmetic is kept in the reference's order
 * (compile with -ffp-contract=off) so that integer/index work is bit-exact and
 * floating-point work differs only where the reference used FFTW/MKL.
 *
 * Parity pins (what checks each function; tests/test_oracle.py):
 *   or_ran_nrc            Numerical Recipes ran2 known answers + structure
 *   or_init_kvec/ksqr     closed form k = j*dk, KE cap, separability
 *   or_init_grid_params   CsPbI3 run.dat grid (44^3, xmin=-11) + rules
 *   or_samp_points        zn.dat (M=2048) golden fixture from the reference
 *   or_gen_newton_coeff   interpolation identity at the Leja nodes, and the
 *                         Gaussian filter reproduced on [-2,2]
 *   or_fft3d              brute-force DFT on tiny grids, numpy.fft
 *   or_hamiltonian        plane waves (T), constant V shift, Hermiticity,
 *                         L.S spectrum (+1/2 x4, -1 x2), dense brute force
 *   or_filter_cycle       Newton filter on an operator with known spectrum
 *   or_calc_sigma_E       sigma = 0 on eigenstates; = ||H psi - E psi||
 *   the whole chain       CsPbI3 reference eval.dat (GPU end-to-end test)
 *
 * Parity unpinned: the exact rounding of the reference's FFTW transforms (no
 * reference test holds transformed grids; only the composed eigenvalues pin
 * them), and the Emin of get_energy_range where the reference's
 * tau -= 0.025 guard (energy.c:253-259) fires, which depends on rounding order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define OR_PIE 3.14159265358979323846
#define OR_TWOPI 6.28318530717958647692
#define OR_EPSDX 1.0e-20
#define sqr(x) ((x) * (x))

typedef struct { double re, im; } or_cplx;

/* Hamiltonian description shared with oracle.py (ctypes mirror). */
typedef struct {
  long nx, ny, nz, ngrid, nspin, nspinngrid;
  int is_complex, so, nl;
  double dv;
  const double *ksqr;      /* ngrid */
  const double *pot_local; /* nspinngrid (duplicated per spin, init.c:666) */
  long n_atoms, n_gridpts; /* n_NL_atoms, per-atom stride n_NL_gridpts */
  const long *n_gpt;       /* n_atoms: nl[] of the reference */
  const long *jxyz;        /* n_atoms*n_gridpts */
  const double *y1;        /* n_atoms*n_gridpts*3 complex */
  const double *proj;      /* n_atoms*n_gridpts*5 (SO radial, x sqrt(SO_par)) */
  const double *nl_proj;   /* n_atoms*n_gridpts*5 */
  const int *nl_sign;      /* n_atoms*5 */
  int n_so_proj;           /* 5 in the reference (phamiltonian.c:225) */
  int n_nl_proj;           /* ist->nproj */
  const double *LS;        /* 36 complex, LS[j_p*6 + j] */
} or_ham;

/* ------------------------------------------------------------------ */
/* ran2 (Numerical Recipes) -- rand.c:22-62.  Static state as in the  */
/* reference; a non-positive seed re-initialises it.                   */
/* ------------------------------------------------------------------ */
#define IM1 2147483563
#define IM2 2147483399
#define AM (1.0 / (double)IM1)
#define IMM1 (IM1 - 1)
#define IA1 40014
#define IA2 40692
#define IQ1 53668
#define IQ2 52774
#define IR1 12211
#define IR2 3791
#define NTAB 32
#define NDIV (1 + IMM1 / NTAB)
#define REPS 1.2e-7
#define RNMX (1.0 - REPS)

static long or_seed2 = 123456789, or_iy = 0, or_iv[NTAB];

double or_ran_nrc(long *rand_seed)
{
  long j, k;
  double temp;
  if (*rand_seed <= 0) {
    if (-(*rand_seed) < 1) *rand_seed = 1;
    else *rand_seed = -(*rand_seed);
    or_seed2 = (*rand_seed);
    for (j = NTAB + 7; j >= 0; j--) {
      k = (*rand_seed) / IQ1;
      *rand_seed = IA1 * (*rand_seed - k * IQ1) - k * IR1;
      if (*rand_seed < 0) *rand_seed += IM1;
      if (j < NTAB) or_iv[j] = *rand_seed;
    }
    or_iy = or_iv[0];
  }
  k = (*rand_seed) / IQ1;
  *rand_seed = IA1 * (*rand_seed - k * IQ1) - k * IR1;
  if (*rand_seed < 0) *rand_seed += IM1;
  k = or_seed2 / IQ2;
  or_seed2 = IA2 * (or_seed2 - k * IQ2) - k * IR2;
  if (or_seed2 < 0) or_seed2 += IM2;
  j = or_iy / NDIV;
  or_iy = or_iv[j] - or_seed2;
  or_iv[j] = *rand_seed;
  if (or_iy < 1) or_iy += IMM1;
  if ((temp = AM * or_iy) > RNMX) return RNMX;
  return temp;
}

/* normalize() -- norm.c:10-76 (calc_norm then divide by N). */
static double or_normalize_cplx(or_cplx *psi, long n, double dv)
{
  double norm = 0.0;
  for (long i = 0; i < n; i++) norm += (sqr(psi[i].re) + sqr(psi[i].im));
  double N = sqrt(norm * dv);
  for (long k = 0; k < n; k++) { psi[k].re /= N; psi[k].im /= N; }
  return N;
}

/* init_psi -- init.c:1256-1305. psi is ngrid complex. */
void or_init_psi(or_cplx *psi, long *rand_seed, long nx, long ny, long nz, int is_complex, double dv)
{
  long randint = *rand_seed;
  for (long jz = 0; jz < nz; jz++)
    for (long jy = 0; jy < ny; jy++) {
      long jzy = nx * (ny * jz + jy);
      for (long jx = 0; jx < nx; jx++) {
        long jxyz = jzy + jx;
        psi[jxyz].re = (-1.0 + 2.0 * or_ran_nrc(&randint));
        if (1 == is_complex) psi[jxyz].im = (-1.0 + 2.0 * or_ran_nrc(&randint));
        else psi[jxyz].im = 0.0;
      }
    }
  or_normalize_cplx(psi, nx * ny * nz, dv);
  *rand_seed = randint;
}

/* init_filter_states (random branch) -- init.c:1207-1253: start state jns at
 * psi_rank[jns*ms*stlen], each spin component drawn and normalised separately. */
void or_init_filter_states(double *psi_rank, long n_starts, long ms, long *rand_seed,
                           long nx, long ny, long nz, long nspin, int is_complex, double dv)
{
  long ngrid = nx * ny * nz, nspinngrid = nspin * ngrid;
  long cidx = is_complex + 1, stlen = cidx * nspinngrid;
  or_cplx *psi = (or_cplx *)calloc(nspinngrid, sizeof(or_cplx));
  for (long jns = 0; jns < n_starts; jns++) {
    for (long jspin = 0; jspin < nspin; jspin++)
      or_init_psi(&psi[jspin * ngrid], rand_seed, nx, ny, nz, is_complex, dv);
    long ns_block = jns * (ms * stlen);
    for (long jg = 0; jg < nspinngrid; jg++) {
      if (is_complex) {
        psi_rank[ns_block + 2 * jg] = psi[jg].re;
        psi_rank[ns_block + 2 * jg + 1] = psi[jg].im;
      } else {
        psi_rank[ns_block + jg] = psi[jg].re;
      }
    }
  }
  free(psi);
}

/* ------------------------------------------------------------------ */
/* Grid -- init.c:5-255 (non-periodic branch), calc_dot_dimension     */
/* init.c:1309-1379, init_kvec init.c:1383-1394, ksqr init.c:259-397. */
/* ------------------------------------------------------------------ */
static double or_dot_dim(const double *R, long n, int axis)
{
  double dr2 = 0.0;
  for (long i = 0; i < n - 1; i++)
    for (long j = i + 1; j < n; j++) {
      double d2 = sqr(R[3 * i + axis] - R[3 * j + axis]);
      if (d2 > dr2) dr2 = d2;
    }
  return sqrt(dr2);
}

/* io[0..2] = nx,ny,nz (in: requested, out: final); d[0..2] = dx,dy,dz;
 * out_min[3], out_dk[3]. */
void or_init_grid_params(const double *R, long natoms, long *io_n, const double *d,
                         double *out_min, double *out_max, double *out_dk)
{
  for (int ax = 0; ax < 3; ax++) {
    double hd = rint(0.5 * or_dot_dim(R, natoms, ax) + 5.0);
    double mn = -hd, mx = hd - d[ax];
    long ntmp = (long)((mx - mn) / d[ax]);
    long n = io_n[ax];
    if (ntmp > n) {
      n = ntmp;
      if (n % 2) n += 1;
    }
    double boxl;
    if (n % 2 == 0) {
      boxl = (double)n * d[ax];
      mn = -boxl / 2.0;
      mx = ((double)(n - 1) * d[ax]) / 2.0;
    } else {
      boxl = (double)(n - 1) * d[ax];
      mn = -boxl / 2.0;
      mx = boxl / 2.0;
    }
    io_n[ax] = n;
    out_min[ax] = mn;
    out_max[ax] = mx;
    out_dk[ax] = OR_TWOPI / boxl;
  }
}

void or_init_kvec(double *k, int n, double dk, double norm)
{
  int npos = (n % 2 == 0) ? n / 2 : (n - 1) / 2;
  k[0] = 0.0;
  for (int j = 1; j <= npos; j++) {
    double val = j * dk * norm;
    k[j] = val;
    k[n - j] = -val;
  }
}

/* ksqr part of build_grid_ksqr (init.c:318-349); grid axes x[],y[],z[] as
 * init.c:291-296 (running sums from xmin). */
void or_build_grid_ksqr(double *ksqr, double *gx, double *gy, double *gz, long nx, long ny, long nz,
                        const double *d, const double *mins, const double *dk, double KE_max_in)
{
  double dx, dy, dz;
  long jx, jy, jz;
  for (jx = 0, dx = mins[0]; jx < nx; jx++, dx += d[0]) gx[jx] = dx;
  for (jy = 0, dy = mins[1]; jy < ny; jy++, dy += d[1]) gy[jy] = dy;
  for (jz = 0, dz = mins[2]; jz < nz; jz++, dz += d[2]) gz[jz] = dz;
  double nx_1 = 1.0 / (double)nx, ny_1 = 1.0 / (double)ny, nz_1 = 1.0 / (double)nz;
  double ngrid_1 = nx_1 * ny_1 * nz_1;
  double *kx = calloc(nx, sizeof(double)), *ky = calloc(ny, sizeof(double)), *kz = calloc(nz, sizeof(double));
  or_init_kvec(kx, nx, dk[0], 1);
  or_init_kvec(ky, ny, dk[1], 1);
  or_init_kvec(kz, nz, dk[2], 1);
  for (jx = 1; jx < nx; jx++) kx[jx] *= 0.5 * kx[jx] * ngrid_1;
  for (jy = 1; jy < ny; jy++) ky[jy] *= 0.5 * ky[jy] * ngrid_1;
  for (jz = 1; jz < nz; jz++) kz[jz] *= 0.5 * kz[jz] * ngrid_1;
  double KE_max = KE_max_in * (ny_1 * nx_1 * nz_1);
  for (jz = 0; jz < nz; jz++)
    for (jy = 0; jy < ny; jy++) {
      long jyz = nx * (ny * jz + jy);
      for (jx = 0; jx < nx; jx++) {
        long jxyz = jyz + jx;
        ksqr[jxyz] = kx[jx] + ky[jy] + kz[jz];
        if (ksqr[jxyz] > KE_max) ksqr[jxyz] = KE_max;
      }
    }
  free(kx); free(ky); free(kz);
}

/* ------------------------------------------------------------------ */
/* Local potential -- build_local_pot init.c:481-703 (non-interpolated, */
/* no strain, no LR, no LSD) with interpolate() interpolate.c:5-47.      */
/* pot: per-type tables r[type*len + i], v[type*len + i], file_lens.     */
/* ------------------------------------------------------------------ */
static double or_interpolate(double r, double dr, const double *vr, const double *pot, long len, long n, long j)
{
  long i = (long)(r / dr);
  if (i > (n - 2)) return 0.0;
  double a = 1.0 * (pot[j * len + i + 1] - pot[j * len + i]) / (vr[j * len + i + 1] - vr[j * len + i]);
  double b = 1.0 * pot[j * len + i] - vr[j * len + i] * a;
  return a * r + b;
}

void or_build_local_pot(double *pot_local, long nspin, const double *gx, const double *gy, const double *gz,
                        long nx, long ny, long nz, const double *R, const int *atype, long natoms,
                        const double *pot_r, const double *pot_v, const long *file_lens, const double *pot_dr,
                        long max_len)
{
  long ngrid = nx * ny * nz;
  for (long jz = 0; jz < nz; jz++)
    for (long jy = 0; jy < ny; jy++) {
      long jyz = nx * (ny * jz + jy);
      for (long jx = 0; jx < nx; jx++) {
        long jxyz = jyz + jx;
        double sum = 0.0;
        for (long jat = 0; jat < natoms; jat++) {
          double dx = gx[jx] - R[3 * jat], dy = gy[jy] - R[3 * jat + 1], dz = gz[jz] - R[3 * jat + 2];
          double dist2 = dx * dx + dy * dy + dz * dz;
          double dist = sqrt(dist2);
          int t = atype[jat];
          sum += or_interpolate(dist, pot_dr[t], pot_r, pot_v, max_len, file_lens[t], t);
        }
        pot_local[jxyz] = sum;
        if (nspin == 2) pot_local[ngrid + jxyz] = sum;
      }
    }
}

/* ------------------------------------------------------------------ */
/* Projector tables on the grid -- count (init.c:351-394) and fill      */
/* (init.c:988-1068).  radial tables: so_tab[p*1024+i], nl_tab per atom  */
/* nl_tab[(a*nproj + p)*1024 + i]; sign per atom.  vr = i*dr_proj.       */
/* ------------------------------------------------------------------ */
#define OR_PROJ_LEN 1024
long or_count_nl_gridpts(const double *R, long natoms, const double *gx, const double *gy, const double *gz,
                         long nx, long ny, long nz, double R_NLcut2)
{
  long best = 0;
  for (long ja = 0; ja < natoms; ja++) {
    long jtmp = 0;
    for (long jz = 0; jz < nz; jz++) {
      double dz = gz[jz] - R[3 * ja + 2];
      for (long jy = 0; jy < ny; jy++) {
        double dy = gy[jy] - R[3 * ja + 1];
        for (long jx = 0; jx < nx; jx++) {
          double dx = gx[jx] - R[3 * ja];
          if (dx * dx + dy * dy + dz * dz < R_NLcut2) jtmp++;
        }
      }
    }
    if (jtmp > best) best = jtmp;
  }
  return best;
}

void or_init_nl_projectors(long *nl, long *jxyz_out, double *y1, double *proj, double *nl_proj, int *sign_out,
                           const double *R, long natoms, const double *gx, const double *gy, const double *gz,
                           long nx, long ny, long nz, double R_NLcut2, long stride, long nproj,
                           const double *so_tab, const double *nl_tab, const int *nl_sgn, const double *SO_par)
{
  const long N = OR_PROJ_LEN;
  double tmp1 = 0.5 * sqrt(3.0 / OR_PIE), tmp2 = 0.5 * sqrt(3.0 / OR_TWOPI);
  double dr = sqrt(R_NLcut2) / ((double)N);
  double *vr = calloc(N, sizeof(double));
  for (long i = 0; i < N; i++) vr[i] = (double)i * dr;
  double dr_proj = vr[1];
  for (long ja = 0; ja < natoms; ja++) {
    nl[ja] = 0;
    for (long p = 0; p < nproj; p++) sign_out[ja * 5 + p] = nl_sgn[ja * 5 + p];
    for (long jz = 0; jz < nz; jz++) {
      double dz = gz[jz] - R[3 * ja + 2];
      double dzeps = dz + OR_EPSDX;
      for (long jy = 0; jy < ny; jy++) {
        long jyz = nx * (ny * jz + jy);
        double dy = gy[jy] - R[3 * ja + 1];
        double dyeps = dy + OR_EPSDX;
        for (long jx = 0; jx < nx; jx++) {
          long jxyz = jyz + jx;
          double dx = gx[jx] - R[3 * ja];
          double dxeps = dx + OR_EPSDX;
          double dr2 = dx * dx + dy * dy + dz * dz;
          if (dr2 < R_NLcut2) {
            long idx = ja * stride + nl[ja];
            double *y = &y1[idx * 6];
            jxyz_out[idx] = jxyz;
            double dr_1 = 1.0 / sqrt(dx * dx + dy * dy + dzeps * dzeps);
            y[2] = tmp1 * dzeps * dr_1;  /* y1[1].re */
            y[3] = 0.0;                  /* y1[1].im */
            dr_1 = 1.0 / sqrt(dxeps * dxeps + dy * dy + dz * dz);
            y[4] = -tmp2 * dxeps * dr_1; /* y1[2].re */
            y[0] = tmp2 * dxeps * dr_1;  /* y1[0].re */
            dr_1 = 1.0 / sqrt(dx * dx + dyeps * dyeps + dz * dz);
            y[5] = -tmp2 * dyeps * dr_1; /* y1[2].im */
            y[1] = -tmp2 * dyeps * dr_1; /* y1[0].im */
            for (long p = 0; p < nproj; p++) {
              proj[idx * 5 + p] = or_interpolate(sqrt(dr2), dr_proj, vr, &so_tab[N * p], 0, N, 0);
              proj[idx * 5 + p] *= sqrt(SO_par[ja]);
              nl_proj[idx * 5 + p] = or_interpolate(sqrt(dr2), dr_proj, vr, &nl_tab[(ja * nproj + p) * N], 0, N, 0);
            }
            nl[ja]++;
          }
        }
      }
    }
  }
  free(vr);
}

/* ------------------------------------------------------------------ */
/* Leja points and Newton coefficients -- coeff.c:8-118, 227-293.       */
/* ------------------------------------------------------------------ */
double or_samp_points_ashkenazy(double *point_re, double *point_im, double min, double max, long ncheby)
{
  long j, k, nc3 = 32 * ncheby, jrnd, kmax, imfrac;
  double dsre, dsim, fkmax, fk, minim, maxim, range = 0.0, del;
  or_cplx *samp = (or_cplx *)calloc(nc3, sizeof(or_cplx));
  double *veca = (double *)calloc(nc3, sizeof(double));
  minim = range * min;
  maxim = range * max;
  imfrac = nc3 / 8;
  dsre = (max - min) / (double)(nc3 - 1 - imfrac);
  dsim = (maxim - minim) / (double)(imfrac - 1);
  for (j = 0; j < nc3; j++) samp[j].re = samp[j].im = 0.0;
  for (j = 0; j < nc3 - imfrac; j++) samp[j].re = min + (double)(j)*dsre;
  for (j = nc3 - imfrac; j < nc3; j++) samp[j].im = minim + (double)(j - nc3 + imfrac) * dsim;
  jrnd = 0;
  point_re[0] = samp[jrnd].re;
  point_im[0] = samp[jrnd].im;
  for (k = 0; k < nc3; k++) {
    del = sqr(samp[k].re - point_re[0]) + sqr(samp[k].im - point_im[0]);
    veca[k] = (del < 1.0e-10) ? -1.0e30 : log(del);
  }
  for (j = 1; j < ncheby; j++) {
    kmax = 0;
    fkmax = -2.0e30;
    for (k = 0; k < nc3; k++)
      if (veca[k] > fkmax) { fkmax = veca[k]; kmax = k; }
    point_re[j] = samp[kmax].re;
    point_im[j] = samp[kmax].im;
    for (k = 0; k < nc3; k++) {
      del = sqr(samp[k].re - point_re[j]) + sqr(samp[k].im - point_im[j]);
      if (del < 1.0e-10) veca[k] = -1.0e30;
      else veca[k] += log(del);
    }
  }
  for (fk = 1.0, j = 0; j < ncheby; j++) fk *= (sqr(point_re[j]) + sqr(point_im[j]));
  fk = sqrt(fk);
  fk = pow(fk, 1.0 / (double)(ncheby));
  free(samp);
  free(veca);
  return fk;
}

#define OR_FILT(x, dt) (sqrt(dt / OR_PIE) * exp(-sqr(x) * dt))

/* gen_newton_coeff (coeff.c:58-115) given the sample points; an is complex
 * interleaved [ie*ncheby + j]. */
void or_gen_newton_coeff(double *an, const double *samp, const double *ene_targets, long ms, long ncheby,
                         double Vmin, double dE, double dt)
{
  double scale = (2.0 - -2.0) / dE;
  for (long ie = 0; ie < ms; ie++) {
    double *a = &an[2 * ncheby * ie];
    double x = (samp[0] + 2.0) / scale + Vmin;
    a[0] = OR_FILT(x - ene_targets[ie], dt);
    a[1] = 0.0;
    x = (samp[1] + 2.0) / scale + Vmin;
    a[2] = (OR_FILT(x - ene_targets[ie], dt) - a[0]) / (samp[1] - samp[0]);
    a[3] = (-a[1]) / (samp[1] - samp[0]);
    for (long j = 2; j < ncheby; j++) {
      double res = 1.0, sumre = 0.0, sumim = 0.0;
      for (long i = 1; i < j; i++) {
        res *= (samp[j] - samp[i - 1]);
        sumre += a[2 * i] * res;
        sumim += a[2 * i + 1] * res;
      }
      res *= (samp[j] - samp[j - 1]);
      x = (samp[j] + 2.0) / scale + Vmin;
      a[2 * j] = (OR_FILT(x - ene_targets[ie], dt) - a[0] - sumre) / res;
      a[2 * j + 1] = (-a[1] - sumim) / res;
    }
  }
}

/* set_ene_targets -- init.c:401-478 (setTargets branch). */
void or_set_ene_targets(double *t, long nVB, long nCB, double VBmin, double VBmax, double CBmin, double CBmax)
{
  double del = (VBmax - VBmin) / (double)nVB;
  for (long jx = 0; jx < nVB; jx++) t[nVB - jx - 1] = VBmax - (double)(jx)*del;
  del = (CBmax - CBmin) / (double)nCB;
  for (long jx = nVB; jx < nVB + nCB; jx++) t[jx] = CBmin + (double)(jx - nVB) * del;
}

/* ------------------------------------------------------------------ */
/* FFT: unnormalised complex DFT with FFTW's sign convention            */
/* (FFTW_FORWARD = -1), x fastest.  Plain recursive mixed radix; used   */
/* in place of FFTW (hamiltonian.c:83,93; aux.c:265-266).                */
/* ------------------------------------------------------------------ */
/* tw[t] = exp(sign 2 pi i t / n), t < n, precomputed once per line length */
static void or_fft1d_rec(or_cplx *out, const or_cplx *in, long n, long stride, long N, const or_cplx *tw)
{
  /* length-n DFT of in[j*stride] using the length-N table (n divides N) */
  if (n == 1) { out[0] = in[0]; return; }
  long p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n; /* prime: direct DFT */
  long m = n / p, step = N / n;
  if (p == n) {
    for (long k = 0; k < n; k++) {
      double sre = 0.0, sim = 0.0;
      for (long j = 0; j < n; j++) {
        const or_cplx w = tw[((j * k) % n) * step];
        sre += in[j * stride].re * w.re - in[j * stride].im * w.im;
        sim += in[j * stride].re * w.im + in[j * stride].im * w.re;
      }
      out[k].re = sre;
      out[k].im = sim;
    }
    return;
  }
  /* decimation in time: p sub-sequences of length m */
  or_cplx *sub = (or_cplx *)malloc(n * sizeof(or_cplx));
  for (long r = 0; r < p; r++) or_fft1d_rec(&sub[r * m], &in[r * stride], m, stride * p, N, tw);
  for (long k = 0; k < m; k++) {
    for (long q = 0; q < p; q++) {
      long kk = k + q * m;
      double sre = 0.0, sim = 0.0;
      for (long r = 0; r < p; r++) {
        const or_cplx w = tw[((r * kk) % n) * step];
        sre += sub[r * m + k].re * w.re - sub[r * m + k].im * w.im;
        sim += sub[r * m + k].re * w.im + sub[r * m + k].im * w.re;
      }
      out[kk].re = sre;
      out[kk].im = sim;
    }
  }
  free(sub);
}

static or_cplx *or_twiddles(long n, int sign)
{
  or_cplx *tw = malloc(n * sizeof(or_cplx));
  for (long t = 0; t < n; t++) {
    double ang = sign * OR_TWOPI * (double)t / (double)n;
    tw[t].re = cos(ang);
    tw[t].im = sin(ang);
  }
  return tw;
}

void or_fft3d(or_cplx *a, long nx, long ny, long nz, int sign)
{
  long nmax = nx > ny ? nx : ny;
  if (nz > nmax) nmax = nz;
  or_cplx *line = malloc(nmax * sizeof(or_cplx)), *res = malloc(nmax * sizeof(or_cplx));
  or_cplx *twx = or_twiddles(nx, sign), *twy = or_twiddles(ny, sign), *twz = or_twiddles(nz, sign);
  for (long jz = 0; jz < nz; jz++)
    for (long jy = 0; jy < ny; jy++) {
      or_cplx *p = &a[nx * (ny * jz + jy)];
      or_fft1d_rec(res, p, nx, 1, nx, twx);
      memcpy(p, res, nx * sizeof(or_cplx));
    }
  for (long jz = 0; jz < nz; jz++)
    for (long jx = 0; jx < nx; jx++) {
      for (long jy = 0; jy < ny; jy++) line[jy] = a[nx * (ny * jz + jy) + jx];
      or_fft1d_rec(res, line, ny, 1, ny, twy);
      for (long jy = 0; jy < ny; jy++) a[nx * (ny * jz + jy) + jx] = res[jy];
    }
  for (long jy = 0; jy < ny; jy++)
    for (long jx = 0; jx < nx; jx++) {
      for (long jz = 0; jz < nz; jz++) line[jz] = a[nx * (ny * jz + jy) + jx];
      or_fft1d_rec(res, line, nz, 1, nz, twz);
      for (long jz = 0; jz < nz; jz++) a[nx * (ny * jz + jy) + jx] = res[jz];
    }
  free(line);
  free(res);
  free(twx);
  free(twy);
  free(twz);
}

/* ------------------------------------------------------------------ */
/* L.S matrix -- def_LS hamiltonian.c:312-442.                          */
/* ------------------------------------------------------------------ */
void or_def_LS(double *LS /* 36 complex */)
{
  int spin_arr[6], m_arr[6];
  or_cplx Lx[3][3], Ly[3][3], Lz[3][3], Sx[2][2], Sy[2][2], Sz[2][2];
  double sq2 = sqrt(0.5);
  memset(Lx, 0, sizeof Lx); memset(Ly, 0, sizeof Ly); memset(Lz, 0, sizeof Lz);
  memset(Sx, 0, sizeof Sx); memset(Sy, 0, sizeof Sy); memset(Sz, 0, sizeof Sz);
  for (int s = 0; s < 2; s++)
    for (int m = 0; m < 3; m++) { spin_arr[s * 3 + m] = s; m_arr[s * 3 + m] = m; }
  Lx[0][1].re = sq2; Lx[1][0].re = sq2; Lx[1][2].re = sq2; Lx[2][1].re = sq2;
  Ly[0][1].im = 1.0 * sq2; Ly[1][0].im = -1.0 * sq2; Ly[1][2].im = 1.0 * sq2; Ly[2][1].im = -1.0 * sq2;
  Lz[0][0].re = -1.0; Lz[2][2].re = 1.00;
  Sx[0][1].re = 0.5; Sx[1][0].re = 0.5;
  Sy[0][1].im = -0.5; Sy[1][0].im = 0.5;
  Sz[0][0].re = 0.5; Sz[1][1].re = -0.5;
  for (int i = 0; i < 72; i++) LS[i] = 0.0;
  for (int jp = 0; jp < 6; jp++) {
    int sp = spin_arr[jp], mp = m_arr[jp];
    for (int j = 0; j < 6; j++) {
      int s = spin_arr[j], m = m_arr[j], jt = jp * 6 + j;
      LS[2 * jt] += (Lx[m][mp].re * Sx[s][sp].re - Lx[m][mp].im * Sx[s][sp].im);
      LS[2 * jt + 1] += (Lx[m][mp].re * Sx[s][sp].im + Lx[m][mp].im * Sx[s][sp].re);
      LS[2 * jt] += (Ly[m][mp].re * Sy[s][sp].re - Ly[m][mp].im * Sy[s][sp].im);
      LS[2 * jt + 1] += (Ly[m][mp].re * Sy[s][sp].im + Ly[m][mp].im * Sy[s][sp].re);
      LS[2 * jt] += (Lz[m][mp].re * Sz[s][sp].re - Lz[m][mp].im * Sz[s][sp].im);
      LS[2 * jt + 1] += (Lz[m][mp].re * Sz[s][sp].im + Lz[m][mp].im * Sz[s][sp].re);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Hamiltonian -- p_hamiltonian phamiltonian.c:10-60, kinetic            */
/* hamiltonian.c:62-99, p_potential phamiltonian.c:65-137, SO            */
/* phamiltonian.c:140-312, NL phamiltonian.c:316-444 (k = 0 branch).     */
/* psi_out holds psi on entry and H psi on exit; psi_tmp is scratch.     */
/* ------------------------------------------------------------------ */
static void or_kinetic(or_cplx *psi_out, const double *ksqr, long nx, long ny, long nz)
{
  long n = nx * ny * nz;
  or_cplx *w = malloc(n * sizeof(or_cplx));
  memcpy(w, psi_out, n * sizeof(or_cplx));
  or_fft3d(w, nx, ny, nz, -1);
  for (long j = 0; j < n; j++) { w[j].re *= ksqr[j]; w[j].im *= ksqr[j]; }
  or_fft3d(w, nx, ny, nz, +1);
  memcpy(psi_out, w, n * sizeof(or_cplx));
  free(w);
}

static void or_so_pot(or_cplx *psi_out, const or_cplx *psi_tmp, const or_ham *h)
{
  int spin_arr[6], m_arr[6];
  for (int s = 0; s < 2; s++)
    for (int m = 0; m < 3; m++) { spin_arr[s * 3 + m] = s; m_arr[s * 3 + m] = m; }
  for (long jat = 0; jat < h->n_atoms; jat++) {
    long off = jat * h->n_gridpts, ng = h->n_gpt[jat];
    for (int ip = 0; ip < h->n_so_proj; ip++)
      for (int jp = 0; jp < 6; jp++) {
        int sp = spin_arr[jp], mp = m_arr[jp];
        double pre = 0.0, pim = 0.0;
        for (long g = 0; g < ng; g++) {
          long ri = off + g, r = h->jxyz[ri] + h->ngrid * sp;
          double prj = h->proj[ri * 5 + ip];
          double yre = h->y1[ri * 6 + 2 * mp], yim = h->y1[ri * 6 + 2 * mp + 1];
          pre += prj * (psi_tmp[r].re * yre + psi_tmp[r].im * yim);
          pim += prj * (psi_tmp[r].im * yre - psi_tmp[r].re * yim);
        }
        pre *= h->dv;
        pim *= h->dv;
        for (int j = 0; j < 6; j++) {
          int jt = jp * 6 + j;
          double lre = h->LS[2 * jt], lim = h->LS[2 * jt + 1];
          if (lre == 0.0 && lim == 0.0) continue;
          int s = spin_arr[j], m = m_arr[j];
          double plre = lre * pre - lim * pim, plim = lre * pim + lim * pre;
          for (long g = 0; g < ng; g++) {
            long ri = off + g, rp = h->jxyz[ri] + h->ngrid * s;
            double prj = h->proj[ri * 5 + ip];
            double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
            psi_out[rp].re += prj * (yre * plre - yim * plim);
            psi_out[rp].im += prj * (yre * plim + yim * plre);
          }
        }
      }
  }
}

static void or_nl_pot(or_cplx *psi_out, const or_cplx *psi_tmp, const or_ham *h)
{
  for (long jat = 0; jat < h->n_atoms; jat++) {
    long off = jat * h->n_gridpts, ng = h->n_gpt[jat];
    for (int ip = 0; ip < h->n_nl_proj; ip++)
      for (long spin = 0; spin < h->nspin; spin++)
        for (int m = 0; m < 3; m++) {
          double pre = 0.0, pim = 0.0;
          for (long g = 0; g < ng; g++) {
            long ri = off + g, r = h->jxyz[ri] + h->ngrid * spin;
            double nlp = h->nl_proj[ri * 5 + ip];
            double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
            pre += nlp * (psi_tmp[r].re * yre + psi_tmp[r].im * yim);
            pim += nlp * (psi_tmp[r].im * yre - psi_tmp[r].re * yim);
          }
          int sgn = h->nl_sign[jat * 5 + ip];
          pre *= sgn * h->dv;
          pim *= sgn * h->dv;
          for (long g = 0; g < ng; g++) {
            long ri = off + g, rp = h->jxyz[ri] + h->ngrid * spin;
            double nlp = h->nl_proj[ri * 5 + ip];
            double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
            psi_out[rp].re += nlp * (yre * pre - yim * pim);
            psi_out[rp].im += nlp * (yre * pim + yim * pre);
          }
        }
  }
}

void or_hamiltonian(or_cplx *psi_out, or_cplx *psi_tmp, const or_ham *h)
{
  memcpy(psi_tmp, psi_out, h->nspinngrid * sizeof(or_cplx));
  for (long js = 0; js < h->nspin; js++) or_kinetic(&psi_out[js * h->ngrid], h->ksqr, h->nx, h->ny, h->nz);
  if (h->so == 1) or_so_pot(psi_out, psi_tmp, h);
  if (h->nl == 1) or_nl_pot(psi_out, psi_tmp, h);
  if (h->is_complex == 1) {
    for (long j = 0; j < h->nspinngrid; j++) {
      psi_out[j].re += (h->pot_local[j] * psi_tmp[j].re);
      psi_out[j].im += (h->pot_local[j] * psi_tmp[j].im);
    }
  } else {
    for (long j = 0; j < h->nspinngrid; j++) psi_out[j].re += (h->pot_local[j] * psi_tmp[j].re);
  }
}

/* H applied to n states in the reference storage layout (stlen doubles per
 * state); the real layout returns Re(H psi) as energy_all/diag_H read it. */
void or_apply_h(const double *psi_in, double *hpsi, long nstates, const or_ham *h)
{
  long nsg = h->nspinngrid, cidx = h->is_complex + 1;
  or_cplx *a = malloc(nsg * sizeof(or_cplx)), *b = malloc(nsg * sizeof(or_cplx));
  for (long s = 0; s < nstates; s++) {
    const double *in = &psi_in[s * cidx * nsg];
    for (long j = 0; j < nsg; j++) {
      a[j].re = in[cidx * j];
      a[j].im = (cidx == 2) ? in[2 * j + 1] : 0.0;
    }
    or_hamiltonian(a, b, h);
    double *out = &hpsi[s * cidx * nsg];
    for (long j = 0; j < nsg; j++) {
      out[cidx * j] = a[j].re;
      if (cidx == 2) out[2 * j + 1] = a[j].im;
    }
  }
  free(a);
  free(b);
}

/* p_hnorm -- filter.c:375-410. */
static void or_p_hnorm(or_cplx *psi_out, or_cplx *psi_tmp, double zn, double Vmin, double dE_1, const or_ham *h)
{
  memcpy(psi_tmp, psi_out, h->nspinngrid * sizeof(or_cplx));
  or_hamiltonian(psi_out, psi_tmp, h);
  for (long j = 0; j < h->nspinngrid; j++) {
    psi_out[j].re = dE_1 * psi_out[j].re - (2.0 + zn + Vmin * dE_1) * psi_tmp[j].re;
    psi_out[j].im = dE_1 * psi_out[j].im - (2.0 + zn + Vmin * dE_1) * psi_tmp[j].im;
  }
}

/* filter_cycle -- filter.c:187-371, for start jns of psi_rank; ncheby_run <
 * ncheby truncates the series (used to keep oracle tests short). */
void or_filter_cycle(double *psi_rank, long jns, long ms, long ncheby, const double *an, const double *zn,
                     double Vmin, double dE_1, const or_ham *h)
{
  long nsg = h->nspinngrid, cidx = h->is_complex + 1, stlen = nsg * cidx;
  long ns_block = jns * (ms * stlen);
  or_cplx *psi = malloc(nsg * sizeof(or_cplx)), *phi = malloc(nsg * sizeof(or_cplx));
  for (long jg = 0; jg < nsg; jg++) {
    if (cidx == 2) { psi[jg].re = psi_rank[ns_block + 2 * jg]; psi[jg].im = psi_rank[ns_block + 2 * jg + 1]; }
    else { psi[jg].re = psi_rank[ns_block + jg]; psi[jg].im = 0.0; }
  }
  for (long jms = 0; jms < ms; jms++) {
    long jst = ns_block + jms * stlen;
    const double *a = &an[2 * ncheby * jms];
    for (long jg = 0; jg < nsg; jg++) {
      if (cidx == 2) {
        psi_rank[jst + 2 * jg] = a[0] * psi[jg].re - a[1] * psi[jg].im;
        psi_rank[jst + 2 * jg + 1] = a[0] * psi[jg].im + a[1] * psi[jg].re;
      } else {
        psi_rank[jst + jg] = a[0] * psi[jg].re - a[1] * psi[jg].im;
      }
    }
  }
  for (long jc = 1; jc < ncheby; jc++) {
    or_p_hnorm(psi, phi, zn[jc - 1], Vmin, dE_1, h);
    for (long jms = 0; jms < ms; jms++) {
      long jst = ns_block + jms * stlen;
      const double *a = &an[2 * (ncheby * jms + jc)];
      for (long jg = 0; jg < nsg; jg++) {
        if (cidx == 2) {
          psi_rank[jst + 2 * jg] += (a[0] * psi[jg].re - a[1] * psi[jg].im);
          psi_rank[jst + 2 * jg + 1] += (a[0] * psi[jg].im + a[1] * psi[jg].re);
        } else {
          psi_rank[jst + jg] += a[0] * psi[jg].re - a[1] * psi[jg].im;
        }
      }
    }
  }
  free(psi);
  free(phi);
}

/* normalize_all -- norm.c:82-180.  Returns 0, or -1 on a zero norm and -2 on
 * a NaN norm (the reference exits there). */
int or_normalize_all(double *psitot, long n_states, long nspinngrid, int is_complex, double dv)
{
  long cidx = is_complex + 1;
  for (long jns = 0; jns < n_states; jns++) {
    long js = jns * cidx * nspinngrid;
    double norm = 0.0;
    for (long jg = 0; jg < nspinngrid; jg++) {
      norm += sqr(psitot[js + cidx * jg]);
      if (is_complex) norm += sqr(psitot[js + cidx * jg + 1]);
    }
    norm *= dv;
    if (norm == 0.0) return -1;
    if (isnan(norm)) return -2;
    norm = 1.0 / sqrt(norm);
    for (long jg = 0; jg < nspinngrid; jg++) {
      psitot[js + cidx * jg] *= norm;
      if (is_complex) psitot[js + cidx * jg + 1] *= norm;
    }
  }
  return 0;
}

/* energy_all -- energy.c:61-165; calc_sigma_E -- energy.c:383-539. */
void or_energy_all(const double *psitot, long n, double *ene, const or_ham *h)
{
  long nsg = h->nspinngrid, cidx = h->is_complex + 1;
  or_cplx *psi = malloc(nsg * sizeof(or_cplx)), *phi = malloc(nsg * sizeof(or_cplx));
  for (long s = 0; s < n; s++) {
    const double *p = &psitot[s * cidx * nsg];
    for (long j = 0; j < nsg; j++) { psi[j].re = p[cidx * j]; psi[j].im = (cidx == 2) ? p[2 * j + 1] : 0.0; }
    memcpy(phi, psi, nsg * sizeof(or_cplx));
    or_hamiltonian(phi, psi, h);
    double e = 0.0;
    for (long j = 0; j < nsg; j++) e += (psi[j].re * phi[j].re + psi[j].im * phi[j].im);
    ene[s] = e * h->dv;
  }
  free(psi);
  free(phi);
}

void or_calc_sigma_E(const double *psitot, long n, double *sigma_E, double *eval_out, const or_ham *h)
{
  long nsg = h->nspinngrid, cidx = h->is_complex + 1;
  or_cplx *psi = malloc(nsg * sizeof(or_cplx)), *phi = malloc(nsg * sizeof(or_cplx));
  for (long s = 0; s < n; s++) {
    const double *p = &psitot[s * cidx * nsg];
    for (long j = 0; j < nsg; j++) { psi[j].re = p[cidx * j]; psi[j].im = (cidx == 2) ? p[2 * j + 1] : 0.0; }
    memcpy(phi, psi, nsg * sizeof(or_cplx));
    or_hamiltonian(phi, psi, h);
    double ev = 0.0;
    for (long j = 0; j < nsg; j++) {
      ev += p[cidx * j] * phi[j].re;
      if (cidx == 2) ev += p[2 * j + 1] * phi[j].im;
    }
    ev *= h->dv;
    memcpy(psi, phi, nsg * sizeof(or_cplx));
    or_hamiltonian(phi, psi, h);
    double ev2 = 0.0;
    for (long j = 0; j < nsg; j++) {
      ev2 += p[cidx * j] * phi[j].re;
      if (cidx == 2) ev2 += p[2 * j + 1] * phi[j].im;
    }
    ev2 *= h->dv;
    ev2 -= sqr(ev);
    sigma_E[s] = sqrt(fabs(ev2));
    if (eval_out) eval_out[s] = ev;
  }
  free(psi);
  free(phi);
}

/* get_energy_range iterative branch -- energy.c:222-300, then the range
 * inflation energy.c:343-369.  energy() (energy.c:6-57) leaves H.psi in its
 * second argument.  out[0..5] = Emin_raw, Emax_raw, Vmin (inflated Emin), dE,
 * iterations for Emin, iterations for Emax.  nl_lambda_min/vloc_min feed the
 * NL-only trimmed inflation (energy.c:343-356); pass 0 to use the blanket one. */
static double or_energy(or_cplx *psi, or_cplx *phi, const or_ham *h)
{
  memcpy(phi, psi, h->nspinngrid * sizeof(or_cplx));
  or_hamiltonian(phi, psi, h);
  double ene = 0.0;
  for (long i = 0; i < h->nspinngrid; i++) ene += (psi[i].re * phi[i].re + psi[i].im * phi[i].im);
  return ene * h->dv;
}

void or_get_energy_range(const or_ham *h, long max_iter, double nl_lambda_min, double vloc_min, double *out)
{
  long nsg = h->nspinngrid, ng = h->ngrid;
  or_cplx *psi = calloc(nsg, sizeof(or_cplx)), *phi = calloc(nsg, sizeof(or_cplx));
  long rand_seed = -874917403;
  double tau = 0.05, ene_old, Emin, Emax;
  long i, imin;
  for (long is = 0; is < h->nspin; is++)
    or_init_psi(&phi[is * ng], &rand_seed, h->nx, h->ny, h->nz, h->is_complex, h->dv);
  Emin = (ene_old = 0.0) + 10.0;
  for (i = 0; (fabs((Emin - ene_old) / Emin) > 1.0e-6) && (i < max_iter); i++) {
    memcpy(psi, phi, nsg * sizeof(or_cplx));
    or_hamiltonian(phi, psi, h);
    for (long j = 0; j < nsg; j++) {
      phi[j].re = psi[j].re - tau * phi[j].re;
      phi[j].im = psi[j].im - tau * phi[j].im;
    }
    or_normalize_cplx(phi, nsg, h->dv);
    ene_old = Emin;
    Emin = or_energy(phi, psi, h);
    if ((i > 5) && (Emin - ene_old) > 0) tau -= 0.025;
  }
  imin = i;
  for (long is = 0; is < h->nspin; is++)
    or_init_psi(&psi[is * ng], &rand_seed, h->nx, h->ny, h->nz, h->is_complex, h->dv);
  Emax = (ene_old = 0.0) + 0.1;
  for (i = 0; (fabs((Emax - ene_old) / Emax) > 1.0e-6) && (i < max_iter); i++) {
    memcpy(phi, psi, nsg * sizeof(or_cplx));
    or_hamiltonian(psi, phi, h);
    or_normalize_cplx(psi, nsg, h->dv);
    ene_old = Emax;
    Emax = or_energy(psi, phi, h);
  }
  out[0] = Emin;
  out[1] = Emax;
  if ((1 == h->nl) && (0 == h->so) && (nl_lambda_min < 0.0)) {
    Emax *= 1.05;
    Emin -= 0.1 * fabs(Emin);
    double weyl_min = 1.05 * (vloc_min + nl_lambda_min);
    if (Emin > weyl_min) Emin = weyl_min;
  } else {
    Emax *= 1.2;
    Emin -= 0.2 * fabs(Emin);
  }
  out[2] = Emin;
  out[3] = Emax - Emin;
  out[4] = (double)imin;
  out[5] = (double)i;
  free(psi);
  free(phi);
}

/* ------------------------------------------------------------------ */
/* Threaded Newton step, for the CPU baseline of the benchmark only     */
/* (tests pin it against the serial functions above).  The reference    */
/* threads p_potential with OpenMP over atoms and adds each atom's      */
/* projector expansion with atomics (phamiltonian.c:288-303,            */
/* 420-435); the same is done here.  The kinetic FFT pair is left to    */
/* the caller (scipy.fft with worker threads in cpu_newton.py).         */
/* ------------------------------------------------------------------ */
#ifdef _OPENMP
#include <omp.h>
#endif

/* SO + NL + V_loc (p_potential, phamiltonian.c:65-137): psi_out holds the
 * kinetic part T psi on entry and H psi on exit; psi_tmp is psi. */
void or_potential_omp(or_cplx *psi_out, const or_cplx *psi_tmp, const or_ham *h, int nthreads)
{
  int spin_arr[6], m_arr[6];
  for (int s = 0; s < 2; s++)
    for (int m = 0; m < 3; m++) { spin_arr[s * 3 + m] = s; m_arr[s * 3 + m] = m; }
  if (h->so == 1) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (long jat = 0; jat < h->n_atoms; jat++) {
      long off = jat * h->n_gridpts, ng = h->n_gpt[jat];
      for (int ip = 0; ip < h->n_so_proj; ip++)
        for (int jp = 0; jp < 6; jp++) {
          int sp = spin_arr[jp], mp = m_arr[jp];
          double pre = 0.0, pim = 0.0;
          for (long g = 0; g < ng; g++) {
            long ri = off + g, r = h->jxyz[ri] + h->ngrid * sp;
            double prj = h->proj[ri * 5 + ip];
            double yre = h->y1[ri * 6 + 2 * mp], yim = h->y1[ri * 6 + 2 * mp + 1];
            pre += prj * (psi_tmp[r].re * yre + psi_tmp[r].im * yim);
            pim += prj * (psi_tmp[r].im * yre - psi_tmp[r].re * yim);
          }
          pre *= h->dv;
          pim *= h->dv;
          for (int j = 0; j < 6; j++) {
            int jt = jp * 6 + j;
            double lre = h->LS[2 * jt], lim = h->LS[2 * jt + 1];
            if (lre == 0.0 && lim == 0.0) continue;
            int s = spin_arr[j], m = m_arr[j];
            double plre = lre * pre - lim * pim, plim = lre * pim + lim * pre;
            for (long g = 0; g < ng; g++) {
              long ri = off + g, rp = h->jxyz[ri] + h->ngrid * s;
              double prj = h->proj[ri * 5 + ip];
              double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
              double vre = prj * (yre * plre - yim * plim), vim = prj * (yre * plim + yim * plre);
#pragma omp atomic
              psi_out[rp].re += vre;
#pragma omp atomic
              psi_out[rp].im += vim;
            }
          }
        }
    }
  }
  if (h->nl == 1) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (long jat = 0; jat < h->n_atoms; jat++) {
      long off = jat * h->n_gridpts, ng = h->n_gpt[jat];
      for (int ip = 0; ip < h->n_nl_proj; ip++)
        for (long spin = 0; spin < h->nspin; spin++)
          for (int m = 0; m < 3; m++) {
            double pre = 0.0, pim = 0.0;
            for (long g = 0; g < ng; g++) {
              long ri = off + g, r = h->jxyz[ri] + h->ngrid * spin;
              double nlp = h->nl_proj[ri * 5 + ip];
              double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
              pre += nlp * (psi_tmp[r].re * yre + psi_tmp[r].im * yim);
              pim += nlp * (psi_tmp[r].im * yre - psi_tmp[r].re * yim);
            }
            int sgn = h->nl_sign[jat * 5 + ip];
            pre *= sgn * h->dv;
            pim *= sgn * h->dv;
            for (long g = 0; g < ng; g++) {
              long ri = off + g, rp = h->jxyz[ri] + h->ngrid * spin;
              double nlp = h->nl_proj[ri * 5 + ip];
              double yre = h->y1[ri * 6 + 2 * m], yim = h->y1[ri * 6 + 2 * m + 1];
              double vre = nlp * (yre * pre - yim * pim), vim = nlp * (yre * pim + yim * pre);
#pragma omp atomic
              psi_out[rp].re += vre;
#pragma omp atomic
              psi_out[rp].im += vim;
            }
          }
    }
  }
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (long j = 0; j < h->nspinngrid; j++) {
    psi_out[j].re += (h->pot_local[j] * psi_tmp[j].re);
    psi_out[j].im += (h->pot_local[j] * psi_tmp[j].im);
  }
}

/* One Newton step after the H application: p_hnorm's combine
 * (filter.c:402-407), psi <- dE_1 hpsi - (2 + zn + Vmin dE_1) psi, then the
 * accumulation into every target (filter.c:318-359):
 * block[jms*stlen ...] += an[jms, jc] psi.  Complex layout (stlen = 2 nsg). */
void or_hnorm_acc_omp(or_cplx *psi, const or_cplx *hpsi, double zn, double Vmin, double dE_1, double *block,
                      long ms, long nsg, const double *an, long ncheby, long jc, int nthreads)
{
  const double c0 = 2.0 + zn + Vmin * dE_1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (long j = 0; j < nsg; j++) {
    double re = dE_1 * hpsi[j].re - c0 * psi[j].re;
    double im = dE_1 * hpsi[j].im - c0 * psi[j].im;
    psi[j].re = re;
    psi[j].im = im;
    for (long jms = 0; jms < ms; jms++) {
      const double *a = &an[2 * (ncheby * jms + jc)];
      double *b = &block[jms * 2 * nsg + 2 * j];
      b[0] += (a[0] * re - a[1] * im);
      b[1] += (a[0] * im + a[1] * re);
    }
  }
}
