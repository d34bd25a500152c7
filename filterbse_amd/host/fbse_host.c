/*
 * fbse_host.c -- host-side inputs of the filter-diagonalization path
 * (declared in include/fbse_host.h).  These are the host computations the
 * reference performs before handing states to the filter; they stay on the
 * CPU here too (they are O(N) or O(M^2) once per run), with the Leja search
 * and the coefficient recursion spread over OpenMP threads.  Integer and index
 * results are bit-identical to the single-threaded reference order: every
 * candidate's log-sum is accumulated in the same order, and the argmax keeps
 * the lowest index among equal maxima (coeff.c:266, first strict '>').
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fbse_host.h"

#define FB_PI 3.14159265358979323846
#define FB_TWOPI 6.28318530717958647692

/* ---------------- ran2: L'Ecuyer pair + Bays-Durham shuffle ---------------- */
enum { R_NTAB = 32 };
static const long R_M1 = 2147483563, R_M2 = 2147483399, R_A1 = 40014, R_A2 = 40692;
static const long R_Q1 = 53668, R_Q2 = 52774, R_R1 = 12211, R_R2 = 3791;
static long r_s2 = 123456789, r_y = 0, r_tab[R_NTAB];

static inline long schrage(long s, long a, long q, long r, long m)
{
  long k = s / q;
  s = a * (s - k * q) - k * r;
  return s < 0 ? s + m : s;
}

double fbse_ran_nrc(long *seed)
{
  const long mm1 = R_M1 - 1, ndiv = 1 + mm1 / R_NTAB;
  if (*seed <= 0) {
    *seed = (-(*seed) < 1) ? 1 : -(*seed);
    r_s2 = *seed;
    for (int j = R_NTAB + 7; j >= 0; j--) {
      *seed = schrage(*seed, R_A1, R_Q1, R_R1, R_M1);
      if (j < R_NTAB) r_tab[j] = *seed;
    }
    r_y = r_tab[0];
  }
  *seed = schrage(*seed, R_A1, R_Q1, R_R1, R_M1);
  r_s2 = schrage(r_s2, R_A2, R_Q2, R_R2, R_M2);
  long j = r_y / ndiv;
  r_y = r_tab[j] - r_s2;
  r_tab[j] = *seed;
  if (r_y < 1) r_y += mm1;
  double t = (1.0 / (double)R_M1) * r_y;
  return t > (1.0 - 1.2e-7) ? (1.0 - 1.2e-7) : t;
}

void fbse_init_psi(double *psi, long *seed, long nx, long ny, long nz, int is_complex, double dv)
{
  long s = *seed, n = nx * ny * nz;
  for (long j = 0; j < n; j++) {
    psi[2 * j] = -1.0 + 2.0 * fbse_ran_nrc(&s);
    psi[2 * j + 1] = is_complex ? -1.0 + 2.0 * fbse_ran_nrc(&s) : 0.0;
  }
  double acc = 0.0;
  for (long j = 0; j < n; j++) acc += psi[2 * j] * psi[2 * j] + psi[2 * j + 1] * psi[2 * j + 1];
  double N = sqrt(acc * dv);
  for (long j = 0; j < 2 * n; j++) psi[j] /= N;
  *seed = s;
}

void fbse_init_filter_states(double *psi_rank, long n_starts, long ms, long *seed, long nx, long ny, long nz,
                             long nspin, int is_complex, double dv)
{
  long ng = nx * ny * nz, cidx = is_complex + 1, stlen = cidx * nspin * ng;
  double *buf = (double *)malloc(2 * ng * sizeof(double));
  for (long jns = 0; jns < n_starts; jns++) {
    double *dst = psi_rank + jns * ms * stlen;
    for (long sp = 0; sp < nspin; sp++) {
      fbse_init_psi(buf, seed, nx, ny, nz, is_complex, dv);
      if (is_complex) memcpy(dst + sp * 2 * ng, buf, 2 * ng * sizeof(double));
      else
        for (long g = 0; g < ng; g++) dst[sp * ng + g] = buf[2 * g];
    }
  }
  free(buf);
}

/* ---------------- grid and kinetic table ---------------- */
void fbse_init_grid_params(const double *R, long natoms, long *n, const double *d, double *mins, double *maxs,
                           double *dk)
{
  for (int ax = 0; ax < 3; ax++) {
    double ext2 = 0.0; /* largest squared pair separation along the axis */
    for (long i = 0; i + 1 < natoms; i++)
      for (long j = i + 1; j < natoms; j++) {
        double t = R[3 * i + ax] - R[3 * j + ax];
        if (t * t > ext2) ext2 = t * t;
      }
    double half = rint(0.5 * sqrt(ext2) + 5.0);
    long need = (long)(((half - d[ax]) - (-half)) / d[ax]);
    long m = n[ax];
    if (need > m) m = need + (need % 2);
    double box = (m % 2 == 0) ? (double)m * d[ax] : (double)(m - 1) * d[ax];
    mins[ax] = -box / 2.0;
    maxs[ax] = (m % 2 == 0) ? ((double)(m - 1) * d[ax]) / 2.0 : box / 2.0;
    dk[ax] = FB_TWOPI / box;
    n[ax] = m;
  }
}

static void kvec_half_sq(double *k, long n, double dk, double ngrid_1)
{
  long npos = (n % 2 == 0) ? n / 2 : (n - 1) / 2;
  for (long j = 0; j < n; j++) k[j] = 0.0;
  for (long j = 1; j <= npos; j++) {
    double v = j * dk * 1.0;
    k[j] = v;
    k[n - j] = -v;
  }
  for (long j = 1; j < n; j++) k[j] *= 0.5 * k[j] * ngrid_1;
}

void fbse_build_grid_ksqr(double *ksqr, double *gx, double *gy, double *gz, const long *n, const double *d,
                          const double *mins, const double *dk, double KE_max)
{
  const long nx = n[0], ny = n[1], nz = n[2];
  double *ax[3] = {gx, gy, gz};
  for (int a = 0; a < 3; a++) {
    double v = mins[a];
    for (long j = 0; j < n[a]; j++, v += d[a]) ax[a][j] = v;
  }
  double nx_1 = 1.0 / (double)nx, ny_1 = 1.0 / (double)ny, nz_1 = 1.0 / (double)nz;
  double ng_1 = nx_1 * ny_1 * nz_1, cap = KE_max * (ny_1 * nx_1 * nz_1);
  double *kx = malloc(nx * sizeof(double)), *ky = malloc(ny * sizeof(double)), *kz = malloc(nz * sizeof(double));
  kvec_half_sq(kx, nx, dk[0], ng_1);
  kvec_half_sq(ky, ny, dk[1], ng_1);
  kvec_half_sq(kz, nz, dk[2], ng_1);
  for (long z = 0; z < nz; z++)
    for (long y = 0; y < ny; y++) {
      double *row = ksqr + nx * (ny * z + y);
      for (long x = 0; x < nx; x++) {
        double v = kx[x] + ky[y] + kz[z];
        row[x] = v > cap ? cap : v;
      }
    }
  free(kx);
  free(ky);
  free(kz);
}

/* ---------------- Leja nodes ---------------- */
double fbse_samp_points_ashkenazy(double *zn, long ncheby, int nthreads)
{
  const double lo = -2.0, hi = 2.0;
  const long nc = 32 * ncheby, nimag = nc / 8, nreal = nc - nimag;
  const double ds = (hi - lo) / (double)(nreal - 1);
  /* candidates: nreal uniform reals, then nimag copies of 0 (the imaginary
     segment has zero length, coeff.c:234-246) */
  double *cand = (double *)malloc(nc * sizeof(double));
  double *score = (double *)malloc(nc * sizeof(double));
  for (long k = 0; k < nreal; k++) cand[k] = lo + (double)k * ds;
  for (long k = nreal; k < nc; k++) cand[k] = 0.0;
  if (nthreads < 1) nthreads = 1;
  zn[0] = cand[0];
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long k = 0; k < nc; k++) {
    double t = cand[k] - zn[0], d2 = t * t;
    score[k] = d2 < 1.0e-10 ? -1.0e30 : log(d2);
  }
  for (long j = 1; j < ncheby; j++) {
    long best = 0;
    double bval = -2.0e30;
#pragma omp parallel num_threads(nthreads)
    {
      long lb = 0;
      double lv = -2.0e30;
      int found = 0;
#pragma omp for schedule(static) nowait
      for (long k = 0; k < nc; k++)
        if (score[k] > lv) {
          lv = score[k];
          lb = k;
          found = 1;
        }
#pragma omp critical
      if (found && (lv > bval || (lv == bval && lb < best))) {
        bval = lv;
        best = lb;
      }
    }
    zn[j] = cand[best];
    const double zj = zn[j];
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long k = 0; k < nc; k++) {
      double t = cand[k] - zj, d2 = t * t;
      if (d2 < 1.0e-10) score[k] = -1.0e30;
      else score[k] += log(d2);
    }
  }
  double fk = 1.0;
  for (long j = 0; j < ncheby; j++) fk *= zn[j] * zn[j];
  fk = pow(sqrt(fk), 1.0 / (double)ncheby);
  free(cand);
  free(score);
  return fk;
}

/* ---------------- Newton divided differences ---------------- */
void fbse_gen_newton_coeff(double *an, const double *zn, const double *targets, long ms, long ncheby, double Vmin,
                           double dE, double dt, int nthreads)
{
  const double scale = 4.0 / dE, pre = sqrt(dt / FB_PI);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
  for (long ie = 0; ie < ms; ie++) {
    double *a = an + 2 * ncheby * ie;
    const double tg = targets[ie];
    for (long j = 0; j < ncheby; j++) {
      const double x = (zn[j] + 2.0) / scale + Vmin, u = x - tg;
      const double f = pre * exp(-(u * u) * dt);
      if (j == 0) {
        a[0] = f;
        a[1] = 0.0;
        continue;
      }
      double res = 1.0, sre = 0.0, sim = 0.0;
      for (long i = 1; i < j; i++) {
        res *= (zn[j] - zn[i - 1]);
        sre += a[2 * i] * res;
        sim += a[2 * i + 1] * res;
      }
      res *= (zn[j] - zn[j - 1]);
      a[2 * j] = (f - a[0] - sre) / res;
      a[2 * j + 1] = (-a[1] - sim) / res;
    }
  }
}

void fbse_set_ene_targets(double *t, long nVB, long nCB, double VBmin, double VBmax, double CBmin, double CBmax)
{
  double step = (VBmax - VBmin) / (double)nVB;
  for (long i = 0; i < nVB; i++) t[nVB - 1 - i] = VBmax - (double)i * step;
  step = (CBmax - CBmin) / (double)nCB;
  for (long i = 0; i < nCB; i++) t[nVB + i] = CBmin + (double)i * step;
}

void fbse_def_LS(double *LS)
{
  /* L.S = Lx Sx + Ly Sy + Lz Sz for l = 1, s = 1/2 with m ordered -1,0,+1 and
     index j = s*3 + m; element [jp*6 + j] = L[m][mp] S[s][sp] summed. */
  const double r = sqrt(0.5);
  double Lre[3][3][3] = {{{0}}}, Lim[3][3][3] = {{{0}}}, Sre[3][2][2] = {{{0}}}, Sim[3][2][2] = {{{0}}};
  Lre[0][0][1] = Lre[0][1][0] = Lre[0][1][2] = Lre[0][2][1] = r; /* Lx */
  Lim[1][0][1] = r;                                              /* Ly */
  Lim[1][1][0] = -r;
  Lim[1][1][2] = r;
  Lim[1][2][1] = -r;
  Lre[2][0][0] = -1.0; /* Lz */
  Lre[2][2][2] = 1.0;
  Sre[0][0][1] = Sre[0][1][0] = 0.5; /* Sx */
  Sim[1][0][1] = -0.5;               /* Sy */
  Sim[1][1][0] = 0.5;
  Sre[2][0][0] = 0.5; /* Sz */
  Sre[2][1][1] = -0.5;
  for (int i = 0; i < 72; i++) LS[i] = 0.0;
  for (int jp = 0; jp < 6; jp++)
    for (int j = 0; j < 6; j++) {
      int sp = jp / 3, mp = jp % 3, s = j / 3, m = j % 3;
      for (int c = 0; c < 3; c++) {
        double ar = Lre[c][m][mp], ai = Lim[c][m][mp], br = Sre[c][s][sp], bi = Sim[c][s][sp];
        LS[2 * (jp * 6 + j)] += ar * br - ai * bi;
        LS[2 * (jp * 6 + j) + 1] += ar * bi + ai * br;
      }
    }
}
