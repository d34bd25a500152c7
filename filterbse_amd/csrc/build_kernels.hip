// build_kernels.hip -- the input side of the path on the GPU: the local
// potential on the grid (build_local_pot, init.c:587-669) and the projector
// tables (the count of init.c:351-394 and init_NL_projectors,
// init.c:988-1068).  Both are O(grid points x atoms) scans in the reference
// (the cutoff test of build_local_pot is commented out, init.c:615-620).
//
// Results are bit-identical to the reference arithmetic: every kernel
// evaluates the same IEEE operations in the same order (per grid point the
// atoms are summed in index order; per atom the grid is scanned z, y, x), with
// floating-point contraction off so that no multiply-add is fused.
#include "fbse_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fbse {

constexpr int kBuildThreads = 256;

// interpolate() of the reference (interpolate.c:5-47 as restated by the
// oracle): linear in the table row, 0 beyond its last interval
__device__ __forceinline__ double interp_table(double r, double dr, const double *vr, const double *pot, long n) {
#pragma clang fp contract(off)
  const long i = (long)(r / dr);
  if (i > n - 2) return 0.0;
  const double a = 1.0 * (pot[i + 1] - pot[i]) / (vr[i + 1] - vr[i]);
  const double b = 1.0 * pot[i] - vr[i] * a;
  return a * r + b;
}

__global__ __launch_bounds__(kBuildThreads) void local_pot_kernel(const double *__restrict__ gx,
                                                                  const double *__restrict__ gy,
                                                                  const double *__restrict__ gz, long nx, long ny,
                                                                  long nz, const double4 *__restrict__ atoms,
                                                                  long natoms, const double *__restrict__ pot_r,
                                                                  const double *__restrict__ pot_v,
                                                                  const long *__restrict__ lens,
                                                                  const double *__restrict__ dr, long max_len,
                                                                  long nspin, double *__restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double4 s_at[kBuildThreads];
  const long ng = nx * ny * nz;
  const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double px = 0.0, py = 0.0, pz = 0.0;
  if (g < ng) {
    const long jx = g % nx, jy = (g / nx) % ny, jz = g / (nx * ny);
    px = gx[jx];
    py = gy[jy];
    pz = gz[jz];
  }
  double sum = 0.0;
  for (long a0 = 0; a0 < natoms; a0 += kBuildThreads) {
    const int na = (int)((natoms - a0) < kBuildThreads ? (natoms - a0) : kBuildThreads);
    __syncthreads();
    if (threadIdx.x < na) s_at[threadIdx.x] = atoms[a0 + threadIdx.x];
    __syncthreads();
    if (g < ng) {
      for (int i = 0; i < na; ++i) {
        const double4 at = s_at[i];
        const double dx = px - at.x, dy = py - at.y, dz = pz - at.z;
        const double dist2 = dx * dx + dy * dy + dz * dz;
        const double dist = sqrt(dist2);
        const long t = (long)at.w;
        sum += interp_table(dist, dr[t], pot_r + t * max_len, pot_v + t * max_len, lens[t]);
      }
    }
  }
  if (g < ng) {
    out[g] = sum;
    if (nspin == 2) out[ng + g] = sum;
  }
}

// Per atom (one work-group), the grid points of its box in the reference's
// z, y, x order; those with r^2 < R_NLcut2 are counted (fill = 0) or written
// in order at slot atom*stride + k (fill = 1).
struct NlArgs {
  const double *gx, *gy, *gz;
  long nx, ny, nz;
  const double *R;
  const int *box;  // per atom jx0, jx1, jy0, jy1, jz0, jz1 (inclusive)
  double rcut2;
  long stride, nproj;
  const double *so_tab, *nl_tab, *so_par;
  long *count;
  long *jxyz;
  double *y1, *proj, *nl_proj;
};

constexpr long kProjLen = 1024;  // radial table rows (OR_PROJ_LEN / the reference's N)

template <bool FILL>
__global__ __launch_bounds__(kBuildThreads) void nl_scan_kernel(NlArgs A) {
#pragma clang fp contract(off)
  __shared__ int s_scan[kBuildThreads];
  const long ja = blockIdx.x;
  const int *b = A.box + ja * 6;
  const long bx = b[1] - b[0] + 1, by = b[3] - b[2] + 1, bz = b[5] - b[4] + 1;
  const long nbox = (bx > 0 && by > 0 && bz > 0) ? bx * by * bz : 0;
  const double Rx = A.R[3 * ja], Ry = A.R[3 * ja + 1], Rz = A.R[3 * ja + 2];
  const double pie = 3.14159265358979323846, twopi = 6.28318530717958647692;
  const double tmp1 = 0.5 * sqrt(3.0 / pie), tmp2 = 0.5 * sqrt(3.0 / twopi);
  const double eps = 1.0e-20;  // EPSDX, the y1 offsets of init.c:1028-1038 (fd.h:542)
  const double drp = sqrt(A.rcut2) / ((double)kProjLen);
  long base = 0;
  for (long c0 = 0; c0 < nbox; c0 += kBuildThreads) {
    const long k = c0 + threadIdx.x;
    int hit = 0;
    long jx = 0, jy = 0, jz = 0;
    double dx = 0.0, dy = 0.0, dz = 0.0, dr2 = 0.0;
    if (k < nbox) {
      jx = b[0] + k % bx;
      jy = b[2] + (k / bx) % by;
      jz = b[4] + k / (bx * by);
      dz = A.gz[jz] - Rz;
      dy = A.gy[jy] - Ry;
      dx = A.gx[jx] - Rx;
      dr2 = dx * dx + dy * dy + dz * dz;
      hit = dr2 < A.rcut2;
    }
    // inclusive scan of the hits (Hillis-Steele in LDS)
    s_scan[threadIdx.x] = hit;
    __syncthreads();
    for (int o = 1; o < kBuildThreads; o <<= 1) {
      const int v = threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0;
      __syncthreads();
      s_scan[threadIdx.x] += v;
      __syncthreads();
    }
    const long pos = base + s_scan[threadIdx.x] - hit;
    const long total = s_scan[kBuildThreads - 1];
    if (FILL && hit) {
      const long idx = ja * A.stride + pos;
      A.jxyz[idx] = A.nx * (A.ny * jz + jy) + jx;
      const double dzeps = dz + eps, dyeps = dy + eps, dxeps = dx + eps;
      double *y = A.y1 + idx * 6;
      double dr_1 = 1.0 / sqrt(dx * dx + dy * dy + dzeps * dzeps);
      y[2] = tmp1 * dzeps * dr_1;
      y[3] = 0.0;
      dr_1 = 1.0 / sqrt(dxeps * dxeps + dy * dy + dz * dz);
      y[4] = -tmp2 * dxeps * dr_1;
      y[0] = tmp2 * dxeps * dr_1;
      dr_1 = 1.0 / sqrt(dx * dx + dyeps * dyeps + dz * dz);
      y[5] = -tmp2 * dyeps * dr_1;
      y[1] = -tmp2 * dyeps * dr_1;
      const double r = sqrt(dr2);
      const double sp = sqrt(A.so_par[ja]);
      // radial grid vr[i] = i * drp, as a table would hold it
      for (long p = 0; p < A.nproj; ++p) {
        const double *so = A.so_tab + kProjLen * p;
        const double *nl = A.nl_tab + (ja * A.nproj + p) * kProjLen;
        const long i = (long)(r / drp);
        double vs = 0.0, vn = 0.0;
        if (!(i > kProjLen - 2)) {
          const double v0 = (double)i * drp, v1 = (double)(i + 1) * drp;
          double a = 1.0 * (so[i + 1] - so[i]) / (v1 - v0);
          double bb = 1.0 * so[i] - v0 * a;
          vs = a * r + bb;
          a = 1.0 * (nl[i + 1] - nl[i]) / (v1 - v0);
          bb = 1.0 * nl[i] - v0 * a;
          vn = a * r + bb;
        }
        A.proj[idx * 5 + p] = vs * sp;
        A.nl_proj[idx * 5 + p] = vn;
      }
    }
    base += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) A.count[ja] = base;
}

}  // namespace fbse

using namespace fbse;

namespace {
struct DevBufs {
  std::vector<void *> p;
  ~DevBufs() {
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
  template <typename T, typename H> hipError_t up(T **d, const H *h, size_t n) {
    static_assert(sizeof(T) == sizeof(H), "element size");
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    p.push_back(q);
    *d = (T *)q;
    return n ? hipMemcpy(q, h, n * sizeof(T), hipMemcpyHostToDevice) : hipSuccess;
  }
  template <typename T> hipError_t alloc(T **d, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    p.push_back(q);
    *d = (T *)q;
    return hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
  }
};
#define BUILD_HIP(call)                                  \
  do {                                                   \
    if ((call) != hipSuccess) return FBSE_ERR_HIP;       \
  } while (0)

// per axis: grid indices j with (g[j] - c)^2 < rcut2, as [lo, hi] (hi < lo: none)
void axis_box(const double *g, long n, double c, double rcut2, int *lo, int *hi) {
  *lo = 0;
  *hi = -1;
  bool found = false;
  for (long j = 0; j < n; ++j) {
    const double d = g[j] - c;
    if (d * d < rcut2) {
      if (!found) *lo = (int)j;
      *hi = (int)j;
      found = true;
    }
  }
  if (!found) {
    *lo = 0;
    *hi = -1;
  }
}

int nl_setup(int device, const double *R, long natoms, const double *gx, const double *gy, const double *gz,
             long nx, long ny, long nz, double rcut2, DevBufs &B, NlArgs &A) {
  if (device >= 0) BUILD_HIP(hipSetDevice(device));
  std::vector<int> box(natoms * 6);
  for (long a = 0; a < natoms; ++a) {
    axis_box(gx, nx, R[3 * a], rcut2, &box[6 * a], &box[6 * a + 1]);
    axis_box(gy, ny, R[3 * a + 1], rcut2, &box[6 * a + 2], &box[6 * a + 3]);
    axis_box(gz, nz, R[3 * a + 2], rcut2, &box[6 * a + 4], &box[6 * a + 5]);
  }
  A = NlArgs{};
  BUILD_HIP(B.up(&A.gx, gx, nx));
  BUILD_HIP(B.up(&A.gy, gy, ny));
  BUILD_HIP(B.up(&A.gz, gz, nz));
  BUILD_HIP(B.up(&A.R, R, natoms * 3));
  BUILD_HIP(B.up(&A.box, box.data(), box.size()));
  BUILD_HIP(B.alloc(&A.count, natoms));
  A.nx = nx;
  A.ny = ny;
  A.nz = nz;
  A.rcut2 = rcut2;
  return FBSE_OK;
}
}  // namespace

extern "C" int fbse_build_local_pot(int device, long nspin, const double *gx, const double *gy, const double *gz,
                                    long nx, long ny, long nz, const double *R, const int *atype, long natoms,
                                    const double *pot_r, const double *pot_v, const long *file_lens,
                                    const double *pot_dr, long n_types, long max_len, double *pot_local) {
  if (!gx || !gy || !gz || !R || !atype || natoms < 0 || !pot_r || !pot_v || !file_lens || !pot_dr ||
      n_types < 1 || max_len < 2 || !pot_local || nx < 1 || ny < 1 || nz < 1 || (nspin != 1 && nspin != 2))
    return FBSE_ERR_ARG;
  for (long a = 0; a < natoms; ++a)
    if (atype[a] < 0 || atype[a] >= n_types) return FBSE_ERR_ARG;
  if (device >= 0) BUILD_HIP(hipSetDevice(device));
  const long ng = nx * ny * nz;
  std::vector<double4> at(natoms);
  for (long a = 0; a < natoms; ++a) at[a] = make_double4(R[3 * a], R[3 * a + 1], R[3 * a + 2], (double)atype[a]);
  DevBufs B;
  double *dgx, *dgy, *dgz, *dpr, *dpv, *ddr, *dout;
  double4 *dat;
  long *dlen;
  BUILD_HIP(B.up(&dgx, gx, nx));
  BUILD_HIP(B.up(&dgy, gy, ny));
  BUILD_HIP(B.up(&dgz, gz, nz));
  BUILD_HIP(B.up(&dat, at.data(), natoms));
  BUILD_HIP(B.up(&dpr, pot_r, n_types * max_len));
  BUILD_HIP(B.up(&dpv, pot_v, n_types * max_len));
  BUILD_HIP(B.up(&dlen, file_lens, n_types));
  BUILD_HIP(B.up(&ddr, pot_dr, n_types));
  BUILD_HIP(B.alloc(&dout, ng * nspin));
  hipLaunchKernelGGL(local_pot_kernel, dim3((unsigned)((ng + kBuildThreads - 1) / kBuildThreads)),
                     dim3(kBuildThreads), 0, 0, dgx, dgy, dgz, nx, ny, nz, dat, natoms, dpr, dpv, dlen, ddr,
                     max_len, nspin, dout);
  BUILD_HIP(hipGetLastError());
  BUILD_HIP(hipMemcpy(pot_local, dout, ng * nspin * sizeof(double), hipMemcpyDeviceToHost));
  return FBSE_OK;
}

extern "C" long fbse_count_nl_gridpts(int device, const double *R, long natoms, const double *gx, const double *gy,
                                      const double *gz, long nx, long ny, long nz, double R_NLcut2) {
  if (!R || natoms < 0 || !gx || !gy || !gz || nx < 1 || ny < 1 || nz < 1 || !(R_NLcut2 > 0.0)) return FBSE_ERR_ARG;
  if (natoms == 0) return 0;
  DevBufs B;
  NlArgs A;
  int rc = nl_setup(device, R, natoms, gx, gy, gz, nx, ny, nz, R_NLcut2, B, A);
  if (rc) return rc;
  hipLaunchKernelGGL((nl_scan_kernel<false>), dim3((unsigned)natoms), dim3(kBuildThreads), 0, 0, A);
  BUILD_HIP(hipGetLastError());
  std::vector<long> cnt(natoms);
  BUILD_HIP(hipMemcpy(cnt.data(), A.count, natoms * sizeof(long), hipMemcpyDeviceToHost));
  return *std::max_element(cnt.begin(), cnt.end());
}

extern "C" int fbse_init_nl_projectors(int device, long *nl, long *jxyz, double *y1, double *proj, double *nl_proj,
                                       int *sign, const double *R, long natoms, const double *gx, const double *gy,
                                       const double *gz, long nx, long ny, long nz, double R_NLcut2, long stride,
                                       long nproj, const double *so_tab, const double *nl_tab, const int *nl_sgn,
                                       const double *SO_par) {
  if (!nl || !jxyz || !y1 || !proj || !nl_proj || !sign || !R || natoms < 0 || !gx || !gy || !gz || nx < 1 ||
      ny < 1 || nz < 1 || !(R_NLcut2 > 0.0) || stride < 0 || nproj < 0 || nproj > 5 || !so_tab || !nl_tab ||
      !nl_sgn || !SO_par)
    return FBSE_ERR_ARG;
  if (natoms == 0) return FBSE_OK;
  for (long a = 0; a < natoms; ++a)
    for (long p = 0; p < nproj; ++p) sign[a * 5 + p] = nl_sgn[a * 5 + p];
  DevBufs B;
  NlArgs A;
  int rc = nl_setup(device, R, natoms, gx, gy, gz, nx, ny, nz, R_NLcut2, B, A);
  if (rc) return rc;
  A.stride = stride;
  A.nproj = nproj;
  const size_t np = (size_t)natoms * stride;
  BUILD_HIP(B.up(&A.so_tab, so_tab, (size_t)kProjLen * std::max(nproj, 1L)));
  BUILD_HIP(B.up(&A.nl_tab, nl_tab, (size_t)natoms * std::max(nproj, 1L) * kProjLen));
  BUILD_HIP(B.up(&A.so_par, SO_par, natoms));
  BUILD_HIP(B.alloc(&A.jxyz, np));
  BUILD_HIP(B.alloc(&A.y1, np * 6));
  BUILD_HIP(B.alloc(&A.proj, np * 5));
  BUILD_HIP(B.alloc(&A.nl_proj, np * 5));
  // the counts first: an atom with more points than stride is an error, not an overflow
  hipLaunchKernelGGL((nl_scan_kernel<false>), dim3((unsigned)natoms), dim3(kBuildThreads), 0, 0, A);
  BUILD_HIP(hipGetLastError());
  std::vector<long> cnt(natoms);
  BUILD_HIP(hipMemcpy(cnt.data(), A.count, natoms * sizeof(long), hipMemcpyDeviceToHost));
  for (long a = 0; a < natoms; ++a)
    if (cnt[a] > stride) return FBSE_ERR_ARG;
  hipLaunchKernelGGL((nl_scan_kernel<true>), dim3((unsigned)natoms), dim3(kBuildThreads), 0, 0, A);
  BUILD_HIP(hipGetLastError());
  BUILD_HIP(hipMemcpy(nl, A.count, natoms * sizeof(long), hipMemcpyDeviceToHost));
  BUILD_HIP(hipMemcpy(jxyz, A.jxyz, np * sizeof(long), hipMemcpyDeviceToHost));
  BUILD_HIP(hipMemcpy(y1, A.y1, np * 6 * sizeof(double), hipMemcpyDeviceToHost));
  BUILD_HIP(hipMemcpy(proj, A.proj, np * 5 * sizeof(double), hipMemcpyDeviceToHost));
  BUILD_HIP(hipMemcpy(nl_proj, A.nl_proj, np * 5 * sizeof(double), hipMemcpyDeviceToHost));
  return FBSE_OK;
}
