// capi.hip -- C ABI of the gfx950 filter-diagonalization library (include/fbse_gpu.h):
// context set-up, H application, the Newton filter driver, normalisation,
// energies, orthogonalisation, Rayleigh-Ritz and sigma_E.
#include "fbse_internal.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

using namespace fbse;

namespace fbse {

int set_err(fbse_ctx *c, int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

static int grow(fbse_ctx *c, void **p, size_t *cur, size_t bytes) {
  if (bytes <= *cur) return FBSE_OK;
  if (c->no_alloc)
    return set_err(c, FBSE_ERR_NOMEM, "a %zu-byte buffer was needed inside the no-allocation entry point", bytes);
  if (*p) {
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    FBSE_HIP(c, hipFree(*p));
    *p = nullptr;
    *cur = 0;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, FBSE_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  *cur = bytes;
  return FBSE_OK;
}
int ensure_scratch(fbse_ctx *c, size_t b) { return grow(c, &c->d_scratch, &c->scratch_bytes, b); }
int ensure_scratch2(fbse_ctx *c, size_t b) { return grow(c, &c->d_scratch2, &c->scratch2_bytes, b); }
int ensure_red(fbse_ctx *c, size_t b) { return grow(c, (void **)&c->d_red, &c->red_bytes, b); }

// Releases the context's work buffers (the filter's ring, FFT work space,
// partial Grams): they are allocated again on demand.
static void release_bufs(fbse_ctx *c, int first) {
  (void)hipStreamSynchronize(c->stream);
  void **bufs[] = {&c->d_scratch, &c->d_scratch2, &c->d_gram};
  size_t *sizes[] = {&c->scratch_bytes, &c->scratch2_bytes, &c->gram_bytes};
  for (int i = first; i < 3; ++i) {
    if (*bufs[i]) (void)hipFree(*bufs[i]);
    *bufs[i] = nullptr;
    *sizes[i] = 0;
  }
}
static void release_work(fbse_ctx *c) {
  release_bufs(c, 0);
  if (c->d_pp_part) (void)hipFree(c->d_pp_part);
  c->d_pp_part = nullptr;
  c->pp_part_bytes = 0;
  c->pp_ready = 0;
}
// Before a filter sizes its ring from the free memory: the second scratch
// buffer and the partial Grams that ortho, Rayleigh-Ritz or the refinement
// left behind (tens of GB on a 256^3 block) are released, so the window depth
// does not depend on what ran before (the ring's own buffer counts as free).
static size_t blas_workspace_bytes(fbse_ctx *c) {
  size_t sz = 0;
  if (rocblas_get_device_memory_size(c->blas, &sz) != rocblas_status_success) return 0;
  return sz;
}
static void release_aux(fbse_ctx *c) {
  if (c->d_scratch2 || c->d_gram) release_bufs(c, 1);
  // rocBLAS keeps the largest workspace any call asked for (the triangular
  // solves and factorisations of ortho / Rayleigh-Ritz); above 1 GiB it is
  // freed here and allocated again on demand
  if (blas_workspace_bytes(c) > (size_t(1) << 30)) {
    (void)hipStreamSynchronize(c->stream);
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
    (void)rocblas_set_device_memory_size(c->blas, 0);
#pragma clang diagnostic pop
  }
}

// hipMalloc of a large block: when the device is full, the work buffers are
// released (after a filter they hold its ring) and the allocation is retried.
static hipError_t big_malloc(fbse_ctx *c, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) == hipSuccess) return hipSuccess;
  (void)hipGetLastError();
  release_work(c);
  return hipMalloc(p, bytes);
}

static int ensure_host_red(fbse_ctx *c, size_t b) {
  if (b <= c->h_red_bytes) return FBSE_OK;
  if (c->h_red) hipHostFree(c->h_red);
  c->h_red = nullptr;
  c->h_red_bytes = 0;
  FBSE_HIP(c, hipHostMalloc((void **)&c->h_red, b, hipHostMallocDefault));
  c->h_red_bytes = b;
  return FBSE_OK;
}

void prof_begin(fbse_ctx *c, int kind, hipEvent_t *ev) {
  if (!c->profile) return;
  hipEvent_t e;
  if (!c->ev_pool.empty()) {
    e = c->ev_pool.back();
    c->ev_pool.pop_back();
  } else {
    hipEventCreate(&e);
  }
  hipEventRecord(e, c->stream);
  *ev = e;
  (void)kind;
}
void prof_end(fbse_ctx *c, int kind, hipEvent_t ev) {
  if (!c->profile || !ev) return;
  hipEvent_t e;
  if (!c->ev_pool.empty()) {
    e = c->ev_pool.back();
    c->ev_pool.pop_back();
  } else {
    hipEventCreate(&e);
  }
  hipEventRecord(e, c->stream);
  c->ev_pending.push_back({kind, {ev, e}});
  if (c->ev_pending.size() > 4096) prof_collect(c);
}
void prof_collect(fbse_ctx *c) {
  if (c->ev_pending.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto &p : c->ev_pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.second.first, p.second.second);
    c->timers[p.first].launches += 1;
    c->timers[p.first].ms += ms;
    c->timers[4].launches += 1;
    c->timers[4].ms += ms;
    c->ev_pool.push_back(p.second.first);
    c->ev_pool.push_back(p.second.second);
  }
  c->ev_pending.clear();
}

// Per-axis passes: for a length with prime factors 2, 3, 5 only, the fewest
// passes over the register butterflies (composite radices included), larger
// radices first; otherwise the schedule below (other primes by direct DFT).
static void factorize(int n, AxisPlan &pl);
static void axis_schedule(int n, AxisPlan &pl) {
  static const int rad[] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};
  std::vector<int> passes(n + 1, 1 << 20), maxr(n + 1, 0), pick(n + 1, 0);
  passes[1] = 0;
  for (int m = 2; m <= n; ++m) {
    if (n % m) continue;
    for (int r : rad) {
      if (m % r || passes[m / r] >= (1 << 20)) continue;
      const int pp = passes[m / r] + 1, mr = std::max(maxr[m / r], r);
      if (pp < passes[m] || (pp == passes[m] && mr < maxr[m])) {
        passes[m] = pp;
        maxr[m] = mr;
        pick[m] = r;
      }
    }
  }
  if (n > 1 && passes[n] >= (1 << 20)) {
    factorize(n, pl);
    return;
  }
  std::vector<int> fs;
  for (int m = n; m > 1; m /= pick[m]) fs.push_back(pick[m]);
  std::sort(fs.begin(), fs.end(), [](int a, int b) { return a > b; });
  pl.n = n;
  pl.nf = (int)std::min<size_t>(fs.size(), kMaxFactors);
  for (int i = 0; i < pl.nf; ++i) pl.f[i] = fs[i];
}

// Radix schedule: 8s first, then 4, 2, 3, 5, then any remaining primes.
static void factorize(int n, AxisPlan &pl) {
  pl.n = n;
  pl.nf = 0;
  int m = n;
  const int pref[] = {8, 4, 2, 3, 5};
  for (int r : pref)
    while (m % r == 0 && pl.nf < kMaxFactors) {
      pl.f[pl.nf++] = r;
      m /= r;
    }
  for (int p = 7; m > 1 && p <= n; p += 2)
    while (m % p == 0 && pl.nf < kMaxFactors) {
      pl.f[pl.nf++] = p;
      m /= p;
    }
}

}  // namespace fbse

// ---------------------------------------------------------------------------
// creation / destruction
// ---------------------------------------------------------------------------
static int upload(fbse_ctx *c, void **dst, const void *src, size_t bytes) {
  FBSE_HIP(c, hipMalloc(dst, bytes ? bytes : 16));
  if (bytes) FBSE_HIP(c, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return FBSE_OK;
}

template <typename T, typename S> static std::vector<T> cast_vec(const S *p, size_t n) {
  std::vector<T> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = (T)p[i];
  return v;
}

// Tables of the projections inside the spin-paired combine, built on the
// first fbse_set_combine_projection(ctx, 1) from the rank records (record e:
// pa[2e] = (y << 16) | x, pa[2e + 1] = atom, values pe[8e ..]; bucket (z, k)
// rows y .. at yoff[(z * nr + k) * (ny + 1) + y]).
static int build_pp_tables(fbse_ctx *c, long na, int nr, const std::vector<int> &pa, const std::vector<double> &pe,
                           const std::vector<int> &yoff) {
  int rc;
  // Projections inside the spin-paired combine (AxisArgs::pp_*): per tile
  // of 8 rows of one plane, its records of every rank grouped by atom
  // (atoms ascending, a group's records in point order, padded with
  // zero-weight records to a multiple of kPpRecs), kPpRecs records per lane;
  // an atom's lanes within one wavefront are one segment, whose sum is one
  // slot, and pp_span[tile] is the tile's longest segment in lanes, rounded
  // up to a power of two.  Slots are numbered atom-major, each atom's in
  // tile order.
  {
    const long tpz = c->ny / 8, ntile = c->nz * tpz;
    struct Piece {
      int atom, end_lane;
    };
    std::vector<Piece> pieces;
    std::vector<long> piece_off(ntile + 1, 0);
    std::vector<int> lane_off(ntile + 1, 0), lane_w, meta, span(ntile, 1);
    std::vector<double> val;
    std::vector<std::array<int, 3>> tmp;  // (atom, (y << 16) | x, record)
    for (long tau = 0; tau < ntile; ++tau) {
      const long z = tau / tpz, y0 = (tau % tpz) * 8;
      tmp.clear();
      for (int k = 0; k < nr; ++k) {
        const size_t b = ((size_t)z * nr + k) * (c->ny + 1);
        for (int e = yoff[b + y0]; e < yoff[b + y0 + 8]; ++e)
          tmp.push_back({pa[2 * (size_t)e + 1], pa[2 * (size_t)e], e});
      }
      std::sort(tmp.begin(), tmp.end());
      const int base = (int)lane_w.size();  // tile's first lane
      for (size_t i = 0; i < tmp.size();) {
        size_t j = i;
        while (j < tmp.size() && tmp[j][0] == tmp[i][0]) ++j;
        const int n = (int)(j - i), nlr = (n + kPpRecs - 1) / kPpRecs;
        int seg = 0;
        for (int q = 0; q < nlr; ++q) {
          const int li = (int)lane_w.size() - base;  // lane within the tile
          const bool head = q == 0 || li % 64 == 0;
          lane_w.push_back(head ? 1 : 0);
          seg = head ? 1 : seg + 1;
          while (span[tau] < seg) span[tau] *= 2;
          for (int u = 0; u < kPpRecs; ++u) {
            const size_t r = i + (size_t)std::min(q * kPpRecs + u, n - 1);
            const bool pad = q * kPpRecs + u >= n;
            const int yx = tmp[r][1];
            meta.push_back((((yx >> 16) - (int)y0) << 16) | (yx & 0xffff));
            for (int v = 0; v < 8; ++v) val.push_back(pad ? 0.0 : pe[(size_t)tmp[r][2] * 8 + v]);
          }
          // the segment ends at the group's last lane or a wavefront's last
          if (q == nlr - 1 || (li + 1) % 64 == 0) pieces.push_back({tmp[i][0], (int)lane_w.size() - 1});
        }
        i = j;
      }
      lane_off[tau + 1] = (int)lane_w.size();
      piece_off[tau + 1] = (long)pieces.size();
    }
    std::vector<int> aoff(na + 1, 0);
    for (const Piece &pc : pieces) ++aoff[pc.atom + 1];
    for (long a = 0; a < na; ++a) aoff[a + 1] += aoff[a];
    std::vector<int> seen(aoff.begin(), aoff.end() - 1);
    for (const Piece &pc : pieces) lane_w[pc.end_lane] |= (seen[pc.atom]++ + 1) << 1;
    c->pp_nslots = aoff[na];
    auto valf = cast_vec<float>(val.data(), val.size());
    if ((rc = upload(c, (void **)&c->d_pp_lane_off, lane_off.data(), lane_off.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_pp_lanes, lane_w.data(), lane_w.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_pp_span, span.data(), span.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_pp_meta, meta.data(), meta.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_pp_val, val.data(), val.size() * sizeof(double))) ||
        (rc = upload(c, (void **)&c->d_pp_val_f, valf.data(), valf.size() * sizeof(float))) ||
        (rc = upload(c, (void **)&c->d_pp_atom_off, aoff.data(), aoff.size() * sizeof(int))))
      return rc;
  }
  return FBSE_OK;
}

static int build_projectors(fbse_ctx *c, const fbse_proj_desc *pd) {
  const long na = pd->n_atoms, stride = pd->n_gridpts;
  std::vector<int> off(na), cnt(na), jx, sign(na * 5);
  long npts = 0;
  for (long a = 0; a < na; ++a) {
    off[a] = (int)npts;
    cnt[a] = (int)pd->n_gpt[a];
    if (pd->n_gpt[a] < 0 || pd->n_gpt[a] > stride)
      return set_err(c, FBSE_ERR_ARG, "atom %ld: n_gpt %ld outside [0, n_gridpts=%ld]", a, pd->n_gpt[a], stride);
    npts += pd->n_gpt[a];
  }
  if (npts > 0x7fffffffL) return set_err(c, FBSE_ERR_UNSUPPORTED, "projector table too large");
  jx.resize(npts);
  std::vector<double> proj(5 * npts, 0.0), nlp(5 * npts, 0.0), y(6 * npts, 0.0);
  for (long a = 0; a < na; ++a) {
    for (int p = 0; p < 5; ++p) sign[a * 5 + p] = pd->nl_sign ? pd->nl_sign[a * 5 + p] : 1;
    for (long g = 0; g < pd->n_gpt[a]; ++g) {
      const long src = a * stride + g, dst = off[a] + g;
      const long jxyz = pd->jxyz[src];
      if (jxyz < 0 || jxyz >= c->ngrid)
        return set_err(c, FBSE_ERR_ARG, "atom %ld point %ld: jxyz %ld outside the grid", a, g, jxyz);
      jx[dst] = (int)jxyz;
      for (int p = 0; p < 5; ++p) {
        if (pd->so && p < pd->n_so_proj) proj[p * npts + dst] = pd->proj[src * 5 + p];
        if (pd->nl && p < pd->n_nl_proj) nlp[p * npts + dst] = pd->nl_proj[src * 5 + p];
      }
      for (int m = 0; m < 3; ++m) {
        y[2 * (m * npts + dst)] = pd->y1[src * 6 + 2 * m];
        y[2 * (m * npts + dst) + 1] = pd->y1[src * 6 + 2 * m + 1];
      }
    }
  }
  double lsd[72];
  for (int i = 0; i < 72; ++i) lsd[i] = pd->LS ? pd->LS[i] : 0.0;
  int rc;
  auto &P = c->pj;
  auto &F = c->pj_f;
  P.n_atoms = F.n_atoms = (int)na;
  P.npts = F.npts = (int)npts;
  P.so = F.so = pd->so;
  P.nl = F.nl = pd->nl;
  P.n_so = F.n_so = pd->n_so_proj;
  P.n_nl = F.n_nl = pd->n_nl_proj;
  if ((rc = upload(c, (void **)&P.a_off, off.data(), na * sizeof(int)))) return rc;
  if ((rc = upload(c, (void **)&P.a_cnt, cnt.data(), na * sizeof(int)))) return rc;
  if ((rc = upload(c, (void **)&P.jxyz, jx.data(), npts * sizeof(int)))) return rc;
  if ((rc = upload(c, (void **)&P.sign, sign.data(), na * 5 * sizeof(int)))) return rc;
  F.a_off = P.a_off;
  F.a_cnt = P.a_cnt;
  F.jxyz = P.jxyz;
  F.sign = P.sign;
  if ((rc = upload(c, (void **)&P.proj, proj.data(), proj.size() * 8))) return rc;
  if ((rc = upload(c, (void **)&P.nlproj, nlp.data(), nlp.size() * 8))) return rc;
  if ((rc = upload(c, (void **)&P.y1, y.data(), y.size() * 8))) return rc;
  if ((rc = upload(c, (void **)&P.ls, lsd, sizeof lsd))) return rc;
  auto pf = cast_vec<float>(proj.data(), proj.size());
  auto nf = cast_vec<float>(nlp.data(), nlp.size());
  auto yf = cast_vec<float>(y.data(), y.size());
  auto lf = cast_vec<float>(lsd, 72);
  if ((rc = upload(c, (void **)&F.proj, pf.data(), pf.size() * 4))) return rc;
  if ((rc = upload(c, (void **)&F.nlproj, nf.data(), nf.size() * 4))) return rc;
  if ((rc = upload(c, (void **)&F.y1, yf.data(), yf.size() * 4))) return rc;
  if ((rc = upload(c, (void **)&F.ls, lf.data(), lf.size() * 4))) return rc;
  // Colour the atoms so that atoms sharing a grid point differ in colour
  // (greedy, in atom order, over the conflict graph of shared points).
  {
    std::vector<std::pair<int, int>> pa(npts);
    for (long a = 0; a < na; ++a)
      for (int g = 0; g < cnt[a]; ++g) pa[off[a] + g] = {jx[off[a] + g], (int)a};
    std::sort(pa.begin(), pa.end());
    std::vector<std::vector<int>> adj(na);
    for (size_t i = 0; i < pa.size();) {
      size_t j = i;
      while (j < pa.size() && pa[j].first == pa[i].first) ++j;
      for (size_t u = i; u < j; ++u)
        for (size_t v = u + 1; v < j; ++v) {
          adj[pa[u].second].push_back(pa[v].second);
          adj[pa[v].second].push_back(pa[u].second);
        }
      i = j;
    }
    std::vector<int> color(na, -1);
    int ncolors = 0;
    for (long a = 0; a < na; ++a) {
      std::sort(adj[a].begin(), adj[a].end());
      adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
      std::vector<char> used(ncolors + 1, 0);
      for (int b : adj[a])
        if (color[b] >= 0) used[color[b]] = 1;
      int k = 0;
      while (used[k]) ++k;
      color[a] = k;
      ncolors = std::max(ncolors, k + 1);
    }
    std::vector<int> order;
    P.color_off.assign(ncolors + 1, 0);
    for (int k = 0; k < ncolors; ++k) {
      P.color_off[k] = (int)order.size();
      for (long a = 0; a < na; ++a)
        if (color[a] == k) order.push_back((int)a);
    }
    P.color_off[ncolors] = (int)order.size();
    if ((rc = upload(c, (void **)&P.color_atoms, order.data(), order.size() * sizeof(int)))) return rc;
    F.color_atoms = P.color_atoms;
    // z-y-x order of the middle sphere point (points are stored in the scan
    // order z, y, x, so the middle one sits near the atom)
    std::vector<std::pair<int, int>> key(na);
    for (long a = 0; a < na; ++a) key[a] = {cnt[a] ? jx[off[a] + cnt[a] / 2] : 0, (int)a};
    std::sort(key.begin(), key.end());
    std::vector<int> sp(na);
    for (long a = 0; a < na; ++a) sp[a] = key[a].second;
    if ((rc = upload(c, (void **)&P.spatial_atoms, sp.data(), sp.size() * sizeof(int)))) return rc;
    F.spatial_atoms = P.spatial_atoms;
    F.color_off = P.color_off;
  }
  c->has_proj = (pd->so || pd->nl) && npts > 0;
  // Projector entries by xy plane and rank: the combine pass adds, in rounds
  // k = 0, 1, ..., the contribution of every point's k-th atom (atoms in
  // index order), so no two additions of one round meet at a point and the
  // summation order per point is fixed.
  if (c->has_proj) {
    const long nxy = c->nx * c->ny;
    std::vector<int> rank(npts), pcnt(c->ngrid, 0);
    for (long a = 0; a < na; ++a)
      for (int k = 0; k < cnt[a]; ++k) {
        const int idx = off[a] + k;
        rank[idx] = pcnt[jx[idx]]++;
      }
    int nr = 0;
    for (long g = 0; g < c->ngrid; ++g) nr = std::max(nr, pcnt[g]);
    c->pt_ranks = nr;
    // bucket (z, rank), then point order inside a bucket
    std::vector<int> roff((size_t)c->nz * nr + 1, 0);
    for (long i = 0; i < npts; ++i) ++roff[(size_t)(jx[i] / nxy) * nr + rank[i] + 1];
    for (size_t b = 0; b + 1 < roff.size(); ++b) roff[b + 1] += roff[b];
    c->rec_seg_max.assign(nr, 0);
    for (long z = 0; z < c->nz; ++z)
      for (int k = 0; k < nr; ++k) {
        const size_t b = (size_t)z * nr + k;
        c->rec_seg_max[k] = std::max(c->rec_seg_max[k], roff[b + 1] - roff[b]);
      }
    std::vector<std::pair<long, int>> keyed(npts);  // (bucket * ngrid + point, entry)
    for (long i = 0; i < npts; ++i)
      keyed[i] = {((long)(jx[i] / nxy) * nr + rank[i]) * c->ngrid + jx[i], (int)i};
    std::sort(keyed.begin(), keyed.end());
    std::vector<int> atom_of(npts);
    for (long a = 0; a < na; ++a)
      for (int k = 0; k < cnt[a]; ++k) atom_of[off[a] + k] = (int)a;
    std::vector<int> pa((size_t)npts * 2);
    std::vector<double> pe((size_t)npts * 8);
    for (long e = 0; e < npts; ++e) {
      const int idx = keyed[e].second;
      const long g = jx[idx], px = g % c->nx, py = (g / c->nx) % c->ny;
      pa[2 * e] = (int)((py << 16) | px);
      pa[2 * e + 1] = atom_of[idx];
      double *v = &pe[(size_t)e * 8];
      v[0] = (pd->so && pd->n_so_proj > 0) ? proj[idx] : 0.0;
      v[1] = (pd->nl && pd->n_nl_proj > 0) ? nlp[idx] : 0.0;
      for (int m = 0; m < 3; ++m) {
        v[2 + 2 * m] = y[2 * (m * npts + idx)];
        v[3 + 2 * m] = y[2 * (m * npts + idx) + 1];
      }
    }
    // row offsets inside each bucket (records of a bucket are in point order,
    // so their rows ascend)
    std::vector<int> yoff((size_t)c->nz * nr * (c->ny + 1));
    for (size_t b = 0; b + 1 < roff.size(); ++b) {
      int e = roff[b];
      for (long y = 0; y <= c->ny; ++y) {
        while (e < roff[b + 1] && (pa[2 * (size_t)e] >> 16) < y) ++e;
        yoff[b * (c->ny + 1) + y] = e;
      }
    }
    auto pef = cast_vec<float>(pe.data(), (long)pe.size());
    if ((rc = upload(c, (void **)&c->d_rec_yoff, yoff.data(), yoff.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_rec_off, roff.data(), roff.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_rec_pa, pa.data(), pa.size() * sizeof(int))) ||
        (rc = upload(c, (void **)&c->d_rec_val, pe.data(), pe.size() * 8)) ||
        (rc = upload(c, (void **)&c->d_rec_val_f, pef.data(), pef.size() * 4)))
      return rc;
  }
  return FBSE_OK;
}

// ---------------------------------------------------------------------------
// Work space of fbse_apply_h_dev_noalloc: one file-scope device array, which
// every GPU the library runs on holds from module load to process end.  In
// host code the array's name is only a stand-in: its device address is looked
// up once per device (hipGetSymbolAddress, at the first fbse_create on that
// device, with the device current) and kept in noalloc_base[device]; kernels
// and copies get that address, never the name.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocBytes = (size_t)640 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static constexpr int kNoAllocDevices = 64;
static std::mutex noalloc_mu;
static void *noalloc_base[kNoAllocDevices] = {};

static int noalloc_setup(fbse_ctx *c) {
  if (c->device < 0 || c->device >= kNoAllocDevices)
    return set_err(c, FBSE_ERR_UNSUPPORTED, "device %d beyond the %d the library tracks", c->device, kNoAllocDevices);
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (noalloc_base[c->device]) return FBSE_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_err(c, FBSE_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", c->device,
                   hipGetErrorString(e));
  noalloc_base[c->device] = ptr;
  return FBSE_OK;
}

static void *noalloc_ptr(int device) {
  if (device < 0 || device >= kNoAllocDevices) return nullptr;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  return noalloc_base[device];
}

extern "C" int fbse_create(const fbse_grid_desc *g, const fbse_proj_desc *pd, int device, fbse_ctx **out) {
  if (!out) return FBSE_ERR_ARG;
  *out = nullptr;
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1 || (g->nspin != 1 && g->nspin != 2) || !g->ksqr || !g->pot_local)
    return FBSE_ERR_ARG;
  if (g->nx > 1024 || g->ny > 1024 || g->nz > 1024) return FBSE_ERR_UNSUPPORTED;
  if (g->nspin == 2 && !g->is_complex) return FBSE_ERR_ARG;  // spinors are complex (read.c:919-926)
  if (pd && pd->so && g->nspin != 2) return FBSE_ERR_ARG;   // SO needs spinors (read.c:914-917)
  fbse_ctx *c = new fbse_ctx();
  int rc = FBSE_OK;
  if (device < 0) {
    if (hipGetDevice(&c->device) != hipSuccess) {
      delete c;
      return FBSE_ERR_HIP;
    }
  } else {
    c->device = device;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FBSE_ERR_HIP;
  }
  c->stream = c->own_stream;
  if ((rc = noalloc_setup(c))) {
    fbse_destroy(c);
    return rc;
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    fbse_destroy(c);
    return FBSE_ERR_HIP;
  }
  if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
    fbse_destroy(c);
    return FBSE_ERR_SOLVER;
  }
  rocblas_set_stream(c->blas, c->stream);
  c->nx = g->nx;
  c->ny = g->ny;
  c->nz = g->nz;
  c->ngrid = g->nx * g->ny * g->nz;
  c->nspin = g->nspin;
  c->nsg = c->nspin * c->ngrid;
  c->is_complex = g->is_complex;
  c->dv = g->dv;

  // ksqr: detect the separable form of build_grid_ksqr (init.c:336-347) so the
  // z pass can rebuild k^2 from three 1D tables instead of reading N doubles.
  {
    const long nx = c->nx, ny = c->ny, nz = c->nz;
    std::vector<double> kx(nx), ky(ny), kz(nz);
    double kmax = 0.0;
    for (long i = 0; i < c->ngrid; ++i) kmax = std::max(kmax, g->ksqr[i]);
    for (long i = 0; i < nx; ++i) kx[i] = g->ksqr[i];
    for (long j = 0; j < ny; ++j) ky[j] = g->ksqr[nx * j];
    for (long k = 0; k < nz; ++k) kz[k] = g->ksqr[nx * ny * k];
    bool sep = g->ksqr[0] == 0.0;
    for (long k = 0; sep && k < nz; ++k)
      for (long j = 0; sep && j < ny; ++j)
        for (long i = 0; i < nx; ++i) {
          double v = kx[i] + ky[j] + kz[k];
          if (v > kmax) v = kmax;
          if (v != g->ksqr[nx * (ny * k + j) + i]) {
            sep = false;
            break;
          }
        }
    // (an axis entry at the cap is exact too: any sum with it reaches the cap;
    // the check above covers every point either way)
    c->ksqr_sep = sep ? 1 : 0;
    c->kcap = kmax;
    if ((rc = upload(c, (void **)&c->d_kx, kx.data(), nx * 8)) || (rc = upload(c, (void **)&c->d_ky, ky.data(), ny * 8)) ||
        (rc = upload(c, (void **)&c->d_kz, kz.data(), nz * 8)) ||
        (rc = upload(c, (void **)&c->d_ksqr, g->ksqr, c->ngrid * 8))) {
      fbse_destroy(c);
      return rc;
    }
    auto fx = cast_vec<float>(kx.data(), nx), fy = cast_vec<float>(ky.data(), ny), fz = cast_vec<float>(kz.data(), nz);
    auto fk = cast_vec<float>(g->ksqr, c->ngrid);
    if ((rc = upload(c, (void **)&c->d_kx_f, fx.data(), nx * 4)) || (rc = upload(c, (void **)&c->d_ky_f, fy.data(), ny * 4)) ||
        (rc = upload(c, (void **)&c->d_kz_f, fz.data(), nz * 4)) ||
        (rc = upload(c, (void **)&c->d_ksqr_f, fk.data(), c->ngrid * 4))) {
      fbse_destroy(c);
      return rc;
    }
  }
  {
    auto vf = cast_vec<float>(g->pot_local, c->nsg);
    if ((rc = upload(c, (void **)&c->d_vloc, g->pot_local, c->nsg * 8)) ||
        (rc = upload(c, (void **)&c->d_vloc_f, vf.data(), c->nsg * 4))) {
      fbse_destroy(c);
      return rc;
    }
  }
  const long dims[3] = {c->nx, c->ny, c->nz};
  for (int ax = 0; ax < 3; ++ax) {
    const int n = (int)dims[ax];
    factorize(n, c->plan[ax]);
    axis_schedule(n, c->aplan[ax]);
    // tile shape: x rows fill ~2048 elements; y/z take 16 consecutive inner
    // indices (256-byte rows for FP64)
    if (ax == 0) c->plan[ax].lines = std::max(1, std::min(64, 2048 / n));
    else c->plan[ax].lines = n > 256 ? 8 : 16;
    std::vector<double2> tw(n);
    std::vector<float2> twf(n);
    for (int t = 0; t < n; ++t) {
      const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
      tw[t] = make_double2((double)cosl(ang), (double)sinl(ang));
      twf[t] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    if ((rc = upload(c, (void **)&c->d_tw[ax], tw.data(), n * sizeof(double2))) ||
        (rc = upload(c, (void **)&c->d_tw_f[ax], twf.data(), n * sizeof(float2)))) {
      fbse_destroy(c);
      return rc;
    }
  }
  if (pd && (pd->so || pd->nl)) {
    if ((rc = build_projectors(c, pd))) {
      std::string msg = c->err;
      fprintf(stderr, "fbse_create: %s\n", msg.c_str());
      fbse_destroy(c);
      return rc;
    }
  }
  *out = c;
  return FBSE_OK;
}

extern "C" void fbse_destroy(fbse_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->d_vloc, c->d_vloc_f, c->d_kx, c->d_ky, c->d_kz, c->d_ksqr, c->d_kx_f, c->d_ky_f, c->d_kz_f,
                  c->d_ksqr_f, c->d_scratch, c->d_scratch2, c->d_red, c->pj.a_off, c->pj.a_cnt, c->pj.jxyz,
                  c->pj.sign, c->pj.proj, c->pj.nlproj, c->pj.y1, c->pj.ls, c->pj_f.proj, c->pj_f.nlproj,
                  c->pj_f.y1, c->pj_f.ls, c->pj.color_atoms, c->pj.spatial_atoms, c->d_step, c->d_rec_off, c->d_rec_pa,
                  c->d_rec_val, c->d_rec_val_f, c->d_pjd, c->d_gram, c->d_rec_yoff, c->d_pp_lane_off,
                  c->d_pp_lanes, c->d_pp_span, c->d_pp_meta, c->d_pp_val, c->d_pp_val_f, c->d_pp_atom_off, c->d_pp_part};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (int a = 0; a < 3; ++a) {
    if (c->d_tw[a]) hipFree(c->d_tw[a]);
    if (c->d_tw_f[a]) hipFree(c->d_tw_f[a]);
  }
  if (c->h_red) hipHostFree(c->h_red);
  io_release(c);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  for (auto &p : c->ev_pending) {
    hipEventDestroy(p.second.first);
    hipEventDestroy(p.second.second);
  }
  if (c->blas) rocblas_destroy_handle(c->blas);
  if (c->side) hipStreamSynchronize(c->side);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
}

extern "C" const char *fbse_last_error(const fbse_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int fbse_set_stream(fbse_ctx *c, void *stream) {
  if (!c) return FBSE_ERR_ARG;
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  rocblas_set_stream(c->blas, c->stream);
  return FBSE_OK;
}
extern "C" int fbse_synchronize(fbse_ctx *c) {
  if (!c) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  return FBSE_OK;
}
extern "C" int fbse_profile(fbse_ctx *c, int enable) {
  if (!c) return FBSE_ERR_ARG;
  prof_collect(c);
  c->profile = enable;
  for (auto &t : c->timers) t = Timer{};
  return FBSE_OK;
}
extern "C" int fbse_profile_read(fbse_ctx *c, int kind, long *launches, double *total_ms) {
  if (!c || kind < 0 || kind > 4) return FBSE_ERR_ARG;
  prof_collect(c);
  if (launches) *launches = c->timers[kind].launches;
  if (total_ms) *total_ms = c->timers[kind].ms;
  return FBSE_OK;
}
extern "C" int fbse_set_window(fbse_ctx *c, int w) {
  if (!c || w < 0 || w > kMaxWindow) return FBSE_ERR_ARG;
  c->window = w;
  return FBSE_OK;
}
extern "C" int fbse_set_fft_kernel(fbse_ctx *c, int variant, int lines) {
  if (!c || variant < -1 || variant > 1 || lines < 0 || (lines > 256 && !(variant == 0 && (lines == 1016 || lines == 1032 || lines == 1064))))
    return FBSE_ERR_ARG;
  c->force_variant = variant;
  c->force_lines = lines;
  return FBSE_OK;
}
extern "C" int fbse_set_slab_split(fbse_ctx *c, int split) {
  if (!c || split < 0 || split > 16) return FBSE_ERR_ARG;
  c->force_zsplit = split;
  return FBSE_OK;
}
extern "C" int fbse_set_graph(fbse_ctx *c, int enable) {
  if (!c || enable < 0 || enable > 1) return FBSE_ERR_ARG;
  c->use_graph = enable;
  return FBSE_OK;
}
extern "C" int fbse_set_axis_fusion(fbse_ctx *c, int enable) {
  if (!c || !(enable == 0 || enable == 1 || enable == 16 || enable == 32 || enable == 64)) return FBSE_ERR_ARG;
  c->axis_fuse = enable;
  return FBSE_OK;
}
extern "C" int fbse_set_combine_projection(fbse_ctx *c, int enable) {
  if (!c || enable < 0 || enable > 1) return FBSE_ERR_ARG;
  if (enable && !c->d_pp_lanes && c->has_proj && c->nspin == 2 && c->ny % 8 == 0) {
    // the tables are built on first use from the rank records on the device
    FBSE_HIP(c, hipSetDevice(c->device));
    const long npts = c->pj.npts, nr = c->pt_ranks;
    std::vector<int> pa((size_t)npts * 2), yoff((size_t)c->nz * nr * (c->ny + 1));
    std::vector<double> pe((size_t)npts * 8);
    FBSE_HIP(c, hipMemcpy(pa.data(), c->d_rec_pa, pa.size() * sizeof(int), hipMemcpyDeviceToHost));
    FBSE_HIP(c, hipMemcpy(yoff.data(), c->d_rec_yoff, yoff.size() * sizeof(int), hipMemcpyDeviceToHost));
    FBSE_HIP(c, hipMemcpy(pe.data(), c->d_rec_val, pe.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int rc = build_pp_tables(c, c->pj.n_atoms, (int)nr, pa, pe, yoff);
    if (rc) return rc;
  }
  c->pp_on = enable;
  c->pp_ready = 0;
  return FBSE_OK;
}
extern "C" int fbse_trim(fbse_ctx *c) {
  if (!c) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  release_work(c);
  io_release(c);
  return FBSE_OK;
}
extern "C" long fbse_graph_replays(const fbse_ctx *c) { return c ? c->graph_replays : -1; }
extern "C" int fbse_last_window(const fbse_ctx *c) { return c ? c->last_window : -1; }
extern "C" long fbse_last_runs(const fbse_ctx *c) { return c ? c->last_runs : -1; }
extern "C" int fbse_set_run_starts(fbse_ctx *c, long n) {
  if (!c || n < 0) return FBSE_ERR_ARG;
  c->run_starts = n;
  return FBSE_OK;
}
extern "C" int fbse_last_group(const fbse_ctx *c) { return c ? c->last_group : -1; }
extern "C" int fbse_held_bytes(fbse_ctx *c, size_t *out) {
  if (!c || !out) return FBSE_ERR_ARG;
  out[0] = c->scratch_bytes;
  out[1] = c->scratch2_bytes;
  out[2] = c->gram_bytes;
  out[3] = blas_workspace_bytes(c);
  return FBSE_OK;
}
extern "C" int fbse_fft_choice(fbse_ctx *c, int precision, int *path, int *instance) {
  if (!c || !path || (precision != FBSE_PREC_FP64 && precision != FBSE_PREC_MIXED)) return FBSE_ERR_ARG;
  *path = fft_path(c, precision == FBSE_PREC_MIXED ? 1 : 0, instance);
  return FBSE_OK;
}
extern "C" int fbse_dev_alloc(fbse_ctx *c, size_t bytes, void **p) {
  if (!c || !p) return FBSE_ERR_ARG;
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) return set_err(c, FBSE_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return FBSE_OK;
}
extern "C" int fbse_dev_free(fbse_ctx *c, void *p) {
  if (!c) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  if (p) FBSE_HIP(c, hipFree(p));
  return FBSE_OK;
}
extern "C" int fbse_memcpy_h2d(fbse_ctx *c, void *d, const void *h, size_t b) {
  if (!c) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, c->stream));
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  return FBSE_OK;
}
extern "C" int fbse_memcpy_d2h(fbse_ctx *c, void *h, const void *d, size_t b) {
  if (!c) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, c->stream));
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  return FBSE_OK;
}

// ---------------------------------------------------------------------------
// H application in the reference state layout
// ---------------------------------------------------------------------------
static long stlen_of(const fbse_ctx *c) { return c->nsg * (c->is_complex + 1); }

// Max states per H chunk: bounded so the scratch stays below ~8 GiB.
static long h_chunk(const fbse_ctx *c) {
  const size_t per = (size_t)c->nsg * sizeof(double2);
  long ch = (long)((8ull << 30) / (3 * per));
  return std::max(1L, std::min(ch, 4096L));
}

static int apply_h_dev_impl(fbse_ctx *c, const double *d_psi, double *d_hpsi, long nstates) {
  int rc;
  const long ch = h_chunk(c);
  const size_t per = (size_t)c->nsg * sizeof(double2);
  if (c->is_complex) {
    if ((rc = ensure_scratch(c, (size_t)std::min(ch, nstates) * per))) return rc;
    for (long s0 = 0; s0 < nstates; s0 += ch) {
      const long ns = std::min(ch, nstates - s0);
      const double2 *in = reinterpret_cast<const double2 *>(d_psi) + s0 * c->nsg;
      double2 *out = reinterpret_cast<double2 *>(d_hpsi) + s0 * c->nsg;
      if ((rc = apply_h_cplx<double>(c, in, out, (double2 *)c->d_scratch, ns, 1.0, 0.0))) return rc;
    }
  } else {
    const long chp = std::max(2L, ch & ~1L);
    const long np = (std::min(chp, nstates) + 1) / 2;
    if ((rc = ensure_scratch(c, (size_t)np * per * 2))) return rc;
    if ((rc = ensure_scratch2(c, (size_t)np * per))) return rc;
    double2 *packed = (double2 *)c->d_scratch;
    double2 *work = packed + np * c->nsg;
    double2 *res = (double2 *)c->d_scratch2;
    for (long s0 = 0; s0 < nstates; s0 += chp) {
      const long ns = std::min(chp, nstates - s0);
      const long nps = (ns + 1) / 2;
      if ((rc = pack_real(c, d_psi + s0 * c->nsg, c->nsg, ns, packed))) return rc;
      if ((rc = apply_h_cplx<double>(c, packed, res, work, nps, 1.0, 0.0))) return rc;
      if ((rc = unpack_real(c, res, ns, d_hpsi + s0 * c->nsg, c->nsg))) return rc;
    }
  }
  return FBSE_OK;
}

extern "C" int fbse_apply_h_dev(fbse_ctx *c, const double *d_psi, double *d_hpsi, long nstates) {
  if (!c || nstates < 0 || (nstates && (!d_psi || !d_hpsi))) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return apply_h_dev_impl(c, d_psi, d_hpsi, nstates);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes at the front of the context's work buffer that apply_h_dev_impl (and
// precondition_dev) use for n states: the FFT work space of one chunk.
static size_t h_work_bytes(const fbse_ctx *c, long nstates) {
  const long ch = h_chunk(c);
  const size_t per = (size_t)c->nsg * sizeof(double2);
  if (c->is_complex) return (size_t)std::min(ch, nstates) * per;
  const long chp = std::max(2L, ch & ~1L);
  return (size_t)((std::min(chp, nstates) + 1) / 2) * per * 2;
}
// After a filter the context's work buffer holds its ring, which the stages
// that follow no longer need.  Their large temporaries (H U of Rayleigh-Ritz,
// the refinement's blocks) are taken from the buffer past the front bytes the
// H applications use, instead of freeing the ring and allocating anew: a fresh
// allocation of tens of GB costs seconds (about 2 s for 75 GB on the MI355X),
// and the next filter then finds its ring in place.
// States of per_state bytes the buffer holds past `front`, and their address.
static long arena_states(const fbse_ctx *c, size_t front, size_t per_state) {
  front = align256(front);
  return c->d_scratch && c->scratch_bytes > front ? (long)((c->scratch_bytes - front) / per_state) : 0;
}
static double *arena_at(const fbse_ctx *c, size_t front) {
  return reinterpret_cast<double *>(static_cast<char *>(c->d_scratch) + align256(front));
}

// fbse_apply_h_dev without device allocation: the FFT work buffer(s) and the
// projection coefficients of apply_h_dev_impl are carved out of the library's
// static work space instead of the context's buffers, and the autotuners
// (which time candidates in a scratch allocation) and profiling events stay
// off for the call; every other step is the same code.  No byte of the work
// space is read before this call writes it.
extern "C" int fbse_apply_h_dev_noalloc(fbse_ctx *c, const double *d_psi, double *d_hpsi, long nstates) {
  if (!c || nstates < 0 || (nstates && (!d_psi || !d_hpsi))) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  char *base = (char *)noalloc_ptr(c->device);
  if (!base) return set_err(c, FBSE_ERR_HIP, "no-allocation work space of device %d is not set up", c->device);
  if (nstates == 0) return FBSE_OK;
  const size_t per = (size_t)c->nsg * sizeof(double2);
  // the buffers apply_h_dev_impl asks for: complex states in one chunk; real
  // states packed two per complex grid (packed input + FFT work, result)
  const long items = c->is_complex ? nstates : (nstates + 1) / 2;
  size_t need0 = c->is_complex ? (size_t)items * per : (size_t)items * per * 2;
  size_t need1 = c->is_complex ? 0 : (size_t)items * per;
  size_t need2 = c->has_proj ? (size_t)c->pj.n_atoms * items * c->nspin * 6 * sizeof(double2) : 0;
  need0 = align256(need0);
  need1 = align256(need1);
  need2 = align256(need2);
  if (need0 + need1 + need2 > kNoAllocBytes || (long)(per ? (8ull << 30) / (3 * per) : 0) < nstates)
    return set_err(c, FBSE_ERR_ARG,
                   "fbse_apply_h_dev_noalloc: %ld states need %zu bytes of work space, more than its %zu", nstates,
                   need0 + need1 + need2, kNoAllocBytes);
  void *sv_s = c->d_scratch, *sv_s2 = c->d_scratch2, *sv_pjd = c->d_pjd;
  const size_t sv_sb = c->scratch_bytes, sv_s2b = c->scratch2_bytes, sv_pjdb = c->pjd_bytes;
  const int sv_tune = c->autotune, sv_prof = c->profile;
  c->d_scratch = base;
  c->scratch_bytes = need0;
  c->d_scratch2 = need1 ? base + need0 : nullptr;
  c->scratch2_bytes = need1;
  c->d_pjd = need2 ? base + need0 + need1 : nullptr;
  c->pjd_bytes = need2;
  c->autotune = 0;
  c->profile = 0;
  c->no_alloc = 1;
  const int rc = apply_h_dev_impl(c, d_psi, d_hpsi, nstates);
  c->no_alloc = 0;
  c->autotune = sv_tune;
  c->profile = sv_prof;
  c->d_scratch = sv_s;
  c->scratch_bytes = sv_sb;
  c->d_scratch2 = sv_s2;
  c->scratch2_bytes = sv_s2b;
  c->d_pjd = sv_pjd;
  c->pjd_bytes = sv_pjdb;
  return rc;
}

// host-pointer wrapper: stage through device buffers in chunks
template <typename F>
static int host_staged(fbse_ctx *c, long nstates, long per_state_doubles, F &&fn) {
  const long ch = std::max(1L, std::min(nstates, h_chunk(c)));
  double *d_a = nullptr, *d_b = nullptr;
  const size_t bytes = (size_t)ch * per_state_doubles * sizeof(double);
  FBSE_HIP(c, hipMalloc(&d_a, bytes));
  if (hipMalloc(&d_b, bytes) != hipSuccess) {
    hipFree(d_a);
    return set_err(c, FBSE_ERR_NOMEM, "staging buffer");
  }
  int rc = FBSE_OK;
  for (long s0 = 0; s0 < nstates && rc == FBSE_OK; s0 += ch) {
    const long ns = std::min(ch, nstates - s0);
    rc = fn(s0, ns, d_a, d_b);
  }
  hipStreamSynchronize(c->stream);
  hipFree(d_a);
  hipFree(d_b);
  return rc;
}

extern "C" int fbse_apply_h(fbse_ctx *c, const double *psi, double *hpsi, long nstates) {
  if (!c || nstates < 0 || (nstates && (!psi || !hpsi))) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  return host_staged(c, nstates, sl, [&](long s0, long ns, double *da, double *db) -> int {
    FBSE_HIP(c, hipMemcpyAsync(da, psi + s0 * sl, ns * sl * 8, hipMemcpyHostToDevice, c->stream));
    int rc = apply_h_dev_impl(c, da, db, ns);
    if (rc) return rc;
    FBSE_HIP(c, hipMemcpyAsync(hpsi + s0 * sl, db, ns * sl * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    return FBSE_OK;
  });
}

// ---------------------------------------------------------------------------
// normalize_all / energy_all / sigma
// ---------------------------------------------------------------------------
static int normalize_dev_impl(fbse_ctx *c, double *d_psi, long n) {
  if (n == 0) return FBSE_OK;
  const long sl = stlen_of(c);
  int rc;
  double *d_n = nullptr;
  FBSE_HIP(c, hipMalloc(&d_n, n * sizeof(double)));
  if ((rc = state_norms(c, d_psi, n, sl, d_n))) {
    hipFree(d_n);
    return rc;
  }
  std::vector<double> h(n);
  FBSE_HIP(c, hipMemcpyAsync(h.data(), d_n, n * 8, hipMemcpyDeviceToHost, c->stream));
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  for (long s = 0; s < n; ++s) {
    double norm = h[s] * c->dv;
    if (norm == 0.0) {
      hipFree(d_n);
      return set_err(c, FBSE_ERR_NORM_ZERO, "Norm is %.2g! exiting! (state %ld)", norm, s);
    }
    if (std::isnan(norm)) {
      hipFree(d_n);
      return set_err(c, FBSE_ERR_NORM_NAN, "Norm is NAN! exiting! (state %ld)", s);
    }
    h[s] = 1.0 / sqrt(norm);
  }
  FBSE_HIP(c, hipMemcpyAsync(d_n, h.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  rc = state_scale(c, d_psi, n, sl, d_n);
  hipStreamSynchronize(c->stream);
  hipFree(d_n);
  return rc;
}
extern "C" int fbse_normalize_all_dev(fbse_ctx *c, double *d_psi, long n) {
  if (!c || n < 0) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return normalize_dev_impl(c, d_psi, n);
}
extern "C" int fbse_normalize_all(fbse_ctx *c, double *psi, long n) {
  if (!c || n < 0) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  return host_staged(c, n, sl, [&](long s0, long ns, double *da, double *) -> int {
    FBSE_HIP(c, hipMemcpyAsync(da, psi + s0 * sl, ns * sl * 8, hipMemcpyHostToDevice, c->stream));
    int rc = normalize_dev_impl(c, da, ns);
    if (rc) return rc;
    FBSE_HIP(c, hipMemcpyAsync(psi + s0 * sl, da, ns * sl * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    return FBSE_OK;
  });
}

// energies (and optionally residual norms) of n device states, chunked; the
// H psi chunk comes from the work buffer past `arena_front` bytes (the front
// holds the H applications' FFT work space, and whatever the caller keeps
// there) when it has room, else it is allocated
static int energies_dev_impl(fbse_ctx *c, const double *d_psi, long n, double *ene, double *sig,
                             size_t arena_front) {
  const long sl = stlen_of(c);
  const long ch = std::max(1L, std::min(n, h_chunk(c)));
  double *d_h = nullptr, *d_e = nullptr, *d_r = nullptr;
  const bool in_arena = arena_front >= h_work_bytes(c, n) && arena_states(c, arena_front, (size_t)sl * 8) >= ch;
  if (in_arena)
    d_h = arena_at(c, arena_front);
  else
    FBSE_HIP(c, hipMalloc(&d_h, (size_t)ch * sl * 8));
  FBSE_HIP(c, hipMalloc(&d_e, (size_t)n * 8));
  FBSE_HIP(c, hipMalloc(&d_r, (size_t)n * 8));
  int rc = FBSE_OK;
  std::vector<double> he(n), hr(n);
  for (long s0 = 0; s0 < n && !rc; s0 += ch) {
    const long ns = std::min(ch, n - s0);
    rc = apply_h_dev_impl(c, d_psi + s0 * sl, d_h, ns);
    if (!rc) rc = state_dots(c, d_psi + s0 * sl, d_h, ns, sl, d_e + s0);
    if (!rc && sig) {
      // eps = <H> (unscaled dot * dv) on device: scale the dots first
      std::vector<double> tmp(ns);
      hipMemcpyAsync(tmp.data(), d_e + s0, ns * 8, hipMemcpyDeviceToHost, c->stream);
      hipStreamSynchronize(c->stream);
      for (long s = 0; s < ns; ++s) tmp[s] *= c->dv;
      hipMemcpyAsync(d_r + s0, tmp.data(), ns * 8, hipMemcpyHostToDevice, c->stream);
      rc = state_resid(c, d_psi + s0 * sl, d_h, d_r + s0, ns, sl, d_r + s0);
      if (!rc) {
        // state_resid reads eps from d_r and overwrites it element-wise after the
        // ordered final reduction; the final kernel runs after the partial one.
      }
    }
  }
  if (!rc) {
    FBSE_HIP(c, hipMemcpyAsync(he.data(), d_e, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (sig) FBSE_HIP(c, hipMemcpyAsync(hr.data(), d_r, n * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    for (long s = 0; s < n; ++s) {
      if (ene) ene[s] = he[s] * c->dv;
      if (sig) sig[s] = sqrt(fabs(hr[s] * c->dv));
    }
  }
  hipStreamSynchronize(c->stream);
  if (!in_arena) hipFree(d_h);
  hipFree(d_e);
  hipFree(d_r);
  return rc;
}
extern "C" int fbse_energy_all_dev(fbse_ctx *c, const double *d_psi, long n, double *ene) {
  if (!c || n < 0 || !ene) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return energies_dev_impl(c, d_psi, n, ene, nullptr, h_work_bytes(c, n));
}
extern "C" int fbse_sigma_dev(fbse_ctx *c, const double *d_psi, long n, double *sigma, double *eval) {
  if (!c || n < 0 || !sigma) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  std::vector<double> e(n);
  int rc = energies_dev_impl(c, d_psi, n, e.data(), sigma, h_work_bytes(c, n));
  if (!rc && eval) memcpy(eval, e.data(), n * 8);
  return rc;
}
template <typename F> static int with_device_copy(fbse_ctx *c, const double *h, long n, F &&fn) {
  const long sl = stlen_of(c);
  double *d = nullptr;
  FBSE_HIP(c, hipMalloc(&d, (size_t)std::max(1L, n) * sl * 8));
  hipError_t e = hipMemcpyAsync(d, h, (size_t)n * sl * 8, hipMemcpyHostToDevice, c->stream);
  int rc = e == hipSuccess ? fn(d) : set_err(c, FBSE_ERR_HIP, "copy: %s", hipGetErrorString(e));
  hipStreamSynchronize(c->stream);
  hipFree(d);
  return rc;
}
extern "C" int fbse_energy_all(fbse_ctx *c, const double *psi, long n, double *ene) {
  if (!c || n < 0 || !ene) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return with_device_copy(c, psi, n, [&](double *d) { return energies_dev_impl(c, d, n, ene, nullptr, h_work_bytes(c, n)); });
}
extern "C" int fbse_sigma(fbse_ctx *c, const double *psi, long n, double *sigma, double *eval) {
  if (!c || n < 0 || !sigma) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return with_device_copy(c, psi, n, [&](double *d) { return fbse_sigma_dev(c, d, n, sigma, eval); });
}

// ---------------------------------------------------------------------------
// Newton filter
// ---------------------------------------------------------------------------
// One filter run (one context, one batch of start states).  The run is a
// small state machine so that several runs (dots) can be interleaved on their
// own streams by fbse_filter_many.
//
// Windows: the first covers steps 1..R-1 (slot 0 holds the start states),
// window w >= 1 steps wR..wR+R-1, writing slots 0..R-1 in order.  Full windows
// w >= 1 therefore launch the same kernels on the same buffers; only the
// Newton shifts a0(j) and the coefficient offset differ, and those are read on
// the device from a0 table[step counter + i] and the counter.  Such windows are
// captured once into a hipGraph and replayed (the counter advances by R at the
// end of each replay).
template <typename T> struct FilterRun {
  using C = cplx_t<T>;
  fbse_ctx *c = nullptr;
  double *d_psi_rank = nullptr;
  long n_starts = 0, ms = 0, M = 0;
  const double *an = nullptr, *zn = nullptr;
  double dE_1 = 0.0, c2 = 0.0;
  int packed = 0, real_an = 1, R = 2;
  long n_items = 0, nsg = 0, sl = 0, slot_elems = 0;
  C *ring = nullptr, *work = nullptr;
  double *d_an = nullptr, *d_a0 = nullptr;
  std::vector<double> an_t;  // host copy of the step-major coefficients (outlives its upload)
  long acc_lo = 0, j = 1, n_full = 0, w_next = 1;
  bool graph = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;

  FilterRun(fbse_ctx *c_, double *psi, long ns, long ms_, long M_, const double *an_, const double *zn_, double Vmin,
            double dE_1_)
      : c(c_), d_psi_rank(psi), n_starts(ns), ms(ms_), M(M_), an(an_), zn(zn_), dE_1(dE_1_) {
    c2 = Vmin * dE_1;
    packed = !c->is_complex;
    n_items = packed ? (n_starts + 1) / 2 : n_starts;
    nsg = c->nsg;
    sl = stlen_of(c);
    slot_elems = n_items * nsg;
  }
  size_t slot_bytes() const { return (size_t)slot_elems * sizeof(C); }
  ~FilterRun() { release(); }
  void release() {
    c->pp_chain = 0;
    c->pp_ready = 0;
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (d_an) (void)hipFree(d_an);
    if (d_a0) (void)hipFree(d_a0);
    ge = nullptr;
    g = nullptr;
    d_an = d_a0 = nullptr;
  }

  // ring (R = K+1 slots) + one work slot; coefficients; start states -> slot 0
  int prepare(int K) {
    int rc;
    // the Newton steps of this run chain their projections through the
    // combine's per-tile sums (each step's input is the step before's output)
    c->pp_chain = 1;
    c->pp_ready = 0;
    R = K + 1;
    c->last_window = K;
    if ((rc = ensure_scratch(c, slot_bytes() * (size_t)(R + 1)))) return rc;
    ring = (C *)c->d_scratch;
    work = ring + (size_t)R * slot_elems;
    real_an = 1;  // all an.im == 0: the accumulation needs only the real parts
    for (long k = 0; k < ms * M && real_an; ++k) real_an = an[2 * k + 1] == 0.0;
    // the accumulation reads the coefficients step-major, anT[j*ms + t] (real
    // parts only when all an.im are 0 or the states are packed real pairs,
    // which accumulate .re alone, filter.c:353-357), else complex pairs
    const bool real_layout = real_an || packed;
    an_t.assign((size_t)M * ms * (real_layout ? 1 : 2), 0.0);
    for (long jj = 0; jj < M; ++jj)
      for (long t = 0; t < ms; ++t) {
        const double *a = an + 2 * (t * M + jj);
        if (real_layout) {
          an_t[jj * ms + t] = a[0];
        } else {
          an_t[2 * (jj * ms + t)] = a[0];
          an_t[2 * (jj * ms + t) + 1] = a[1];
        }
      }
    FBSE_HIP(c, hipMalloc(&d_an, an_t.size() * sizeof(double)));
    FBSE_HIP(c, hipMemcpyAsync(d_an, an_t.data(), an_t.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // start states -> ring slot 0 (complex FP32: converted in place of a
    // staging copy).  psi_rank is not cleared: the first window's
    // accumulation writes it.
    if (!packed && sizeof(T) == 4) {
      if ((rc = to_f32_strided(c, (const double2 *)d_psi_rank, ms * sl / 2, (float2 *)ring, nsg, n_starts)))
        return rc;
    } else {
      double2 *stage;
      if constexpr (sizeof(T) == 8) stage = (double2 *)ring;
      else {
        if ((rc = ensure_scratch2(c, (size_t)slot_elems * sizeof(double2)))) return rc;
        stage = (double2 *)c->d_scratch2;
      }
      if (packed) {
        if ((rc = pack_real(c, d_psi_rank, ms * sl, n_starts, stage))) return rc;
      } else {
        FBSE_HIP(c, hipMemcpy2DAsync(stage, nsg * sizeof(double2), d_psi_rank, ms * sl * sizeof(double),
                                     nsg * sizeof(double2), n_starts, hipMemcpyDeviceToDevice, c->stream));
      }
      if constexpr (sizeof(T) == 4) {
        if ((rc = to_f32(c, stage, (float2 *)ring, slot_elems))) return rc;
      }
    }
    n_full = M / R - 1;  // full windows w = 1 .. n_full
    graph = c->use_graph && !c->profile && n_full >= 2;
    c->graph_replays = 0;
    if (graph) {
      std::vector<double> a0h(M, 0.0);
      for (long jj = 1; jj < M; ++jj) a0h[jj] = -(2.0 + zn[jj - 1] + c2);
      if (!c->d_step) FBSE_HIP(c, hipMalloc(&c->d_step, sizeof(long)));
      FBSE_HIP(c, hipMalloc(&d_a0, (size_t)M * sizeof(double)));
      FBSE_HIP(c, hipMemcpyAsync(d_a0, a0h.data(), (size_t)M * sizeof(double), hipMemcpyHostToDevice, c->stream));
      // a0h must outlive the copy
      FBSE_HIP(c, hipStreamSynchronize(c->stream));
    }
    return FBSE_OK;
  }

  // operands of Newton step j: input and output ring slots, Newton shift
  const C *step_in() const { return ring + (size_t)((j - 1) % R) * slot_elems; }
  C *step_out() const { return ring + (size_t)(j % R) * slot_elems; }
  double step_a0() const { return -(2.0 + zn[j - 1] + c2); }
  // after step j's H application: the window's accumulation when due, next j
  int finish_step() {
    if (j + 1 - acc_lo == R || j == M - 1) {
      int r;
      if ((r = accumulate_window<T>(c, ring, R, slot_elems, acc_lo, j + 1, d_an, ms, M, d_psi_rank, n_starts,
                                    n_items, packed, real_an, nullptr, acc_lo == 0)))
        return r;
      acc_lo = j + 1;
    }
    ++j;
    return FBSE_OK;
  }
  // Newton step j (eager launches), with the window's accumulation when due
  int step() {
    int r = apply_h_chain<T>(c, step_in(), step_out(), work, n_items, (T)dE_1, (T)step_a0(), j == 1);
    return r ? r : finish_step();
  }
  bool in_window0() const { return j < R && j < M; }

  // window capture in pieces (a group of runs captures them into one graph):
  // the device step counter before the capture, the Newton shifts read from
  // the device table for in-window step i, the window's accumulation and
  // counter update at the end of it
  int capture_prologue() {
    int rc;
    // counter = R (start of window 1), set by kernels so the stream order holds
    if ((rc = fill_zero(c, c->d_step, sizeof(long))) || (rc = step_advance(c, c->d_step, R))) return rc;
    return FBSE_OK;
  }
  void capture_step(int i) {
    c->capturing = 1;
    c->dyn_a0_tab = d_a0;
    c->dyn_step_base = c->d_step;
    c->dyn_step_off = i;
  }
  const C *window_in(int i) const { return ring + (size_t)((i + R - 1) % R) * slot_elems; }
  C *window_out(int i) const { return ring + (size_t)i * slot_elems; }
  int capture_epilogue() {
    c->dyn_a0_tab = nullptr;
    c->dyn_step_base = nullptr;
    int rc = accumulate_window<T>(c, ring, R, slot_elems, R, 2 * R, d_an, ms, M, d_psi_rank, n_starts, n_items,
                                  packed, real_an, c->d_step);
    return rc ? rc : step_advance(c, c->d_step, R);
  }
  void capture_done() {
    c->capturing = 0;
    c->dyn_a0_tab = nullptr;
    c->dyn_step_base = nullptr;
  }

  // capture one full window (call after window 0)
  int capture() {
    if (!graph) return FBSE_OK;
    int rc;
    if ((rc = capture_prologue())) return rc;
    FBSE_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int crc = FBSE_OK;
    for (int i = 0; i < R && !crc; ++i) {
      capture_step(i);
      crc = apply_h_chain<T>(c, window_in(i), window_out(i), work, n_items, (T)dE_1, (T)0, false);
    }
    if (!crc) crc = capture_epilogue();
    hipError_t ee = hipStreamEndCapture(c->stream, &g);
    capture_done();
    hipError_t ei = (crc || ee != hipSuccess) ? hipSuccess : hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (crc) return crc;
    if (ee != hipSuccess || ei != hipSuccess)
      return set_err(c, FBSE_ERR_HIP, "Newton window capture: %s (fbse_set_graph(ctx, 0) launches eagerly)",
                     hipGetErrorString(ee != hipSuccess ? ee : ei));
    return FBSE_OK;
  }
  bool replays_left() const { return graph && ge && w_next <= n_full; }
  // bookkeeping of one replayed window (the graph's or the group's)
  void replayed() {
    ++c->graph_replays;
    ++w_next;
    if (w_next > n_full) {  // all full windows queued: the tail continues eagerly
      j = (n_full + 1) * R;
      acc_lo = j;
    }
  }
  int replay() {
    FBSE_HIP(c, hipGraphLaunch(ge, c->stream));
    replayed();
    return FBSE_OK;
  }
  bool tail_left() const { return j < M; }
  int finish() {
    int rc = FBSE_OK;
    if (M == 1)
      rc = accumulate_window<T>(c, ring, R, slot_elems, 0, 1, d_an, ms, M, d_psi_rank, n_starts, n_items, packed,
                                real_an, nullptr, true);
    (void)hipStreamSynchronize(c->stream);  // replays done before the graph exec goes away
    release();
    if (!rc) {
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = set_err(c, FBSE_ERR_HIP, "filter: %s", hipGetErrorString(e));
    }
    return rc;
  }
};

// window depth: forced, or as deep as 75% of free memory allows for all the
// runs' rings together (<= kMaxWindow); the rest covers the FP32 staging
// slot, the autotuners' scratch and the projection coefficients.  The rings
// stay within kMaxRingBytes: a 165 GB FP64 ring (K = 89 at 96^3) made the
// filter twice as slow as a 110 GB one (K = 59), while rings up to 117 GB ran
// at full speed (FP32, K = 127).
constexpr double kMaxRingBytes = 120e9;
static int window_depth(int forced, size_t ring_bytes_per_slot, size_t held) {
  if (forced > 0) return std::min(forced, kMaxWindow);
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  // rings reuse the contexts' scratch buffers: what they hold counts as free
  fr += held;
  const double room = std::min(0.75 * (double)fr, kMaxRingBytes);
  long kk = (long)(room / (double)ring_bytes_per_slot) - 2;
  return (int)std::max(1L, std::min((long)kMaxWindow, kk));
}

// Window depth of one filter run's ring (K + 2 slots): window_depth(), lowered
// to what the context's work buffer already holds when that costs at most 15%
// of the depth (growing a buffer of ~100 GB means a free and a fresh
// allocation, seconds on the MI355X).
static int ring_depth(fbse_ctx *c, size_t slot) {
  const int K = window_depth(c->window, slot, c->scratch_bytes);
  if (c->window > 0 || !c->scratch_bytes || slot == 0) return K;
  const long fit = (long)(c->scratch_bytes / slot) - 2;
  return fit >= 1 && fit < K && fit >= (long)(0.85 * K) ? (int)fit : K;
}

// Batch items (start states; real states: pairs) per filter run.  Per Newton
// step and grid point of an item the windowed accumulation reads the ring
// slot (b bytes) and, once per window of K steps, reads and writes the FP64
// accumulators of the ms targets (32 ms / K bytes).  A ring of fewer items per
// slot is deeper, so a large batch runs as consecutive sub-batches of the
// start states (each its own filter of the same Newton length) when that
// lowers the step's traffic; each extra run reads the data its items share
// through the L2 once more (V_loc, the projector records).  With projectors
// a run keeps at least one full column chunk of the projection pass (16
// grids: fewer leave its lanes idle while it still reads every atom's
// tables).  Forced window depths keep the batch whole.
template <typename T> static long filter_items_per_run(fbse_ctx *c, long items, long ms, long M) {
  if (c->window > 0 || items < 2 || M < 3) return items;
  const double b = (double)sizeof(cplx_t<T>), N = (double)c->nsg;
  const double shared = (double)c->ngrid * c->nspin * sizeof(T) +
                        (c->has_proj ? (double)c->pj.npts * (8.0 + 8.0 * sizeof(T)) : 0.0);
  const long min_items = c->has_proj ? std::max(1L, (long)kProjColChunk / c->nspin) : 1;
  long best_items = items;
  double best = 1e300;
  for (long m = 1; m <= std::min(items, 16L); ++m) {
    const long per = (items + m - 1) / m;
    if (m > 1 && ((items + per - 1) / per != m || per < min_items)) continue;  // (the same split as a smaller m)
    const int K = ring_depth(c, (size_t)(per * N * b));
    const double Ke = (double)std::min<long>(K, M - 1);
    const double bytes = (double)items * N * (b + 32.0 * ms / Ke) + (double)(m - 1) * shared;
    if (bytes < best * 0.98) {  // a split must save more than 2%
      best = bytes;
      best_items = per;
    }
  }
  return best_items;
}

template <typename T>
static int filter_impl(fbse_ctx *c, double *d_psi_rank, long n_starts, long ms, long M, const double *an,
                       const double *zn, double Vmin, double dE_1) {
  release_aux(c);
  const long sl = stlen_of(c);
  const long unit = c->is_complex ? 1 : 2;  // real states run as packed pairs
  const long items = (n_starts + unit - 1) / unit;
  const long per_run = c->run_starts > 0 ? (c->run_starts + unit - 1) / unit * unit
                                          : filter_items_per_run<T>(c, items, ms, M) * unit;
  c->last_runs = (n_starts + per_run - 1) / per_run;
  int rc = FBSE_OK;
  for (long s0 = 0; s0 < n_starts && !rc; s0 += per_run) {
    FilterRun<T> r(c, d_psi_rank + s0 * ms * sl, std::min(per_run, n_starts - s0), ms, M, an, zn, Vmin, dE_1);
    rc = r.prepare(ring_depth(c, r.slot_bytes()));
    while (!rc && r.in_window0()) rc = r.step();  // eager (runs the autotuners)
    if (!rc) rc = r.capture();
    while (!rc && r.replays_left()) rc = r.replay();
    while (!rc && r.tail_left()) rc = r.step();
    const int rc2 = r.finish();
    if (!rc) rc = rc2;
  }
  return rc;
}

// Several dots at once: every run on its own context's stream, launches
// interleaved step by step and window by window so that the dots' kernels
// overlap on the device's hardware queues.
// Scheduling units of fbse_filter_many: one dot's run, or a group of runs of
// one grid size whose FFT passes are grouped launches.
struct ManyRunner {
  virtual ~ManyRunner() {}
  virtual int dot() const = 0;  // index of the (first) dot, for error reports
  virtual bool in_window0() const = 0;
  virtual int step() = 0;
  virtual int capture() = 0;
  virtual bool replays_left() const = 0;
  virtual int replay() = 0;
  virtual bool tail_left() const = 0;
  virtual int finish() = 0;
};

template <typename T> struct SingleRunner : ManyRunner {
  FilterRun<T> *r;
  int d;
  SingleRunner(FilterRun<T> *r_, int d_) : r(r_), d(d_) { r->c->last_group = 1; }
  int dot() const override { return d; }
  bool in_window0() const override { return r->in_window0(); }
  int step() override { return r->step(); }
  int capture() override { return r->capture(); }
  bool replays_left() const override { return r->replays_left(); }
  int replay() override { return r->replay(); }
  bool tail_left() const override { return r->tail_left(); }
  int finish() override { return r->finish(); }
};

// Dots of one grid size on the FP32 fixed plane/slab instance: every Newton
// step issues the dots' projections and then ONE launch per FFT pass for all
// of them (apply_h_group), on the first dot's stream, which the other dots'
// streams join for the duration; a full window of all dots is captured into
// one hipGraph.  The dots share M, the window and the step counter position,
// so they advance in lockstep.  Results equal the dots' separate runs.
struct GroupRunner : ManyRunner {
  std::vector<FilterRun<float> *> m;
  std::vector<int> idx;
  std::vector<hipStream_t> saved;
  hipStream_t gs = nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  bool graph = false;
  GroupRunner(std::vector<FilterRun<float> *> runs, std::vector<int> dots) : m(std::move(runs)), idx(std::move(dots)) {}
  ~GroupRunner() override {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  }
  fbse_ctx *c0() const { return m[0]->c; }
  int dot() const override { return idx[0]; }
  // the members' streams join the first one (after their prepare work)
  int begin() {
    gs = c0()->stream;
    graph = true;
    for (auto *r : m) {
      graph = graph && r->graph;
      r->c->last_group = (int)m.size();
      saved.push_back(r->c->stream);
      if (r->c->stream != gs) {
        FBSE_HIP(c0(), hipEventRecord(r->c->ev_fork, r->c->stream));
        FBSE_HIP(c0(), hipStreamWaitEvent(gs, r->c->ev_fork, 0));
        r->c->stream = gs;
      }
    }
    return FBSE_OK;
  }
  int apply(bool first, bool window, int i) {
    const int n = (int)m.size();
    fbse_ctx *cs[kTileGroupMax];
    const float2 *in[kTileGroupMax];
    float2 *out[kTileGroupMax], *work[kTileGroupMax];
    long ns[kTileGroupMax];
    float a1[kTileGroupMax], a0[kTileGroupMax];
    for (int k = 0; k < n; ++k) {
      FilterRun<float> *r = m[k];
      cs[k] = r->c;
      in[k] = window ? r->window_in(i) : r->step_in();
      out[k] = window ? r->window_out(i) : r->step_out();
      work[k] = r->work;
      ns[k] = r->n_items;
      a1[k] = (float)r->dE_1;
      a0[k] = window ? 0.f : (float)r->step_a0();
    }
    return apply_h_group(n, cs, in, out, work, ns, a1, a0, first);
  }
  bool in_window0() const override { return m[0]->in_window0(); }
  int step() override {
    int rc = apply(m[0]->j == 1, false, 0);
    for (size_t k = 0; k < m.size() && !rc; ++k) rc = m[k]->finish_step();
    return rc;
  }
  int capture() override {
    if (!graph) return FBSE_OK;
    int rc;
    for (auto *r : m)
      if ((rc = r->capture_prologue())) return rc;
    FBSE_HIP(c0(), hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
    int crc = FBSE_OK;
    const int R = m[0]->R;
    for (int i = 0; i < R && !crc; ++i) {
      for (auto *r : m) r->capture_step(i);
      crc = apply(false, true, i);
    }
    for (size_t k = 0; k < m.size() && !crc; ++k) crc = m[k]->capture_epilogue();
    hipError_t ee = hipStreamEndCapture(gs, &g);
    for (auto *r : m) r->capture_done();
    hipError_t ei = (crc || ee != hipSuccess) ? hipSuccess : hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (crc) return crc;
    if (ee != hipSuccess || ei != hipSuccess)
      return set_err(c0(), FBSE_ERR_HIP, "grouped Newton window capture: %s",
                     hipGetErrorString(ee != hipSuccess ? ee : ei));
    return FBSE_OK;
  }
  bool replays_left() const override { return graph && ge && m[0]->w_next <= m[0]->n_full; }
  int replay() override {
    FBSE_HIP(c0(), hipGraphLaunch(ge, gs));
    for (auto *r : m) r->replayed();
    return FBSE_OK;
  }
  bool tail_left() const override { return m[0]->tail_left(); }
  int finish() override {
    int rc = FBSE_OK;
    for (auto *r : m) {
      int r2 = r->finish();
      if (r2 && !rc) rc = r2;
    }
    // the members' own streams again, ordered after the group's work
    for (size_t k = 0; k < m.size(); ++k) {
      fbse_ctx *c = m[k]->c;
      if (saved[k] != gs) {
        c->stream = saved[k];
        if (hipEventRecord(c->ev_fork, gs) != hipSuccess || hipStreamWaitEvent(c->stream, c->ev_fork, 0) != hipSuccess)
          if (!rc) rc = set_err(c0(), FBSE_ERR_HIP, "grouped filter: stream join failed");
      }
    }
    return rc;
  }
};

// Dots that may share grouped launches: FP32 fixed plane/slab instance,
// projectors absent or expanded in the plane passes.
static bool groupable(const fbse_ctx *c) { return tile_groupable(c) && (!c->has_proj || proj_fusable(c)); }

template <typename T>
static int filter_many_impl(int n, fbse_ctx *const *ctx, double *const *psi, const long *n_starts, long ms, long M,
                            const double *const *an, const double *const *zn, const double *Vmin,
                            const double *dE_1) {
  std::vector<std::unique_ptr<FilterRun<T>>> runs;
  size_t ring_slot = 0, held = 0;
  int forced = 0;
  for (int d = 0; d < n; ++d) {
    runs.emplace_back(new FilterRun<T>(ctx[d], psi[d], n_starts[d], ms, M, an[d], zn[d], Vmin[d], dE_1[d]));
    ring_slot += runs.back()->slot_bytes();
    release_aux(ctx[d]);
    held += ctx[d]->scratch_bytes;
    if (ctx[d]->window > 0) forced = forced ? std::min(forced, ctx[d]->window) : ctx[d]->window;
  }
  const int K = window_depth(forced, ring_slot, held);
  int rc = FBSE_OK, bad = -1;
  auto fail = [&](int d, int r) {
    if (r && !rc) {
      rc = r;
      bad = d;
    }
  };
  for (int d = 0; d < n && !rc; ++d) fail(d, runs[d]->prepare(K));
  // scheduling units: FP32 dots of one fixed-instance size in groups of up
  // to kTileGroupMax (two or more), every other dot on its own
  std::vector<std::unique_ptr<ManyRunner>> units;
  std::vector<char> taken(n, 0);
  if constexpr (sizeof(T) == 4) {
    for (int d = 0; d < n && !rc; ++d) {
      if (taken[d] || !groupable(ctx[d]) || ctx[d]->profile) continue;
      std::vector<FilterRun<float> *> grp;
      std::vector<int> di;
      for (int e = d; e < n && (int)grp.size() < kTileGroupMax; ++e)
        if (!taken[e] && groupable(ctx[e]) && !ctx[e]->profile && ctx[e]->nx == ctx[d]->nx && ctx[e]->ny == ctx[d]->ny &&
            ctx[e]->nz == ctx[d]->nz && ctx[e]->use_graph == ctx[d]->use_graph) {
          grp.push_back(runs[e].get());
          di.push_back(e);
        }
      if (grp.size() < 2) continue;
      for (int e : di) taken[e] = 1;
      auto *gr = new GroupRunner(grp, di);
      units.emplace_back(gr);
      fail(d, gr->begin());
    }
  }
  for (int d = 0; d < n; ++d)
    if (!taken[d]) units.emplace_back(new SingleRunner<T>(runs[d].get(), d));
  const int nu = (int)units.size();
  for (bool more = true; more && !rc;) {
    more = false;
    for (int u = 0; u < nu && !rc; ++u)
      if (units[u]->in_window0()) {
        fail(units[u]->dot(), units[u]->step());
        more = true;
      }
  }
  for (int u = 0; u < nu && !rc; ++u) fail(units[u]->dot(), units[u]->capture());
  for (bool more = true; more && !rc;) {
    more = false;
    for (int u = 0; u < nu && !rc; ++u)
      if (units[u]->replays_left()) {
        fail(units[u]->dot(), units[u]->replay());
        more = true;
      }
  }
  for (bool more = true; more && !rc;) {
    more = false;
    for (int u = 0; u < nu && !rc; ++u)
      if (units[u]->tail_left()) {
        fail(units[u]->dot(), units[u]->step());
        more = true;
      }
  }
  for (int u = 0; u < nu; ++u) {
    int r = units[u]->finish();
    if (r) fail(units[u]->dot(), r);
  }
  if (rc && bad >= 0 && ctx[0] != ctx[bad])
    set_err(ctx[0], rc, "dot %d: %s", bad, ctx[bad]->err.c_str());
  return rc;
}

extern "C" int fbse_filter_dev(fbse_ctx *c, double *d_psi_rank, long n_starts, long ms, long M, const double *an,
                               const double *zn, double Vmin, double dE_1, int precision) {
  if (!c || n_starts < 0 || ms < 1 || M < 1 || !an || (M > 1 && !zn)) return FBSE_ERR_ARG;
  if (n_starts == 0) return FBSE_OK;
  FBSE_HIP(c, hipSetDevice(c->device));
  if (precision == FBSE_PREC_FP64) return filter_impl<double>(c, d_psi_rank, n_starts, ms, M, an, zn, Vmin, dE_1);
  if (precision == FBSE_PREC_MIXED) return filter_impl<float>(c, d_psi_rank, n_starts, ms, M, an, zn, Vmin, dE_1);
  return FBSE_ERR_ARG;
}
extern "C" int fbse_filter_many(int n_dots, fbse_ctx *const *ctx, double *const *d_psi_rank, const long *n_starts,
                                long ms, long M, const double *const *an, const double *const *zn,
                                const double *Vmin, const double *dE_1, int precision) {
  if (n_dots < 1 || !ctx || !d_psi_rank || !n_starts || !an || !zn || !Vmin || !dE_1 || ms < 1 || M < 1)
    return FBSE_ERR_ARG;
  for (int d = 0; d < n_dots; ++d) {
    if (!ctx[d] || !d_psi_rank[d] || n_starts[d] < 1 || !an[d] || (M > 1 && !zn[d])) return FBSE_ERR_ARG;
    if (ctx[d]->device != ctx[0]->device)
      return set_err(ctx[0], FBSE_ERR_UNSUPPORTED, "dot %d is on device %d, dot 0 on %d", d, ctx[d]->device,
                     ctx[0]->device);
    for (int e = 0; e < d; ++e)
      if (ctx[e] == ctx[d]) return set_err(ctx[0], FBSE_ERR_ARG, "dots %d and %d share a context", e, d);
  }
  FBSE_HIP(ctx[0], hipSetDevice(ctx[0]->device));
  if (precision == FBSE_PREC_FP64)
    return filter_many_impl<double>(n_dots, ctx, d_psi_rank, n_starts, ms, M, an, zn, Vmin, dE_1);
  if (precision == FBSE_PREC_MIXED)
    return filter_many_impl<float>(n_dots, ctx, d_psi_rank, n_starts, ms, M, an, zn, Vmin, dE_1);
  return FBSE_ERR_ARG;
}
extern "C" int fbse_filter(fbse_ctx *c, double *psi_rank, long n_starts, long ms, long M, const double *an,
                           const double *zn, double Vmin, double dE_1, int precision) {
  if (!c || !psi_rank) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const size_t bytes = (size_t)n_starts * ms * stlen_of(c) * sizeof(double);
  double *d = nullptr;
  FBSE_HIP(c, hipMalloc(&d, bytes ? bytes : 16));
  int rc = FBSE_OK;
  hipError_t e = hipMemcpyAsync(d, psi_rank, bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) rc = set_err(c, FBSE_ERR_HIP, "copy: %s", hipGetErrorString(e));
  if (!rc) rc = fbse_filter_dev(c, d, n_starts, ms, M, an, zn, Vmin, dE_1, precision);
  if (!rc) {
    e = hipMemcpyAsync(psi_rank, d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) rc = set_err(c, FBSE_ERR_HIP, "copy: %s", hipGetErrorString(e));
  }
  hipStreamSynchronize(c->stream);
  hipFree(d);
  return rc;
}

// ---------------------------------------------------------------------------
// Orthogonalisation and Rayleigh-Ritz (FP64, rocBLAS GEMMs + rocSOLVER)
// ---------------------------------------------------------------------------
__global__ void hermitize_kernel(double *H, long n, int cplx) {
  // H <- (H + H^H)/2, column-major
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n * n; idx += (long)gridDim.x * blockDim.x) {
    const long i = idx % n, j = idx / n;
    if (i > j) continue;
    if (cplx) {
      double2 *A = (double2 *)H;
      double2 a = A[i + j * n], b = A[j + i * n];
      double2 r = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      A[i + j * n] = r;
      A[j + i * n] = make_double2(r.x, -r.y);
    } else {
      const double r = 0.5 * (H[i + j * n] + H[j + i * n]);
      H[i + j * n] = r;
      H[j + i * n] = r;
    }
  }
}

// C[i + j*ldc] = sum_p part[p][i + j*n1] (p ascending), doubles of cw per
// element; tot = n1 * n2 * cw doubles per partial.
__global__ void sum_parts_kernel(const double *part, long parts, long n1, long tot, int cw, double *C, long ldc) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < tot; idx += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (long p = 0; p < parts; ++p) acc += part[p * tot + idx];
    const long e = idx / cw, t = idx - e * cw;
    const long i = e % n1, j = e / n1;
    C[(i + j * ldc) * cw + t] = acc;
  }
}

// Lower triangle of the column-major n x n matrix H from its upper triangle:
// H[i, j] = conj(H[j, i]) for i > j.
__global__ void mirror_upper_kernel(double *H, long n, int cplx) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n * n; idx += (long)gridDim.x * blockDim.x) {
    const long i = idx % n, j = idx / n;
    if (i <= j) continue;
    if (cplx) {
      const double2 u = ((const double2 *)H)[j + i * n];
      ((double2 *)H)[i + j * n] = make_double2(u.x, -u.y);
    } else {
      H[i + j * n] = H[j + i * n];
    }
  }
}

// Split-K target of the Gram products: about four work-groups per CU (two
// are resident at a time), slabs of at least this many rows.
static const long kGramGroups = 2048;
static const long kGramMinRows = 4096;

struct LinAlg {
  fbse_ctx *c;
  bool cplx;
  long m;   // rows (complex elements or reals)
  size_t es() const { return cplx ? 16 : 8; }
  int cw() const { return cplx ? 2 : 1; }  // doubles per element
  // C (n1 x n2, ldc) = A^H B over the m rows (A, B with leading dimension m)
  // on the FP64 matrix cores (gemm_kernels.hip).  Each work-group forms one
  // 64 x 64 output tile over one slab of rows; with few tiles the rows are
  // split into P slabs whose partial Grams are summed in a fixed order, so
  // results stay deterministic.  upper (n1 == n2): only the tiles on or above
  // the diagonal are formed (the lower triangle of C is then left to the caller).
  int gram_core(const double *A, long n1, const double *B, long n2, double *Cm, long ldc, bool upper) {
    const long tiles = fbse::gram_tile_count(n1, n2, upper);
    long P = std::max(1L, std::min((kGramGroups + tiles - 1) / tiles, m / kGramMinRows));
    long slab = (m + P - 1) / P;
    slab = (slab + 15) / 16 * 16;
    P = (m + slab - 1) / slab;
    if (P <= 1) {
      if (fbse::launch_gram_mfma(c->stream, cplx, A, m, B, m, m, n1, n2, upper, m, 1, Cm, ldc, 0))
        return set_err(c, FBSE_ERR_HIP, "Gram kernel launch (%ld x %ld over %ld rows)", n1, n2, m);
      return FBSE_OK;
    }
    const long blk = n1 * n2;  // elements of one partial Gram
    int rc = grow(c, &c->d_gram, &c->gram_bytes, (size_t)P * blk * es());
    if (rc) return rc;
    double *part = (double *)c->d_gram;
    if (fbse::launch_gram_mfma(c->stream, cplx, A, m, B, m, m, n1, n2, upper, slab, P, part, n1, blk))
      return set_err(c, FBSE_ERR_HIP, "Gram kernel launch (%ld x %ld over %ld rows, %ld slabs)", n1, n2, m, P);
    const long tot = blk * cw();
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)std::min(2048L, (tot + 255) / 256)), dim3(256), 0, c->stream,
                       (const double *)part, P, n1, tot, cw(), Cm, ldc);
    FBSE_HIP(c, hipGetLastError());
    return FBSE_OK;
  }
  // C(n1 x n2, ldc) = A^H B   with A (m x n1), B (m x n2), lda = ldb = m.
  int gram(const double *A, long n1, const double *B, long n2, double *Cm, long ldc) {
    return gram_core(A, n1, B, n2, Cm, ldc, false);
  }
  // G (n x n) = A^H B for a product known to be Hermitian (B = A, or B = H A).
  // Only the tiles on and above the diagonal are formed, then the lower
  // triangle is mirrored from the upper one.
  int gram_upper(const double *A, const double *B, long n, double *G) {
    int rc = gram_core(A, n, B, n, G, n, true);
    if (rc) return rc;
    hipLaunchKernelGGL(mirror_upper_kernel, dim3(256), dim3(256), 0, c->stream, G, n, cplx ? 1 : 0);
    FBSE_HIP(c, hipGetLastError());
    return FBSE_OK;
  }
  int gram_self(const double *A, long n, double *G) { return gram_upper(A, A, n, G); }
  // Out (m x nout) (+)= A[:, 0:nin] * V (nin x nout, ldv), on the FP64 matrix cores
  int product(const double *A, long nin, const double *V, long ldv, long nout, double *Out, bool acc = false) {
    if (fbse::launch_rot_mfma(c->stream, cplx, A, m, m, nin, V, ldv, nout, Out, m, acc))
      return set_err(c, FBSE_ERR_HIP, "rotation kernel launch (%ld rows, %ld -> %ld)", m, nin, nout);
    return FBSE_OK;
  }
  // Out (r x nout) = X (r x r2, ldx = r) * Y (r2 x nout, ldy = r2): small dense products
  int small_product(const double *X, long r, long r2, const double *Y, long nout, double *Out) {
    if (cplx) {
      const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
      FBSE_BLAS(c, rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none, (int)r, (int)nout, (int)r2,
                                 &one, (const rocblas_double_complex *)X, (int)r, (const rocblas_double_complex *)Y,
                                 (int)r2, &zero, (rocblas_double_complex *)Out, (int)r));
    } else {
      const double one = 1.0, zero = 0.0;
      FBSE_BLAS(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, (int)r, (int)nout, (int)r2,
                                 &one, X, (int)r, Y, (int)r2, &zero, Out, (int)r));
    }
    return FBSE_OK;
  }
  // A[:, 0:nout] = A[:, 0:nin] * V (nin x nout, ldv), in place by row blocks:
  // each block's product goes to a scratch block (at most 1 GiB) on the FP64
  // matrix cores, then back into A
  int rotate(double *A, long nin, const double *V, long ldv, long nout) {
    long rb = (long)((1ull << 30) / ((size_t)nout * es()));
    rb = std::max(64L, std::min(rb, m)) / 64 * 64;
    if (rb > m) rb = m;
    int rc = ensure_scratch2(c, (size_t)rb * nout * es());
    if (rc) return rc;
    for (long r0 = 0; r0 < m; r0 += rb) {
      const long nr = std::min(rb, m - r0);
      if (fbse::launch_rot_mfma(c->stream, cplx, A + (size_t)r0 * cw(), m, nr, nin, V, ldv, nout,
                                (double *)c->d_scratch2, nr))
        return set_err(c, FBSE_ERR_HIP, "rotation kernel launch (%ld rows, %ld -> %ld)", nr, nin, nout);
      char *dst = (char *)A + (size_t)r0 * es();
      FBSE_HIP(c, hipMemcpy2DAsync(dst, (size_t)m * es(), c->d_scratch2, (size_t)nr * es(), (size_t)nr * es(), nout,
                                   hipMemcpyDeviceToDevice, c->stream));
    }
    return FBSE_OK;
  }
  // eigen-decomposition of the Hermitian n x n matrix H (in place -> vectors)
  int eigh(double *H, long n, double *d_w) {
    double *d_e = nullptr;
    int *d_info = nullptr;
    FBSE_HIP(c, hipMalloc(&d_e, n * 8));
    FBSE_HIP(c, hipMalloc(&d_info, sizeof(int)));
    rocblas_status st;
    if (cplx)
      st = rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, (int)n, (rocblas_double_complex *)H,
                            (int)n, d_w, d_e, d_info);
    else
      st = rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, (int)n, H, (int)n, d_w, d_e, d_info);
    int info = 0;
    hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    hipFree(d_e);
    hipFree(d_info);
    if (st != rocblas_status_success || info != 0)
      return set_err(c, FBSE_ERR_SOLVER, "eigensolver failed: status %d info %d (n=%ld)", (int)st, info, n);
    return FBSE_OK;
  }
};

static int ortho_dev_impl(fbse_ctx *c, double *A, long n, double svd_eps, long *rank, double *sing) {
  LinAlg la{c, c->is_complex != 0, c->nsg};
  const size_t es = la.es();
  int rc;
  double *G = nullptr, *w = nullptr, *Z = nullptr;
  FBSE_HIP(c, hipMalloc(&G, (size_t)n * n * es));
  FBSE_HIP(c, hipMalloc(&Z, (size_t)n * n * es));
  FBSE_HIP(c, hipMalloc(&w, (size_t)n * 8));
  std::vector<double> hw(n);
  long k = 0;
  auto cleanup = [&]() {
    hipStreamSynchronize(c->stream);
    hipFree(G);
    hipFree(Z);
    hipFree(w);
  };
  // pass 1: G = A^H A -> V (columns of B = A V are orthogonal, |B_i| = s_i)
  if ((rc = la.gram_self(A, n, G)) || (rc = la.eigh(G, n, w))) {
    cleanup();
    return rc;
  }
  // The first Gram eigensolve resolves eigenvalues only to eps*||G||, i.e.
  // singular values down to ~1e-8 s_0.  Columns of B = A V whose eigenvalue is
  // below kSmall*lambda_max (the leading, ascending ones) are formed
  // explicitly, B_s = A V_s, and re-measured by a second Gram over that block
  // alone, where the eigensolver's error scales with the block's own norm;
  // this resolves s_i/s_0 to ~1e-12, enough for the 1e-10 cut of
  // ortho.c:68-75.
  FBSE_HIP(c, hipMemcpyAsync(hw.data(), w, n * 8, hipMemcpyDeviceToHost, c->stream));
  FBSE_HIP(c, hipStreamSynchronize(c->stream));
  const double kSmall = 1e-8;
  const double lmax = std::max(hw[n - 1], 0.0);
  long ns = 0;
  while (ns < n && hw[ns] < kSmall * lmax) ++ns;
  std::vector<double> mu;
  if (ns > 0) {
    double *Bs = nullptr;
    // from the work buffer (the filter's ring) when it holds the block
    const bool in_arena = arena_states(c, 0, (size_t)la.m * es) >= ns;
    if (in_arena) {
      Bs = arena_at(c, 0);
    } else if (big_malloc(c, (void **)&Bs, (size_t)la.m * ns * es) != hipSuccess) {
      cleanup();
      return set_err(c, FBSE_ERR_NOMEM, "ortho: %ld x %ld block of small singular directions", la.m, ns);
    }
    rc = la.product(A, n, G, n, ns, Bs);
    if (!rc) rc = la.gram_self(Bs, ns, Z);
    hipStreamSynchronize(c->stream);
    if (!in_arena) hipFree(Bs);
    if (rc) {
      cleanup();
      return rc;
    }
    hipLaunchKernelGGL(hermitize_kernel, dim3(256), dim3(256), 0, c->stream, Z, ns, la.cplx ? 1 : 0);
    if ((rc = la.eigh(Z, ns, w))) {
      cleanup();
      return rc;
    }
    mu.resize(ns);
    FBSE_HIP(c, hipMemcpyAsync(mu.data(), w, ns * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
  }
  // singular values, descending: the large block from pass 1, then the small
  // block from pass 2 (all of which are below every large one)
  std::vector<double> sv(n);
  for (long i = 0; i < n - ns; ++i) sv[i] = sqrt(std::max(0.0, hw[n - 1 - i]));
  for (long i = 0; i < ns; ++i) sv[n - ns + i] = sqrt(std::max(0.0, mu[ns - 1 - i]));
  if (sing)
    for (long i = 0; i < n; ++i) sing[i] = sv[i];
  k = n;
  for (long i = 0; i < n; ++i)
    if (sv[0] == 0.0 || (sv[i] / sv[0]) < svd_eps) {
      k = i;
      break;
    }
  *rank = k;
  if (k == 0) {
    cleanup();
    return set_err(c, FBSE_ERR_SOLVER, "ortho: all singular values vanish");
  }
  // Zsel (n x k) in the basis of B's columns: large column i -> e_col/s ;
  // small kept j -> W_s[:, j]/s.  U = B Zsel = A (V Zsel): one pass over A.
  {
    std::vector<double> zh((size_t)n * k * (la.cplx ? 2 : 1), 0.0);
    std::vector<double> ws;
    const long nsk = std::max(0L, k - (n - ns));
    if (nsk > 0) {
      ws.resize((size_t)ns * ns * (la.cplx ? 2 : 1));
      FBSE_HIP(c, hipMemcpyAsync(ws.data(), Z, ws.size() * 8, hipMemcpyDeviceToHost, c->stream));
      FBSE_HIP(c, hipStreamSynchronize(c->stream));
    }
    const int cw = la.cplx ? 2 : 1;
    for (long i = 0; i < k; ++i) {
      const double inv = 1.0 / sv[i];
      if (i < n - ns) {
        const long col = n - 1 - i;  // column of B holding this singular value
        zh[((size_t)i * n + col) * cw] = inv;
      } else {
        const long j = ns - 1 - (i - (n - ns));  // eigenvector index in the small block
        for (long r = 0; r < ns; ++r)
          for (int t = 0; t < cw; ++t) zh[((size_t)i * n + r) * cw + t] = ws[((size_t)j * ns + r) * cw + t] * inv;
      }
    }
    FBSE_HIP(c, hipMemcpyAsync(Z, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, c->stream));
    // W = V Zsel (n x k)
    double *W = nullptr;
    FBSE_HIP(c, hipMalloc(&W, (size_t)n * k * es));
    rc = la.small_product(G, n, n, Z, k, W);
    if (!rc) rc = la.rotate(A, n, W, n, k);
    hipStreamSynchronize(c->stream);
    hipFree(W);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  // pass 3: Cholesky re-orthogonalisation U <- U R^-1 with U^H U = R^H R
  if ((rc = la.gram_self(A, k, G))) {
    cleanup();
    return rc;
  }
  {
    int *d_info = nullptr;
    FBSE_HIP(c, hipMalloc(&d_info, sizeof(int)));
    rocblas_status st;
    if (la.cplx)
      st = rocsolver_zpotrf(c->blas, rocblas_fill_upper, (int)k, (rocblas_double_complex *)G, (int)k, d_info);
    else
      st = rocsolver_dpotrf(c->blas, rocblas_fill_upper, (int)k, G, (int)k, d_info);
    int info = 0;
    hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    hipFree(d_info);
    if (st != rocblas_status_success || info != 0) {
      cleanup();
      return set_err(c, FBSE_ERR_SOLVER, "ortho: Cholesky re-orthogonalisation failed (info %d)", info);
    }
    // right-side solves act on rows independently: blocks of rows keep the
    // workspace rocBLAS allocates (and its handle then keeps) at rows x k
    // elements instead of the whole block (137 GB for 256 states of 256^3)
    constexpr long kTrsmRows = 262144;
    for (long r0 = 0; r0 < la.m && st == rocblas_status_success; r0 += kTrsmRows) {
      const int nr = (int)std::min(kTrsmRows, (long)la.m - r0);
      if (la.cplx) {
        const rocblas_double_complex one = {1.0, 0.0};
        st = rocblas_ztrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, nr, (int)k, &one, (const rocblas_double_complex *)G, (int)k,
                           (rocblas_double_complex *)A + r0, (int)la.m);
      } else {
        const double one = 1.0;
        st = rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                           rocblas_diagonal_non_unit, nr, (int)k, &one, G, (int)k, A + r0, (int)la.m);
      }
    }
    if (st != rocblas_status_success) {
      cleanup();
      return set_err(c, FBSE_ERR_SOLVER, "ortho: trsm status %d", (int)st);
    }
  }
  cleanup();
  return normalize_dev_impl(c, A, k);
}

extern "C" int fbse_ortho_dev(fbse_ctx *c, double *d_psitot, long n_cols, double svd_eps, long *rank,
                              double *sing) {
  if (!c || !d_psitot || n_cols < 1 || !rank) return FBSE_ERR_ARG;
  if (c->nsg > 0x7fffffffL) return FBSE_ERR_UNSUPPORTED;
  FBSE_HIP(c, hipSetDevice(c->device));
  return ortho_dev_impl(c, d_psitot, n_cols, svd_eps, rank, sing);
}
extern "C" int fbse_ortho(fbse_ctx *c, double *psitot, long n_cols, double svd_eps, long *rank, double *sing) {
  if (!c || !psitot || n_cols < 1 || !rank) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  return with_device_copy(c, psitot, n_cols, [&](double *d) -> int {
    int rc = ortho_dev_impl(c, d, n_cols, svd_eps, rank, sing);
    if (rc) return rc;
    FBSE_HIP(c, hipMemcpyAsync(psitot, d, (size_t)n_cols * sl * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    return FBSE_OK;
  });
}

// States of per_state bytes that fit in frac of the free device memory (>= 1).
static long states_that_fit(const fbse_ctx *c, size_t per_state, double frac) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 1;
  return std::max(1L, (long)(frac * (double)fr / (double)per_state));
}

static int rr_dev_impl(fbse_ctx *c, double *U, long n, double *eval) {
  LinAlg la{c, c->is_complex != 0, c->nsg};
  const long sl = stlen_of(c);
  const size_t es = la.es();
  int rc = FBSE_OK;
  double *Hs = nullptr, *w = nullptr, *HU = nullptr;
  // H U is held for as many columns as memory allows: U^H (H U) over a narrow
  // block of columns runs far below the GEMM peak, so one wide product is used
  // whenever it fits (1024 states of a 96^3 spinor grid take 29 GB).  It comes
  // from the work buffer when that holds at least as many columns as a fresh
  // block from the free memory would.
  const size_t hw = h_work_bytes(c, n);
  long ch = std::min(n, states_that_fit(c, (size_t)sl * 8, 0.5));
  const long ch_arena = std::min(n, arena_states(c, hw, (size_t)sl * 8));
  const bool in_arena = ch_arena >= 1 && ch_arena >= ch;
  if (in_arena) {
    ch = ch_arena;
  } else if (ch < n && c->scratch_bytes) {  // the filter's ring is no longer needed: make room for one wide product
    release_work(c);
    ch = std::min(n, states_that_fit(c, (size_t)sl * 8, 0.5));
  }
  FBSE_HIP(c, hipMalloc(&Hs, (size_t)n * n * es));
  FBSE_HIP(c, hipMalloc(&w, (size_t)n * 8));
  if (in_arena) {
    // the FFT work space of apply_h_dev_impl stays at the front of the buffer
    if ((rc = ensure_scratch(c, hw))) {
      hipFree(Hs);
      hipFree(w);
      return rc;
    }
    HU = arena_at(c, hw);
  } else {
    FBSE_HIP(c, big_malloc(c, (void **)&HU, (size_t)ch * sl * 8));
  }
  if (ch == n) {
    // all of H U at once: U^H (H U) is Hermitian, its upper blocks suffice
    rc = apply_h_dev_impl(c, U, HU, n);
    if (!rc) rc = la.gram_upper(U, HU, n, Hs);
  } else {
    for (long c0 = 0; c0 < n && !rc; c0 += ch) {
      const long nc = std::min(ch, n - c0);
      rc = apply_h_dev_impl(c, U + c0 * sl, HU, nc);
      if (!rc) rc = la.gram(U, n, HU, nc, (double *)((char *)Hs + (size_t)c0 * n * es), n);
    }
    if (!rc) hipLaunchKernelGGL(hermitize_kernel, dim3(256), dim3(256), 0, c->stream, Hs, n, la.cplx ? 1 : 0);
  }
  if (!rc) rc = la.eigh(Hs, n, w);
  if (!rc) rc = la.rotate(U, n, Hs, n, n);
  if (!rc) {
    std::vector<double> wh(n);
    hipMemcpyAsync(wh.data(), w, n * 8, hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    for (long i = 0; i < n; ++i) eval[i] = wh[i] * c->dv;
  }
  hipStreamSynchronize(c->stream);
  hipFree(Hs);
  hipFree(w);
  if (!in_arena) hipFree(HU);
  if (rc) return rc;
  return normalize_dev_impl(c, U, n);
}
extern "C" int fbse_rayleigh_ritz_dev(fbse_ctx *c, double *d_psitot, long n, double *eval) {
  if (!c || !d_psitot || n < 1 || !eval) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  return rr_dev_impl(c, d_psitot, n, eval);
}
extern "C" int fbse_rayleigh_ritz(fbse_ctx *c, double *psitot, long n, double *eval) {
  if (!c || !psitot || n < 1 || !eval) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  return with_device_copy(c, psitot, n, [&](double *d) -> int {
    int rc = rr_dev_impl(c, d, n, eval);
    if (rc) return rc;
    FBSE_HIP(c, hipMemcpyAsync(psitot, d, (size_t)n * sl * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    return FBSE_OK;
  });
}

// Subspace products on the FP64 matrix cores, for callers that form their
// own subspace matrices (the dotp / dotpreal loops of diag_H, Hmat.c:141-164,
// 301-338, and its rotation into the eigenvectors, Hmat.c:230-281).
extern "C" int fbse_overlap_dev(fbse_ctx *c, const double *d_a, long n1, const double *d_b, long n2, double *s) {
  if (!c || !d_a || !d_b || !s || n1 < 1 || n2 < 1) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  LinAlg la{c, c->is_complex != 0, c->nsg};
  const size_t bytes = (size_t)n1 * n2 * la.es();
  double *d_s = nullptr;
  FBSE_HIP(c, hipMalloc(&d_s, bytes));
  int rc = la.gram(d_a, n1, d_b, n2, d_s, n1);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(s, d_s, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = set_err(c, FBSE_ERR_HIP, "overlap copy: %s", hipGetErrorString(e));
  }
  hipStreamSynchronize(c->stream);
  hipFree(d_s);
  if (rc) return rc;
  const size_t nd = (size_t)n1 * n2 * la.cw();
  for (size_t i = 0; i < nd; ++i) s[i] *= c->dv;
  return FBSE_OK;
}

extern "C" int fbse_rotate_dev(fbse_ctx *c, double *d_a, long nin, const double *v, long nout) {
  if (!c || !d_a || !v || nin < 1 || nout < 1 || nout > nin) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  LinAlg la{c, c->is_complex != 0, c->nsg};
  const size_t bytes = (size_t)nin * nout * la.es();
  double *d_v = nullptr;
  FBSE_HIP(c, hipMalloc(&d_v, bytes));
  int rc = FBSE_OK;
  hipError_t e = hipMemcpyAsync(d_v, v, bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) rc = set_err(c, FBSE_ERR_HIP, "rotation matrix copy: %s", hipGetErrorString(e));
  if (!rc) rc = la.rotate(d_a, nin, d_v, nin, nout);
  hipStreamSynchronize(c->stream);
  hipFree(d_v);
  return rc;
}

// ---------------------------------------------------------------------------
// FP64 residual refinement of the Ritz pairs in a window (mixed-precision path)
// ---------------------------------------------------------------------------
int fbse::precondition_dev(fbse_ctx *c, const double *d_r, double *d_out, long n, double e_p) {
  int rc;
  const long ch = h_chunk(c);
  const size_t per = (size_t)c->nsg * sizeof(double2);
  if (c->is_complex) {
    if ((rc = ensure_scratch(c, (size_t)std::min(ch, n) * per))) return rc;
    for (long s0 = 0; s0 < n; s0 += ch) {
      const long ns = std::min(ch, n - s0);
      if ((rc = precondition_cplx<double>(c, reinterpret_cast<const double2 *>(d_r) + s0 * c->nsg,
                                          reinterpret_cast<double2 *>(d_out) + s0 * c->nsg, (double2 *)c->d_scratch,
                                          ns, e_p)))
        return rc;
    }
  } else {
    const long chp = std::max(2L, ch & ~1L);
    const long np = (std::min(chp, n) + 1) / 2;
    if ((rc = ensure_scratch(c, (size_t)np * per * 2))) return rc;
    if ((rc = ensure_scratch2(c, (size_t)np * per))) return rc;
    double2 *packed = (double2 *)c->d_scratch, *work = packed + np * c->nsg, *res = (double2 *)c->d_scratch2;
    for (long s0 = 0; s0 < n; s0 += chp) {
      const long ns = std::min(chp, n - s0);
      if ((rc = pack_real(c, d_r + s0 * c->nsg, c->nsg, ns, packed))) return rc;
      if ((rc = precondition_cplx<double>(c, packed, res, work, (ns + 1) / 2, e_p))) return rc;
      if ((rc = unpack_real(c, res, ns, d_out + s0 * c->nsg, c->nsg))) return rc;
    }
  }
  return FBSE_OK;
}

// M[i + i*ld] = d[i] for i < n (raw diagonal of the Rayleigh-Ritz block)
__global__ void set_diag_kernel(double *M, long ld, long n, const double *d, int cplx) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (cplx)
      ((double2 *)M)[i + i * ld] = make_double2(d[i], 0.0);
    else
      M[i + i * ld] = d[i];
  }
}

// G[b] (nt x nt, column-major) = dv (M2 - 2 theta_b M) + theta_b^2 I for the
// raw Rayleigh-Ritz matrix M = Q^H H Q and M2 = (H Q)^H (H Q) of a dv-orthonormal
// basis Q: the matrix of ||(H - theta_b) Q y||^2 (dv-weighted) over unit y.
__global__ void refined_matrix_kernel(const double *M, const double *M2, const double *theta, long nt, long nb,
                                      double dv, int cplx, double *G) {
  const long per = nt * nt, tot = per * nb;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < tot; idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / per, e = idx - b * per;
    const long i = e % nt, j = e / nt;
    const double th = theta[b];
    if (cplx) {
      const double2 m = ((const double2 *)M)[e], m2 = ((const double2 *)M2)[e];
      double re = dv * (m2.x - 2.0 * th * m.x), im = dv * (m2.y - 2.0 * th * m.y);
      if (i == j) {
        re += th * th;
        im = 0.0;
      }
      ((double2 *)G)[idx] = make_double2(re, im);
    } else {
      double v = dv * (M2[e] - 2.0 * th * M[e]);
      if (i == j) v += th * th;
      G[idx] = v;
    }
  }
}

// One guarded FP64 refinement sweep over the Ritz pairs of a preceding
// Rayleigh-Ritz (refined Ritz vectors over a Davidson-extended basis):
//   1. the k selected pairs (eval in [e_lo, e_hi], sigma < sigma_max, lowest
//      sigma first; k limited by memory) get residuals r = H u - theta u and
//      corrections w = P r, P = 1 / (1 + 0.5|k|^2 / e_p) in k-space;
//   2. the corrections are orthogonalised against all n Ritz vectors (twice)
//      and among themselves (rank cut), giving Q = [U, W];
//   3. A1 = Q^H H Q and A2 = (H Q)^H (H Q) on the matrix-core Grams;
//   4. for each selected pair two candidates: its Rayleigh-Ritz vector over
//      span(Q) (when that carries more than half its weight on the pair), and
//      its refined vector, the unit vector with the least ||(H - theta) v||
//      (lowest eigenvectors of A2 - 2 theta A1 + theta^2, near-degenerate
//      pairs together, theta updated twice to the refined vectors' Rayleigh
//      quotient); since u itself lies in span(Q), the refined residual is never
//      above the pair's sigma, also where the spectrum is too dense for
//      interior Rayleigh-Ritz to help; the lower predicted residual wins;
//   5. a candidate that would duplicate another Ritz vector or an earlier
//      candidate (overlap above one half) is dropped, and a pair is replaced
//      only when its new vector's own residual is lower.
// Every pair stays in its slot; none gets a larger residual.
static int refine_impl(fbse_ctx *c, double *d_psi, long n, const double *eval, const double *sigma,
                       double sigma_max, double e_lo, double e_hi, double e_p, double *eval_out, double *sigma_out,
                       long *n_replaced) {
  LinAlg la{c, c->is_complex != 0, c->nsg};
  const long sl = stlen_of(c);
  const size_t es = la.es();
  const int cw = la.cw();
  const double dv = c->dv;
  for (long i = 0; i < n; ++i) {
    eval_out[i] = eval[i];
    sigma_out[i] = sigma[i];
  }
  *n_replaced = 0;
  std::vector<long> sel;
  for (long i = 0; i < n; ++i)
    if (eval[i] >= e_lo && eval[i] <= e_hi && sigma[i] < sigma_max) sel.push_back(i);
  if (sel.empty()) return FBSE_OK;
  std::stable_sort(sel.begin(), sel.end(), [&](long a, long b) { return sigma[a] < sigma[b]; });
  // two blocks of k states (corrections W and H W / the new vectors), later
  // two chunks of H U and H^2 U: from the work buffer (the filter's ring, not
  // needed any more) past the H applications' FFT work space when it holds
  // four blocks of min(k, 32) states or more, else freshly allocated after
  // the ring is released
  const size_t hw = h_work_bytes(c, n), st_b = (size_t)sl * 8;
  const long want = (long)sel.size();
  const long k_arena = arena_states(c, hw, st_b) / 4;
  const bool in_arena = k_arena >= std::min(want, 32L);
  long k;
  if (in_arena) {
    k = std::min(want, k_arena);
  } else {
    if (c->scratch_bytes) release_work(c);
    k = std::max(1L, std::min(want, states_that_fit(c, 2 * st_b, 0.6)));
  }
  sel.resize(k);
  int rc = FBSE_OK;
  double *X = nullptr, *Y = nullptr, *d_th = nullptr, *Cm = nullptr, *M = nullptr, *d_w = nullptr, *M2 = nullptr;
  double *T1 = nullptr, *T2 = nullptr, *G = nullptr, *Gw = nullptr, *Ge = nullptr;
  int *Ginfo = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(c->stream);
    for (double *p : {d_th, Cm, M, d_w, M2, G, Gw, Ge}) hipFree(p);
    if (!in_arena)
      for (double *p : {X, Y, T1, T2}) hipFree(p);
    hipFree(Ginfo);
  };
  if (in_arena) {
    if ((rc = ensure_scratch(c, hw))) return rc;  // (a no-op: the buffer is larger)
    X = arena_at(c, hw);
    Y = X + (size_t)k * sl;
  } else if (big_malloc(c, (void **)&X, (size_t)k * st_b) != hipSuccess ||
             big_malloc(c, (void **)&Y, (size_t)k * st_b) != hipSuccess) {
    cleanup();
    return set_err(c, FBSE_ERR_NOMEM, "refine: two blocks of %ld states", k);
  }
  std::vector<double> th(std::max(n, k));
  for (long j = 0; j < k; ++j) th[j] = eval[sel[j]];
#define REF_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      rc = set_err(c, FBSE_ERR_HIP, "refine: %s: %s", #call, hipGetErrorString(e_));                 \
      cleanup();                                                                                       \
      return rc;                                                                                       \
    }                                                                                                  \
  } while (0)
#define REF_RC(expr)     \
  do {                   \
    if ((rc = (expr))) { \
      cleanup();         \
      return rc;         \
    }                    \
  } while (0)
  REF_HIP(hipMalloc(&d_th, (size_t)std::max(n, k) * 8));
  REF_HIP(hipMemcpyAsync(d_th, th.data(), k * 8, hipMemcpyHostToDevice, c->stream));
  // 1. residuals of the selected pairs and their corrections W (in X)
  for (long j = 0; j < k; ++j)
    REF_HIP(hipMemcpyAsync(X + j * sl, d_psi + sel[j] * sl, sl * 8, hipMemcpyDeviceToDevice, c->stream));
  REF_RC(apply_h_dev_impl(c, X, Y, k));
  REF_RC(states_sub_scaled(c, X, Y, k, sl, d_th));
  REF_RC(precondition_dev(c, Y, X, k, e_p));
  // 2. W <- W - U (dv U^H W), twice; then orthonormal W Z (W^H W dv = 1)
  REF_HIP(hipMalloc(&Cm, (size_t)std::max(n, k) * k * es));
  for (int pass = 0; pass < 2; ++pass) {
    REF_RC(la.gram(d_psi, n, X, k, Cm, n));
    std::vector<double> ch((size_t)n * k * cw);
    REF_HIP(hipMemcpyAsync(ch.data(), Cm, ch.size() * 8, hipMemcpyDeviceToHost, c->stream));
    REF_HIP(hipStreamSynchronize(c->stream));
    for (auto &v : ch) v *= -dv;
    REF_HIP(hipMemcpyAsync(Cm, ch.data(), ch.size() * 8, hipMemcpyHostToDevice, c->stream));
    REF_RC(la.product(d_psi, n, Cm, n, k, X, true));
    REF_HIP(hipStreamSynchronize(c->stream));
  }
  REF_RC(la.gram_self(X, k, Cm));
  REF_HIP(hipMalloc(&d_w, (size_t)std::max(n, k) * 8));
  REF_RC(la.eigh(Cm, k, d_w));
  std::vector<double> lam(k), q((size_t)k * k * cw);
  REF_HIP(hipMemcpyAsync(lam.data(), d_w, k * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipMemcpyAsync(q.data(), Cm, q.size() * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipStreamSynchronize(c->stream));
  const double lmax = std::max(lam[k - 1], 0.0);
  long k0 = 0;  // eigenvalues ascending: drop directions at the rounding level
  while (k0 < k && lam[k0] <= 1e-10 * lmax) ++k0;
  const long k2 = k - k0;
  if (k2 == 0 || lmax == 0.0) {
    cleanup();
    return FBSE_OK;
  }
  {
    std::vector<double> z((size_t)k * k2 * cw);
    for (long jj = 0; jj < k2; ++jj) {
      const double f = 1.0 / sqrt(lam[k0 + jj] * dv);
      for (long r = 0; r < k * cw; ++r) z[(size_t)jj * k * cw + r] = q[(size_t)(k0 + jj) * k * cw + r] * f;
    }
    REF_HIP(hipMemcpyAsync(Cm, z.data(), z.size() * 8, hipMemcpyHostToDevice, c->stream));
    REF_RC(la.rotate(X, k, Cm, k, k2));
  }
  // 3. H W; with Q = [U, W] (order nt = n + k2) the matrices A1 = Q^H H Q
  // (U^H H U = diag(theta): U are Ritz vectors) and A2 = (H Q)^H (H Q), raw
  // (unweighted) sums: U^H H^2 U and (H U)^H (H W) from H U and H^2 U formed
  // chunk by chunk, (H W)^H (H W) directly.
  REF_RC(apply_h_dev_impl(c, X, Y, k2));
  const long nt = n + k2;
  REF_HIP(hipMalloc(&M, (size_t)nt * nt * es));
  REF_HIP(hipMalloc(&M2, (size_t)nt * nt * es));
  REF_RC(fill_zero(c, M, (size_t)nt * nt * es));
  for (long i = 0; i < n; ++i) th[i] = eval[i] / dv;
  REF_HIP(hipMemcpyAsync(d_th, th.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(set_diag_kernel, dim3(64), dim3(256), 0, c->stream, M, nt, n, (const double *)d_th, cw - 1);
  REF_HIP(hipGetLastError());
  REF_RC(la.gram(d_psi, n, Y, k2, M + (size_t)n * nt * cw, nt));
  REF_RC(la.gram(X, k2, Y, k2, M + ((size_t)n * nt + n) * cw, nt));
  hipLaunchKernelGGL(mirror_upper_kernel, dim3(256), dim3(256), 0, c->stream, M, nt, cw - 1);
  REF_HIP(hipGetLastError());
  REF_RC(la.gram(Y, k2, Y, k2, M2 + ((size_t)n * nt + n) * cw, nt));
  {
    long cch;
    if (in_arena) {
      // past X and Y: at least 2 k_arena states, so cch >= k
      cch = std::max(1L, std::min(n, (arena_states(c, hw, st_b) - 2 * k) / 2));
      T1 = Y + (size_t)k * sl;
      T2 = T1 + (size_t)cch * sl;
    } else {
      cch = std::max(1L, std::min(n, states_that_fit(c, 2 * st_b, 0.7)));
      if (big_malloc(c, (void **)&T1, (size_t)cch * st_b) != hipSuccess ||
          big_malloc(c, (void **)&T2, (size_t)cch * st_b) != hipSuccess) {
        cleanup();
        return set_err(c, FBSE_ERR_NOMEM, "refine: two blocks of %ld states for H U", cch);
      }
    }
    for (long a0 = 0; a0 < n; a0 += cch) {
      const long cc = std::min(cch, n - a0);
      REF_RC(apply_h_dev_impl(c, d_psi + a0 * sl, T1, cc));
      REF_RC(apply_h_dev_impl(c, T1, T2, cc));
      REF_RC(la.gram(d_psi, n, T2, cc, M2 + (size_t)a0 * nt * cw, nt));
      REF_RC(la.gram(T1, cc, Y, k2, M2 + ((size_t)n * nt + a0) * cw, nt));
    }
    REF_HIP(hipStreamSynchronize(c->stream));
    if (!in_arena) {
      hipFree(T1);
      hipFree(T2);
    }
    T1 = T2 = nullptr;
  }
  hipLaunchKernelGGL(mirror_upper_kernel, dim3(256), dim3(256), 0, c->stream, M2, nt, cw - 1);
  REF_HIP(hipGetLastError());
  // 4. candidates for every selected pair, from two constructions over span(Q):
  //  (a) Rayleigh-Ritz: the eigenvectors y of A1 whose weight on the pair's
  //      Ritz vector exceeds one half (one to one), predicted residual^2
  //      y^H A2 y - mu^2;
  //  (b) refined vectors: the s lowest eigenvectors of
  //      G(theta) = A2 - 2 theta A1 + theta^2 for a cluster of s pairs with
  //      nearly equal theta (the least ||(H - theta) Q y|| over unit y; the
  //      pair's own Ritz vector is one such y, so the lowest value never
  //      exceeds its sigma^2), theta updated twice to the Rayleigh quotient of
  //      the cluster's refined vectors.
  //  Each pair takes the candidate with the lower predicted residual.
  const long nt2 = nt * nt;
  std::vector<double> A1h(nt2 * cw), A2h(nt2 * cw);
  REF_HIP(hipMemcpyAsync(A1h.data(), M, A1h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipMemcpyAsync(A2h.data(), M2, A2h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipStreamSynchronize(c->stream));
  for (auto &v : A1h) v *= dv;  // A1, A2 in the dv-weighted metric (energies, energies^2)
  for (auto &v : A2h) v *= dv;
  // y^H A y = sum_ij conj(y_i) A_ij y_j (real for Hermitian A), A column-major on the host
  auto quad = [&](const std::vector<double> &A, const double *y) {
    double acc = 0.0;
    for (long j = 0; j < nt; ++j) {
      const double yr = y[j * cw], yi = cw == 2 ? y[j * cw + 1] : 0.0;
      for (long i = 0; i < nt; ++i) {
        const double ar = A[(j * nt + i) * cw], ai = cw == 2 ? A[(j * nt + i) * cw + 1] : 0.0;
        const double xr = y[i * cw], xi = cw == 2 ? y[i * cw + 1] : 0.0;
        const double pr = ar * yr - ai * yi, pi = ar * yi + ai * yr;  // A_ij y_j
        acc += xr * pr + xi * pi;                                     // Re(conj(y_i) A_ij y_j)
      }
    }
    return acc;
  };
  // clusters: consecutive values closer than the larger of the two residuals
  // (at most 1e-3 Ha): the Ritz values of one degenerate level, split by the
  // noise of the mixed-precision filter, refine together, so that each gets
  // its own vector of the level instead of a copy of its neighbour's
  std::vector<long> order(sel.begin(), sel.end());
  std::sort(order.begin(), order.end(), [&](long a, long b) { return eval[a] < eval[b]; });
  std::vector<std::vector<long>> clusters;
  for (long i : order) {
    if (!clusters.empty()) {
      const long j = clusters.back().back();
      const double tol = std::max(1e-9 * std::max(1.0, fabs(eval[i])), std::min(std::max(sigma[i], sigma[j]), 1e-3));
      if (fabs(eval[i] - eval[j]) <= tol) {
        clusters.back().push_back(i);
        continue;
      }
    }
    clusters.push_back({i});
  }
  const long nb = (long)clusters.size();
  std::vector<double> cth(nb);
  for (long b = 0; b < nb; ++b) {
    double sum = 0.0;
    for (long i : clusters[b]) sum += eval[i];
    cth[b] = sum / (double)clusters[b].size();
    std::sort(clusters[b].begin(), clusters[b].end(), [&](long a, long bb) { return sigma[a] < sigma[bb]; });
  }
  REF_HIP(hipMalloc(&G, (size_t)nb * nt2 * es));
  REF_HIP(hipMalloc(&Gw, (size_t)nb * nt * 8));
  REF_HIP(hipMalloc(&Ge, (size_t)nb * nt * 8));
  REF_HIP(hipMalloc(&Ginfo, (size_t)nb * sizeof(int)));
  std::vector<double> lam_b((size_t)nb * nt), vec_b((size_t)nb * nt2 * cw);
  std::vector<int> info(nb);
  auto refined_solve = [&]() -> int {
    REF_HIP(hipMemcpyAsync(d_th, cth.data(), nb * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(refined_matrix_kernel, dim3(256), dim3(256), 0, c->stream, (const double *)M,
                       (const double *)M2, (const double *)d_th, nt, nb, dv, cw - 1, G);
    REF_HIP(hipGetLastError());
    rocblas_status st;
    if (cw == 2)
      st = rocsolver_zheevd_strided_batched(c->blas, rocblas_evect_original, rocblas_fill_upper, (int)nt,
                                            (rocblas_double_complex *)G, (int)nt, (rocblas_stride)nt2, Gw,
                                            (rocblas_stride)nt, Ge, (rocblas_stride)nt, Ginfo, (int)nb);
    else
      st = rocsolver_dsyevd_strided_batched(c->blas, rocblas_evect_original, rocblas_fill_upper, (int)nt, G, (int)nt,
                                            (rocblas_stride)nt2, Gw, (rocblas_stride)nt, Ge, (rocblas_stride)nt,
                                            Ginfo, (int)nb);
    if (st != rocblas_status_success) {
      rc = set_err(c, FBSE_ERR_SOLVER, "refine: batched eigensolver status %d", (int)st);
      cleanup();
      return rc;
    }
    REF_HIP(hipMemcpyAsync(lam_b.data(), Gw, lam_b.size() * 8, hipMemcpyDeviceToHost, c->stream));
    REF_HIP(hipMemcpyAsync(vec_b.data(), G, vec_b.size() * 8, hipMemcpyDeviceToHost, c->stream));
    REF_HIP(hipMemcpyAsync(info.data(), Ginfo, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    REF_HIP(hipStreamSynchronize(c->stream));
    return FBSE_OK;
  };
  for (int round = 0; round < 3; ++round) {
    if ((rc = refined_solve())) return rc;
    if (round == 2) break;
    for (long b = 0; b < nb; ++b) {  // theta <- mean Rayleigh quotient of the cluster's refined vectors
      if (info[b] != 0) continue;
      double sum = 0.0;
      const long sz = (long)clusters[b].size();
      for (long r = 0; r < sz; ++r) sum += quad(A1h, &vec_b[((size_t)b * nt + r) * nt * cw]);
      cth[b] = sum / (double)sz;
    }
  }
  // Rayleigh-Ritz eigenvectors of A1 (the raw M, in place) and their predicted residuals
  std::vector<double> mu(nt), Yrr(nt2 * cw);
  hipFree(d_w);
  d_w = nullptr;
  REF_HIP(hipMalloc(&d_w, (size_t)nt * 8));
  REF_RC(la.eigh(M, nt, d_w));
  REF_HIP(hipMemcpyAsync(mu.data(), d_w, nt * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipMemcpyAsync(Yrr.data(), M, Yrr.size() * 8, hipMemcpyDeviceToHost, c->stream));
  REF_HIP(hipStreamSynchronize(c->stream));
  std::vector<char> is_sel(n, 0);
  for (long j = 0; j < k; ++j) is_sel[sel[j]] = 1;
  auto w2 = [&](const double *y, long l) {
    return cw == 2 ? y[2 * l] * y[2 * l] + y[2 * l + 1] * y[2 * l + 1] : y[l] * y[l];
  };
  // cluster of a pair when the cluster holds several (one degenerate level
  // split by the noise), else -1
  std::vector<long> grp(n, -1);
  for (long b = 0; b < nb; ++b)
    if (clusters[b].size() > 1)
      for (long i : clusters[b]) grp[i] = b;
  struct Cand {
    long slot;
    double pred;
    const double *y;
    long group;
  };
  std::vector<Cand> best(n, Cand{-1, 1e300, nullptr, -1});
  // single pairs: the Rayleigh-Ritz or the refined candidate, whichever has
  // the lower predicted residual (when below the pair's sigma)
  for (long j = 0; j < nt; ++j) {
    const double *y = &Yrr[(size_t)j * nt * cw];
    long bi = -1;
    double bw = 0.0;
    for (long i = 0; i < n; ++i)
      if (w2(y, i) > bw) {
        bw = w2(y, i);
        bi = i;
      }
    if (bi < 0 || bw <= 0.5 || !is_sel[bi] || grp[bi] >= 0) continue;
    const double m = mu[j] * dv;
    const double pred = sqrt(std::max(0.0, quad(A2h, y) - m * m));
    if (pred < sigma[bi] && pred < best[bi].pred) best[bi] = Cand{bi, pred, y, -1};
  }
  // a cluster of s pairs: its s refined vectors, which span the level as a
  // whole (the r-th lowest for the pair of r-th lowest sigma, reassigned by
  // their measured residuals at the end)
  for (long b = 0; b < nb; ++b) {
    if (info[b] != 0) continue;
    const bool multi = clusters[b].size() > 1;
    for (size_t r = 0; r < clusters[b].size(); ++r) {
      const long slot = clusters[b][r];
      const double pred = sqrt(std::max(0.0, lam_b[(size_t)b * nt + r]));
      if (multi || (pred < sigma[slot] && pred < best[slot].pred))
        best[slot] = Cand{slot, pred, &vec_b[((size_t)b * nt + r) * nt * cw], multi ? b : -1};
    }
  }
  // units: a single pair's candidate, or all vectors of one cluster, taken in
  // order of their lowest predicted residual.  A unit is dropped when one of
  // its vectors carries more than half its weight on a Ritz vector outside
  // the unit, or on a vector taken before: it would duplicate that state.
  std::vector<std::vector<Cand>> units;
  {
    std::vector<long> unit_of(nb, -1);
    for (long i = 0; i < n; ++i) {
      const Cand &cd = best[i];
      if (cd.slot < 0) continue;
      if (cd.group < 0) {
        units.push_back({cd});
      } else {
        if (unit_of[cd.group] < 0) {
          unit_of[cd.group] = (long)units.size();
          units.push_back({});
        }
        units[unit_of[cd.group]].push_back(cd);
      }
    }
  }
  auto unit_key = [](const std::vector<Cand> &u) {
    double m = 1e300;
    for (const Cand &cd : u) m = std::min(m, cd.pred);
    return m;
  };
  std::stable_sort(units.begin(), units.end(),
                   [&](const std::vector<Cand> &a, const std::vector<Cand> &b) { return unit_key(a) < unit_key(b); });
  std::vector<Cand> acc;
  std::vector<char> in_unit(n, 0);
  for (const auto &u : units) {
    bool ok = true;
    for (const Cand &cd : u) in_unit[cd.slot] = 1;
    for (const Cand &cd : u) {
      for (long l = 0; l < n && ok; ++l)
        if (!in_unit[l] && w2(cd.y, l) > 0.5) ok = false;
      for (const Cand &a : acc) {
        if (!ok) break;
        double re = 0.0, im = 0.0;
        for (long l = 0; l < nt; ++l) {
          if (cw == 2) {
            re += a.y[2 * l] * cd.y[2 * l] + a.y[2 * l + 1] * cd.y[2 * l + 1];
            im += a.y[2 * l] * cd.y[2 * l + 1] - a.y[2 * l + 1] * cd.y[2 * l];
          } else {
            re += a.y[l] * cd.y[l];
          }
        }
        if (re * re + im * im > 0.5) ok = false;
      }
      if (!ok) break;
    }
    for (const Cand &cd : u) in_unit[cd.slot] = 0;
    if (ok) acc.insert(acc.end(), u.begin(), u.end());
  }
  const long nc = (long)acc.size();
  if (nc == 0) {
    cleanup();
    return FBSE_OK;
  }
  // the accepted vectors Q y = U y_top + W y_bottom into Y (H W is not needed any more)
  {
    std::vector<double> yt((size_t)n * nc * cw), yb((size_t)k2 * nc * cw);
    for (long q2 = 0; q2 < nc; ++q2) {
      memcpy(&yt[(size_t)q2 * n * cw], acc[q2].y, (size_t)n * cw * 8);
      memcpy(&yb[(size_t)q2 * k2 * cw], acc[q2].y + (size_t)n * cw, (size_t)k2 * cw * 8);
    }
    double *d_yt = Cm;  // n x k >= n x nc
    REF_HIP(hipMemcpyAsync(d_yt, yt.data(), yt.size() * 8, hipMemcpyHostToDevice, c->stream));
    REF_RC(la.product(d_psi, n, d_yt, n, nc, Y, false));
    REF_HIP(hipStreamSynchronize(c->stream));
    REF_HIP(hipMemcpyAsync(d_yt, yb.data(), yb.size() * 8, hipMemcpyHostToDevice, c->stream));
    REF_RC(la.product(X, k2, d_yt, k2, nc, Y, true));
    REF_HIP(hipStreamSynchronize(c->stream));
  }
  REF_RC(normalize_dev_impl(c, Y, nc));
  std::vector<double> ev2(nc), sg2(nc);
  // (its H U chunk past X and Y when they are in the work buffer)
  REF_RC(energies_dev_impl(c, Y, nc, ev2.data(), sg2.data(), in_arena ? align256(hw) + 2 * k * st_b : hw));
  // a single pair is replaced when its new residual is lower; a cluster's
  // vectors, in ascending residual, replace its pairs in ascending sigma when
  // every one of them is lower (a level is replaced as a whole, so no pair of
  // it keeps a vector that a new one duplicates)
  long nrep = 0;
  auto put = [&](long q2, long i) -> int {
    REF_HIP(hipMemcpyAsync(d_psi + i * sl, Y + q2 * sl, sl * 8, hipMemcpyDeviceToDevice, c->stream));
    eval_out[i] = ev2[q2];
    sigma_out[i] = sg2[q2];
    ++nrep;
    return FBSE_OK;
  };
  for (long q0 = 0; q0 < nc;) {
    long q1 = q0 + 1;
    while (acc[q0].group >= 0 && q1 < nc && acc[q1].group == acc[q0].group) ++q1;
    if (acc[q0].group < 0) {
      const long i = acc[q0].slot;
      if (std::isfinite(sg2[q0]) && sg2[q0] < sigma[i] && (rc = put(q0, i))) return rc;
    } else {
      std::vector<long> qs, slots;
      for (long q2 = q0; q2 < q1; ++q2) {
        qs.push_back(q2);
        slots.push_back(acc[q2].slot);
      }
      std::sort(qs.begin(), qs.end(), [&](long a, long b) { return sg2[a] < sg2[b]; });
      std::sort(slots.begin(), slots.end(), [&](long a, long b) { return sigma[a] < sigma[b]; });
      bool all = true;
      for (size_t r = 0; r < qs.size() && all; ++r) all = std::isfinite(sg2[qs[r]]) && sg2[qs[r]] < sigma[slots[r]];
      if (all)
        for (size_t r = 0; r < qs.size(); ++r)
          if ((rc = put(qs[r], slots[r]))) return rc;
    }
    q0 = q1;
  }
#undef REF_HIP
#undef REF_RC
  *n_replaced = nrep;
  cleanup();
  return FBSE_OK;
}

extern "C" int fbse_refine_dev(fbse_ctx *c, double *d_psi, long n_ritz, const double *eval, const double *sigma,
                               double sigma_max, double e_lo, double e_hi, double e_p, double *eval_out,
                               double *sigma_out, long *n_replaced) {
  if (!c || !d_psi || !eval || !sigma || !eval_out || !sigma_out || !n_replaced || n_ritz < 1) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  if (e_p <= 0.0) e_p = 1.0;
  return refine_impl(c, d_psi, n_ritz, eval, sigma, sigma_max, e_lo, e_hi, e_p, eval_out, sigma_out, n_replaced);
}
extern "C" int fbse_refine(fbse_ctx *c, double *psitot, long n_ritz, const double *eval, const double *sigma,
                           double sigma_max, double e_lo, double e_hi, double e_p, double *eval_out,
                           double *sigma_out, long *n_replaced) {
  if (!c || !psitot || !eval || !sigma || !eval_out || !sigma_out || !n_replaced || n_ritz < 1) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  return with_device_copy(c, psitot, n_ritz, [&](double *d) -> int {
    int rc = fbse_refine_dev(c, d, n_ritz, eval, sigma, sigma_max, e_lo, e_hi, e_p, eval_out, sigma_out, n_replaced);
    if (rc) return rc;
    FBSE_HIP(c, hipMemcpyAsync(psitot, d, (size_t)n_ritz * sl * 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    return FBSE_OK;
  });
}

// ---------------------------------------------------------------------------
// Spectrum bounds (energy.c:222-300 and 343-369)
// ---------------------------------------------------------------------------
extern "C" int fbse_energy_range(fbse_ctx *c, const double *psi0_min, const double *psi0_max, long max_iter,
                                 double nl_lambda_min, double vloc_min, double *out) {
  if (!c || !psi0_min || !psi0_max || !out || max_iter < 0) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_of(c);
  double *psi = nullptr, *phi = nullptr, *d_s = nullptr;
  FBSE_HIP(c, hipMalloc(&psi, sl * 8));
  FBSE_HIP(c, hipMalloc(&phi, sl * 8));
  FBSE_HIP(c, hipMalloc(&d_s, 2 * 8));
  int rc = FBSE_OK;
  double s2[2];
  auto dot = [&](const double *a, const double *b, double *res) -> int {
    int r = b ? state_dots(c, a, b, 1, sl, d_s) : state_norms(c, a, 1, sl, d_s);
    if (r) return r;
    FBSE_HIP(c, hipMemcpyAsync(s2, d_s, 8, hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipStreamSynchronize(c->stream));
    *res = s2[0];
    return FBSE_OK;
  };
  auto normalize = [&](double *a) -> int {
    double nn;
    int r = dot(a, nullptr, &nn);
    if (r) return r;
    const double N = sqrt(nn * c->dv);
    return axpby_states(c, a, a, sl, 0.0, 1.0 / N);
  };
  double tau = 0.05, ene_old, Emin, Emax;
  long i = 0, imin = 0;
  FBSE_HIP(c, hipMemcpyAsync(phi, psi0_min, sl * 8, hipMemcpyHostToDevice, c->stream));
  Emin = (ene_old = 0.0) + 10.0;
  for (i = 0; !rc && (fabs((Emin - ene_old) / Emin) > 1.0e-6) && (i < max_iter); i++) {
    // psi <- phi ; phi <- H psi ; phi <- psi - tau phi ; normalise ; energy
    rc = apply_h_dev_impl(c, phi, psi, 1);  // psi = H phi_old
    if (!rc) rc = axpby_states(c, phi, psi, sl, 1.0, -tau);  // psi = phi_old - tau H phi_old
    if (!rc) {
      double *t = phi;
      phi = psi;
      psi = t;
      rc = normalize(phi);
    }
    if (!rc) rc = apply_h_dev_impl(c, phi, psi, 1);
    double e = 0.0;
    if (!rc) rc = dot(phi, psi, &e);
    ene_old = Emin;
    Emin = e * c->dv;
    if ((i > 5) && (Emin - ene_old) > 0) tau -= 0.025;
  }
  imin = i;
  if (!rc) {
    FBSE_HIP(c, hipMemcpyAsync(psi, psi0_max, sl * 8, hipMemcpyHostToDevice, c->stream));
    Emax = (ene_old = 0.0) + 0.1;
    for (i = 0; !rc && (fabs((Emax - ene_old) / Emax) > 1.0e-6) && (i < max_iter); i++) {
      rc = apply_h_dev_impl(c, psi, phi, 1);  // phi = H psi
      if (!rc) {
        double *t = psi;
        psi = phi;
        phi = t;
        rc = normalize(psi);
      }
      if (!rc) rc = apply_h_dev_impl(c, psi, phi, 1);
      double e = 0.0;
      if (!rc) rc = dot(psi, phi, &e);
      ene_old = Emax;
      Emax = e * c->dv;
    }
  }
  if (!rc) {
    out[0] = Emin;
    out[1] = Emax;
    if (c->has_proj && c->pj.nl && !c->pj.so && nl_lambda_min < 0.0) {
      Emax *= 1.05;
      Emin -= 0.1 * fabs(Emin);
      const double weyl_min = 1.05 * (vloc_min + nl_lambda_min);
      if (Emin > weyl_min) Emin = weyl_min;
    } else {
      Emax *= 1.2;
      Emin -= 0.2 * fabs(Emin);
    }
    out[2] = Emin;
    out[3] = Emax - Emin;
    out[4] = (double)imin;
    out[5] = (double)i;
  }
  hipStreamSynchronize(c->stream);
  hipFree(psi);
  hipFree(phi);
  hipFree(d_s);
  return rc;
}
