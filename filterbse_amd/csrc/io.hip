// io.hip -- state files streamed between HBM and disk (include/fbse_io.h).
//
// Writes: the selected states are cut into chunks of at most io_chunk bytes;
// chunk k is copied device-to-host on the context's stream into pinned
// buffer k % 2 (so it follows every kernel that produced the states) while
// the host writes chunk k - 1 from the other buffer.  Reads: the host reads
// chunk k into one pinned buffer while the host-to-device copy of chunk k - 1
// runs from the other.  Two buffers, one event each: a buffer is reused only
// after its previous copy has completed.
#include "fbse_internal.h"
#include "../../include/fbse_io.h"

#include <algorithm>
#include <vector>

namespace fbse {

void io_release(fbse_ctx *c) {
  for (int b = 0; b < 2; ++b) {
    if (c->h_io[b]) (void)hipHostFree(c->h_io[b]);
    if (c->io_ev[b]) (void)hipEventDestroy(c->io_ev[b]);
    c->h_io[b] = nullptr;
    c->io_ev[b] = nullptr;
  }
  c->io_bytes = 0;
}

static int io_prepare(fbse_ctx *c) {
  if (c->io_bytes == c->io_chunk && c->h_io[0] && c->h_io[1]) return FBSE_OK;
  io_release(c);
  for (int b = 0; b < 2; ++b) {
    if (hipHostMalloc(&c->h_io[b], c->io_chunk, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      io_release(c);
      return set_err(c, FBSE_ERR_NOMEM, "pinned staging buffer of %zu bytes", c->io_chunk);
    }
    FBSE_HIP(c, hipEventCreateWithFlags(&c->io_ev[b], hipEventDisableTiming));
  }
  c->io_bytes = c->io_chunk;
  return FBSE_OK;
}

namespace {
// one chunk of the transfer: bytes [off, off + len) of state `state`
struct Chunk {
  long state;
  size_t off, len;
};

std::vector<Chunk> chunks_of(const long *idx, long n, size_t state_bytes, size_t cap) {
  std::vector<Chunk> out;
  for (long k = 0; k < n; ++k) {
    const long s = idx ? idx[k] : k;
    for (size_t off = 0; off < state_bytes; off += cap) out.push_back({s, off, std::min(cap, state_bytes - off)});
  }
  return out;
}

int write_zeros(FILE *pf, size_t bytes) {
  static const double zero[4096] = {};
  while (bytes > 0) {
    const size_t b = std::min(bytes, sizeof zero);
    if (fwrite(zero, 1, b, pf) != b) return -1;
    bytes -= b;
  }
  return 0;
}
}  // namespace

static long stlen_io(const fbse_ctx *c) { return c->nsg * (c->is_complex + 1); }

}  // namespace fbse

using namespace fbse;

extern "C" int fbse_set_io_chunk(fbse_ctx *c, size_t bytes) {
  if (!c || bytes < 4096 || bytes % 8) return FBSE_ERR_ARG;
  c->io_chunk = bytes;
  return FBSE_OK;
}

extern "C" int fbse_fwrite_states(fbse_ctx *c, const double *d_psi, const long *idx, long n, int n_every, FILE *pf) {
  if (!c || !pf || n < 0 || n_every < 1 || (n > 0 && !d_psi)) return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const size_t sb = (size_t)stlen_io(c) * sizeof(double);
  int rc;
  if ((rc = io_prepare(c))) return rc;
  const std::vector<Chunk> ch = chunks_of(idx, n, sb, c->io_bytes);
  const size_t nc = ch.size();
  auto issue = [&](size_t k) -> int {
    const Chunk &q = ch[k];
    FBSE_HIP(c, hipMemcpyAsync(c->h_io[k % 2], (const char *)d_psi + (size_t)q.state * sb + q.off, q.len,
                               hipMemcpyDeviceToHost, c->stream));
    FBSE_HIP(c, hipEventRecord(c->io_ev[k % 2], c->stream));
    return FBSE_OK;
  };
  for (size_t k = 0; k < std::min<size_t>(2, nc); ++k)
    if ((rc = issue(k))) {
      (void)hipStreamSynchronize(c->stream);
      return rc;
    }
  for (size_t k = 0; k < nc; ++k) {
    FBSE_HIP(c, hipEventSynchronize(c->io_ev[k % 2]));
    if ((ch[k].off == 0 && n_every > 1 && write_zeros(pf, (size_t)(n_every - 1) * sb)) ||
        fwrite(c->h_io[k % 2], 1, ch[k].len, pf) != ch[k].len) {
      (void)hipStreamSynchronize(c->stream);  // no copy may still target the staging buffers
      return set_err(c, FBSE_ERR_IO, "state file: write of state %ld failed", ch[k].state);
    }
    if (k + 2 < nc && (rc = issue(k + 2))) {
      (void)hipStreamSynchronize(c->stream);
      return rc;
    }
  }
  return FBSE_OK;
}

extern "C" int fbse_write_states(fbse_ctx *c, const double *d_psi, const long *idx, long n, int n_every,
                                 const char *path) {
  if (!c || !path) return FBSE_ERR_ARG;
  FILE *pf = fopen(path, "wb");
  if (!pf) return set_err(c, FBSE_ERR_IO, "cannot open %s for writing", path);
  int rc = fbse_fwrite_states(c, d_psi, idx, n, n_every, pf);
  if (fclose(pf) != 0 && rc == FBSE_OK) rc = set_err(c, FBSE_ERR_IO, "closing %s failed", path);
  return rc;
}

extern "C" int fbse_read_start_states(fbse_ctx *c, const char *path, long start, long end, int n_every,
                                      double *d_psi_rank, long n_starts, long ms, long *n_read) {
  if (n_read) *n_read = 0;
  if (!c || !path || start < 0 || end < start || n_every < 1 || n_starts < 0 || ms < 1 ||
      (n_starts > 0 && !d_psi_rank))
    return FBSE_ERR_ARG;
  FBSE_HIP(c, hipSetDevice(c->device));
  const long sl = stlen_io(c);
  const size_t sb = (size_t)sl * sizeof(double);
  // the states read_psi_dat selects: start + n_every - 1, then every n_every-th, below end
  std::vector<long> sel;
  for (long i = start; i < end && (long)sel.size() < n_starts; i += n_every) sel.push_back(i + n_every - 1);
  if (sel.empty()) return FBSE_OK;
  int rc;
  if ((rc = io_prepare(c))) return rc;
  FILE *pf = fopen(path, "rb");
  if (!pf) return set_err(c, FBSE_ERR_IO, "cannot open %s", path);
  std::vector<Chunk> ch;
  for (size_t k = 0; k < sel.size(); ++k)
    for (size_t off = 0; off < sb; off += c->io_bytes) ch.push_back({(long)k, off, std::min(c->io_bytes, sb - off)});
  bool used[2] = {false, false};
  for (size_t k = 0; k < ch.size() && !rc; ++k) {
    const int b = (int)(k % 2);
    if (used[b] && hipEventSynchronize(c->io_ev[b]) != hipSuccess) {
      rc = set_err(c, FBSE_ERR_HIP, "state file: copy to the device failed");
      break;
    }
    const long fstate = sel[ch[k].state];
    if (fseek(pf, (long)((size_t)fstate * sb + ch[k].off), SEEK_SET) != 0 ||
        fread(c->h_io[b], 1, ch[k].len, pf) != ch[k].len) {
      rc = set_err(c, FBSE_ERR_IO, "%s: state %ld is beyond the end of the file", path, fstate);
      break;
    }
    double *dst = d_psi_rank + (size_t)ch[k].state * ms * sl;
    if (hipMemcpyAsync((char *)dst + ch[k].off, c->h_io[b], ch[k].len, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipEventRecord(c->io_ev[b], c->stream) != hipSuccess) {
      rc = set_err(c, FBSE_ERR_HIP, "state file: copy to the device failed");
      break;
    }
    used[b] = true;
  }
  fclose(pf);
  // the staging buffers are reused by the next call: wait for the copies
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = set_err(c, FBSE_ERR_HIP, "state file copy");
  if (!rc && n_read) *n_read = (long)sel.size();
  return rc;
}
