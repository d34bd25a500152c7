/* fbse_io.h -- state files of the filter-diagonalization path, streamed
 * between HBM and disk.
 *
 * The reference writes its state blocks with one fwrite from host memory
 * (write_psi_dat, write.c:289-299; the converged-state psi.dat of mod_output,
 * mod_output.c:81-100; psi-filt-<jns>-<rank>.dat, filter.c:118-138; the psitot
 * block of output.dat, save.c:446-503) and reads start states back with
 * read_psi_dat (aux.c:131-156) under inputPsiFilt (init.c:1168-1205).  With the
 * states in HBM (29 GB at 64 x 16 states of a 96^3 spinor grid, 137 GB at
 * 256 states of 256^3) these entry points stream them through two pinned host
 * buffers: the device-to-host copy of one chunk overlaps the write of the
 * previous one, and nothing of the block is staged whole in host memory.
 *
 * File layout is the reference's: each state is stlen = complex_idx * nspin *
 * ngrid doubles, raw native-endian binary, states back to back.
 */
#ifndef FBSE_IO_H
#define FBSE_IO_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

struct fbse_ctx;

/* Writes n states of the device block d_psi (state k at d_psi + k*stlen) to
 * the open stream pf, in the order of idx (host array of state indices; NULL
 * writes states 0 .. n-1), each as stlen raw doubles.  n_every > 1 writes
 * n_every - 1 zero states before each state: the layout that read_psi_dat
 * (aux.c:131-156) reads back with stride n_every (the reference's spinor runs
 * read with n_every = 2, init.c:1170-1171).  The stream's own buffering is
 * kept (pf may hold a header written with fprintf, as output.dat does).
 * Replaces the fwrite of write.c:296, mod_output.c:98, filter.c:130-134 and
 * save.c:492.  Returns FBSE_OK or an error code (FBSE_ERR_IO when the stream
 * refuses the bytes). */
int fbse_fwrite_states(struct fbse_ctx *ctx, const double *d_psi, const long *idx, long n, int n_every, FILE *pf);

/* fbse_fwrite_states to a new file (truncated) at path, closed afterwards. */
int fbse_write_states(struct fbse_ctx *ctx, const double *d_psi, const long *idx, long n, int n_every,
                      const char *path);

/* Start states from a file (the inputPsiFilt path, init.c:1168-1205): the
 * states read_psi_dat(pf, psi, start, end, stlen, n_every) selects (aux.c:131-156:
 * states start + n_every - 1, start + 2*n_every - 1, ... below end) are
 * placed where the filter reads start states (filter.c:253): selected state c
 * goes to d_psi_rank + c*ms*stlen, for c < n_starts (the reference stores them
 * contiguously, c*stlen, so that the filter uses only every ms-th one).  No
 * cube files are written (init.c:1177-1200 writes one per state).  *n_read
 * (may be NULL) receives the number of states placed.  Returns FBSE_OK, or
 * FBSE_ERR_IO when the file cannot be opened or is shorter than the
 * selection. */
int fbse_read_start_states(struct fbse_ctx *ctx, const char *path, long start, long end, int n_every,
                           double *d_psi_rank, long n_starts, long ms, long *n_read);

/* Bytes of each pinned staging buffer (two per context; default 64 MiB,
 * allocated on first use, released by fbse_trim). */
int fbse_set_io_chunk(struct fbse_ctx *ctx, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif
