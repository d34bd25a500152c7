"""CPU oracle of the reference's state-file layout (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
filterbse_amd/csrc/io.hip; nothing in the product path imports it.

write_psi_dat    the bytes the reference writes for a block of states: each
                 state's stlen doubles, raw and back to back (write_psi_dat
                 write.c:289-299, the converged-state loop of mod_output.c:81-100,
                 psi-filt-<jns>-<rank>.dat filter.c:118-138, the psitot block of
                 save_output save.c:492), with n_every - 1 zero states before
                 each one when a stride is asked for.
read_psi_dat     the states read_psi_dat (aux.c:131-156) reads from such
                 bytes: seek to start, then per i = start, start + n_every, ...
                 < end skip n_every - 1 states and read one.
start_blocks     where the filter reads start state jns (filter.c:253,
                 offset jns*ms*stlen), i.e. where the fixed reader places the
                 c-th state it reads.
"""
import numpy as np


def write_psi_dat(psitot, stlen, idx=None, n_every=1):
    psitot = np.asarray(psitot, dtype=np.float64).reshape(-1, stlen)
    order = range(psitot.shape[0]) if idx is None else idx
    parts = []
    for k in order:
        if n_every > 1:
            parts.append(np.zeros((n_every - 1) * stlen))
        parts.append(psitot[k])
    return np.concatenate(parts).astype("<f8").tobytes() if parts else b""


def read_psi_dat(raw, start, end, stlen, n_every):
    """States (rows) read as aux.c:131-156 reads them; a state past the end
    of the data ends the list (the reference's fread would come back short)."""
    data = np.frombuffer(raw, dtype="<f8")
    st = stlen
    pos = start * st                   # fseek(pf, start * st_mem, SEEK_SET)
    out = []
    for _ in range(start, end, n_every):
        pos += (n_every - 1) * st      # fseek(pf, offset, SEEK_CUR)
        if pos + st > data.size:
            break
        out.append(data[pos:pos + st].copy())   # fread(&psi[cntr * len], ...)
        pos += st
    return np.array(out).reshape(-1, stlen)


def start_blocks(states, n_starts, ms, stlen):
    """psi_rank (n_starts*ms*stlen doubles, zero elsewhere) with state c at
    offset c*ms*stlen (filter.c:253)."""
    out = np.zeros(n_starts * ms * stlen)
    for c, s in enumerate(np.asarray(states).reshape(-1, stlen)[:n_starts]):
        out[c * ms * stlen:(c * ms + 1) * stlen] = s
    return out
