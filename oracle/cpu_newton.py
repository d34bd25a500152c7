"""Threaded CPU Newton filter step: the CPU baseline of bench.py.

TEST INFRASTRUCTURE ONLY (the benchmark's cpu_baseline leg and tests/):
nothing in the filterbse_amd package imports it.  It runs the reference
filter's inner loop for one start state, filter_cycle's step jc
(filter.c:318-359): p_hnorm (filter.c:375-410) = p_hamiltonian
(phamiltonian.c:10-60: kinetic term by a forward FFT, x ksqr, backward FFT
per spin, hamiltonian.c:62-99; SO + NL projectors and V_loc,
phamiltonian.c:65-444), the Newton shift, and the accumulation of the step
into every energy target.  The FFTs are scipy's (pocketfft, O(N log N), one
worker thread per core); projectors, V_loc, shift and accumulation are the
oracle's C restatement threaded with OpenMP over atoms / grid points, the
projector expansion added with atomics as the reference does
(phamiltonian.c:288-303).  tests/test_cpu_newton.py pins it against the
serial oracle (or_hamiltonian, or_filter_cycle).
"""
import ctypes as C
import os
import time

import numpy as np

from . import oracle as O


def host_cores():
    """CPU threads this process may use (its affinity set, at most 16: the
    CPU share of one GPU on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class CpuNewton:
    """The Hamiltonian of a workload (the arrays filterbse_amd.dots.build_dot
    returns) and the threaded Newton step for one start state."""

    def __init__(self, nx, ny, nz, nspin, is_complex, dv, ksqr, pot_local, proj=None, nthreads=None):
        import scipy.fft  # noqa: F401  (fail early when scipy is absent)
        self.ham = O.Ham(nx, ny, nz, nspin, is_complex, dv, ksqr, pot_local, proj)
        self.nx, self.ny, self.nz, self.nspin = nx, ny, nz, nspin
        self.ngrid = nx * ny * nz
        self.nsg = nspin * self.ngrid
        self.ksqr = np.ascontiguousarray(ksqr, dtype=np.float64).reshape(nz, ny, nx)
        self.nthreads = nthreads or host_cores()
        L = O.lib()
        L.or_potential_omp.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(O.OrHam), C.c_int]
        L.or_hnorm_acc_omp.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                       C.c_long, C.c_long, O.DP, C.c_long, C.c_long, C.c_int]

    def hamiltonian(self, psi, out):
        """out <- H psi for one complex state (nsg complex128 values)."""
        import scipy.fft as F
        shape = (self.nz, self.ny, self.nx)
        for s in range(self.nspin):
            g = psi[s * self.ngrid:(s + 1) * self.ngrid].reshape(shape)
            w = F.fftn(g, workers=self.nthreads)                  # FFTW_FORWARD, unnormalised
            w *= self.ksqr                                         # ksqr carries the 1/N
            out[s * self.ngrid:(s + 1) * self.ngrid] = F.ifftn(w, norm="forward",
                                                               workers=self.nthreads).ravel()
        O.lib().or_potential_omp(out.ctypes.data, psi.ctypes.data, C.byref(self.ham.h), self.nthreads)
        return out

    def newton_steps(self, block, ms, an, zn, Vmin, dE_1, j_first, j_last, psi):
        """Newton steps jc = j_first .. j_last of filter_cycle for one start
        state: psi (complex, in/out) is psi_{jc-1} on entry, block (ms *
        2*nsg float64, the start's psi_rank block) accumulates
        an[jms, jc] psi_jc."""
        hpsi = np.empty_like(psi)
        an = np.ascontiguousarray(an, dtype=np.float64)
        M = len(zn)
        for jc in range(j_first, j_last + 1):
            self.hamiltonian(psi, hpsi)
            O.lib().or_hnorm_acc_omp(psi.ctypes.data, hpsi.ctypes.data, float(zn[jc - 1]), Vmin, dE_1,
                                     block.ctypes.data, ms, self.nsg, O._dp(an), M, jc, self.nthreads)
        return psi


def baseline(job, ms, an, zn, Vmin, dE_1, start, budget_s=20.0, max_steps=64):
    """Times Newton steps of one start state (the first ran2 start) on all
    host cores for about budget_s seconds; returns a dict for the benchmark
    line (rate in H psi per second: state x Newton step pairs)."""
    cn = CpuNewton(job["nx"], job["ny"], job["nz"], job["nspin"], job["is_complex"], job["dv"], job["ksqr"],
                   job["pot_local"], job["proj"])
    nsg = cn.nsg
    psi = np.ascontiguousarray(start, dtype=np.float64).view(np.complex128).copy()
    assert psi.size == nsg
    block = np.zeros(ms * 2 * nsg)
    # term 0 as filter_cycle (filter.c:284-312) is setup, not timed
    cn.newton_steps(block, ms, an, zn, Vmin, dE_1, 1, 1, psi)  # one untimed step (thread pools, pages)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps:
        cn.newton_steps(block, ms, an, zn, Vmin, dE_1, 2 + n, 2 + n, psi)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return dict(value=n / el, unit="Hψ/s", cores=cn.nthreads, kind="port",
                sample="%d Newton steps of one start state (H: scipy.fft FFT pair per spin with %d workers, "
                       "SO+NL projectors and V_loc in the oracle's C on %d OpenMP threads; shift and "
                       "accumulation into %d targets), %.1f s" % (n, cn.nthreads, cn.nthreads, ms, el),
                host_cpus=os.cpu_count())
