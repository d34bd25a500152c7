"""The threaded CPU Newton step of the benchmark's CPU baseline
(oracle/cpu_newton.py) against the serial oracle: H (scipy FFT pair + the
OpenMP projector / V_loc restatement) against or_hamiltonian, and a short
Newton filter cycle against or_filter_cycle, on spinor systems with SO and NL
projectors (5 and 1 radial projectors)."""
import numpy as np
import pytest

from oracle import oracle as O
from systems import VARIANTS, oracle_ham, random_states, small_system


@pytest.mark.parametrize("variant", ["spinor_so_nl", "spinor_so_nl_1p", "spinor_vloc"])
def test_cpu_hamiltonian_matches_oracle(variant):
    from oracle.cpu_newton import CpuNewton
    s = small_system(**VARIANTS[variant])
    cn = CpuNewton(s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"], s["ksqr"], s["pot_local"],
                   s["proj"], nthreads=4)
    psi = random_states(s, 1, seed=4)
    ref = oracle_ham(s).apply(psi)
    z = psi.view(np.complex128).copy()
    out = np.empty_like(z)
    cn.hamiltonian(z, out)
    err = np.abs(out.view(np.float64) - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


def test_cpu_newton_cycle_matches_oracle():
    from oracle.cpu_newton import CpuNewton
    s = small_system(**VARIANTS["spinor_so_nl"])
    h = oracle_ham(s)
    cn = CpuNewton(s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"], s["ksqr"], s["pot_local"],
                   s["proj"], nthreads=3)
    M, ms = 7, 3
    psi_rank, _ = O.init_filter_states(1, ms, -1, s["nx"], s["ny"], s["nz"], s["nspin"], s["is_complex"], s["dv"])
    Vmin = s["pot_local"].min() - 3.0
    dE = s["pot_local"].max() + 25.0 - Vmin
    zn = O.leja_points(M)
    an = O.newton_coeff(zn, np.linspace(Vmin + 2.5, Vmin + 4.0, ms), Vmin, dE, (M / (2.5 * dE)) ** 2)
    ref = h.filter_cycles(psi_rank, 1, ms, an, zn, Vmin, 4.0 / dE)
    stlen = h.stlen
    psi0 = psi_rank[:stlen].view(np.complex128).copy()
    block = np.zeros(ms * stlen)
    for t in range(ms):  # term 0 (filter.c:284-312)
        a = an[2 * M * t] + 1j * an[2 * M * t + 1]
        block[t * stlen:(t + 1) * stlen] = (a * psi0).view(np.float64)
    cn.newton_steps(block, ms, an, zn, Vmin, 4.0 / dE, 1, M - 1, psi0)
    err = np.abs(block - ref).max() / np.abs(ref).max()
    assert err < 1e-11, err
