"""Closed forms of the per-trajectory observables (calc_chars_std, core.cpp:665-725;
get_mean_simple / get_m2 / get_dispersion_simple / get_energy, experiment.cpp:1395-1469)
on spin-coherent states of the dimer, used to pin the oracle and the GPU.

A coherent state puts every boson in the mode sqrt(1-p) b1^dag + sqrt(p) e^{i phi} b2^dag.
Its amplitudes are psi_k = c sqrt(C(N,k) p^k (1-p)^(N-k)) e^{i k phi} (k = bosons in site
2), so |psi_k|^2 / ||psi||^2 is the binomial pmf:
  norm = c^2,  <k> = N p,  <(k - <k>)^2> = N p (1 - p),
and with H_kk = 2 (U/N) [k(k-1) + (N-k)(N-k-1)], H_{k,k+1} = -J sqrt((k+1)(N-k)),
H_drv = diag(N - 2k) (newdel.cpp:450-498) the energy <H + E H_drv> is
  -2 J N sqrt(p (1-p)) cos(phi) + 2 U (N-1) (p^2 + (1-p)^2) + E N (1 - 2p),
from <b1^dag b2> = N sqrt(p(1-p)) e^{i phi} and the binomial factorial moments
E[n(n-1)] = N (N-1) q^2.
"""
import numpy as np
from scipy.special import comb


def coherent_state(N, p, phi, c=1.0):
    k = np.arange(N + 1)
    amp = np.sqrt(comb(N, k) * p ** k * (1.0 - p) ** (N - k))
    return c * amp * np.exp(1j * k * phi)


def coherent_chars(N, p, phi, c, prm_E, prm_U, prm_J):
    k = np.arange(N + 1)
    pmf = comb(N, k) * p ** k * (1.0 - p) ** (N - k)
    energy = (-2.0 * prm_J * N * np.sqrt(p * (1.0 - p)) * np.cos(phi)
              + 2.0 * prm_U * (N - 1) * (p * p + (1.0 - p) ** 2) + prm_E * N * (1.0 - 2.0 * p))
    return {"norm": c * c, "mean": N * p, "m2": N * p * (1.0 - p), "energy": energy, "adr": pmf}


def coherent_ensemble(N, M, seed):
    """M coherent states with random p, phi and norm c^2 in (0.2, 1]."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.02, 0.98, M)
    phi = rng.uniform(-np.pi, np.pi, M)
    c = np.sqrt(rng.uniform(0.2, 1.0, M))
    psi = np.stack([coherent_state(N, p[i], phi[i], c[i]) for i in range(M)])
    return p, phi, c, psi
