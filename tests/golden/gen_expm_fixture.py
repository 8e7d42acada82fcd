"""Write tests/golden/expm_dimer.npz: scipy.linalg.expm of the dimer's
-i H_eff,b dt for small N (this container only; scipy 1.15.3).

The matrices are built here from the second-quantised definitions
H = -J (b1^+ b2 + b2^+ b1) + 2 (U/N) sum_j n_j (n_j - 1) + E_b (n1 - n2),
L = -(b1^+ + b2^+)(b1 - b2) in the basis |N - k, k> (k = bosons on site 2),
H_eff = H - (i/2)(gamma/N) L^+ L, independently of the oracle's code.
"""
import os

import numpy as np
import scipy
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "expm_dimer.npz")


def two_mode_ops(N):
    """b1^+ b2, n1, n2 on the N-particle sector, basis index k = n2."""
    d = N + 1
    b1d_b2 = np.zeros((d, d))
    for k in range(1, d):
        # b1^+ b2 |N-k, k> = sqrt((N-k+1) k) |N-k+1, k-1>
        b1d_b2[k - 1, k] = np.sqrt((N - k + 1) * k)
    n2 = np.diag(np.arange(d, dtype=float))
    n1 = np.diag(N - np.arange(d, dtype=float))
    return b1d_b2, n1, n2


def dimer(N, J, U, E, gamma):
    b1d_b2, n1, n2 = two_mode_ops(N)
    I = np.eye(N + 1)
    H = -J * (b1d_b2 + b1d_b2.T) + 2.0 * (U / N) * (n1 @ (n1 - I) + n2 @ (n2 - I)) + E * (n1 - n2)
    # (b1^+ + b2^+)(b1 - b2) = n1 - n2 + b2^+ b1 - b1^+ b2
    Lp = (n1 - n2) + b1d_b2.T - b1d_b2
    L = -Lp
    Heff = H - 0.5j * (gamma / N) * (L.conj().T @ L)
    return H, L, Heff


def main():
    rows = {}
    T = 2 * np.pi
    cases = []
    for N in (1, 2, 4, 10):
        for b, E in enumerate((1.0 + 1.5, 1.0 - 1.5)):
            for dt in (T / 4, T / 32, T / 2 ** 17, T / 200):
                cases.append((N, b, dt))
                _, _, Heff = dimer(N, 1.0, 0.5, E, 0.1)
                rows["U_N%d_b%d_dt%.12e" % (N, b, dt)] = sl.expm(-1j * Heff * dt)
    np.savez_compressed(OUT, scipy_version=scipy.__version__, **rows)
    print(OUT, len(rows))


if __name__ == "__main__":
    main()
