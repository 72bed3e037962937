"""compute_xi: anisotropic 2PCF multipoles from the NN and RR multipoles with
the survey-boundary correction of Slepian & Eisenstein (2015) eq. 10 (the
reference's fftcorr/correlate/utils.py:5-69).  Host-side post-processing of
the correlator outputs; numpy, FP64, vectorised over separation bins.

With f_j = RR_j / RR_0 and x_l = NN_l / RR_0, SE15 eq. 10 reads
    x_k = xi_k + sum_l sum_{j>=2} (2k+1) (l j k; 0 0 0)^2 f_j xi_l,
so xi solves (I + C f) xi = x per bin, where only even k, l, j enter (the odd
ones vanish for an NN that is even in mu).
"""

import math

import numpy as np


def wigner3j_000(j1, j2, j3):
    """(j1 j2 j3; 0 0 0) from the closed form for J = j1+j2+j3 even:
    (-1)^(J/2) sqrt[(J-2j1)!(J-2j2)!(J-2j3)!/(J+1)!] (J/2)!/[(J/2-j1)!(J/2-j2)!(J/2-j3)!]."""
    J = j1 + j2 + j3
    if J % 2 or j1 + j2 < j3 or j2 + j3 < j1 or j3 + j1 < j2:
        return 0.0
    g = J // 2
    f = math.factorial
    root = math.sqrt(f(J - 2 * j1) * f(J - 2 * j2) * f(J - 2 * j3) / f(J + 1))
    return (-1) ** g * root * f(g) / (f(g - j1) * f(g - j2) * f(g - j3))


def coupling_tensor(num_ell):
    """C[k, l, j] = (2k+1) (2l 2j 2k; 0 0 0)^2 over even multipoles 2k, 2l, 2j."""
    C = np.zeros((num_ell, num_ell, num_ell))
    for k in range(num_ell):
        for l in range(num_ell):
            for j in range(1, num_ell):  # the j = 0 term is the identity
                C[k, l, j] = (4 * k + 1) * wigner3j_000(2 * l, 2 * j, 2 * k) ** 2
    return C


def compute_xi(hist_corrNN, hist_corrRR):
    """xi[num_s, num_ell] from NN and RR multipoles given as [num_s, num_ell]."""
    nn = np.asarray(hist_corrNN, np.float64)
    rr = np.asarray(hist_corrRR, np.float64)
    num_s, num_ell = rr.shape
    f = rr / rr[:, [0]]
    x = nn / rr[:, [0]]
    A = np.eye(num_ell)[None] + np.einsum("klj,sj->skl", coupling_tensor(num_ell), f)
    return np.linalg.solve(A, x[..., None])[..., 0]
