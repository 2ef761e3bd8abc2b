"""Pins of the oracle's implicit-diffusion operator (MLABecLaplacian with
constant coefficients: alpha a phi - beta b lap(phi) = rhs, the solve of
DiffusionOps.cpp:154-160 with setScalars(1, dt), a = rho, b = mu).

* A dense solve of the 7-point matrix assembled here from the BC rules
  (periodic wrap; Neumann: ghost = first cell; Dirichlet with the zero value
  on the face: ghost = minus the first cell), for three BC sets.
* A discrete Fourier mode on a periodic box: the solution is the mode over
  the exact eigenvalue alpha a + beta b sum_d 4 sin^2(pi k_d / n_d) / h_d^2.
* apply(solve(rhs)) = rhs, and the solve converges on the ABL aspect ratio
  with semicoarsening.
"""
import itertools

import numpy as np
import pytest

import oracle

oracle.build()
P_, N_, D_ = 0, 1, 2


def _dense(n, h, lobc, hibc, alpha, beta):
    nn = int(np.prod(n))
    A = np.zeros((nn, nn))

    def idx(i, j, k):
        return i + n[0] * (j + n[1] * k)

    for k, j, i in itertools.product(range(n[2]), range(n[1]), range(n[0])):
        r = idx(i, j, k)
        A[r, r] += alpha
        c = (i, j, k)
        for d in range(3):
            w = beta / (h[d] * h[d])
            for s, bc in ((-1, lobc[d]), (1, hibc[d])):
                q = list(c)
                q[d] += s
                A[r, r] += w
                if 0 <= q[d] < n[d]:
                    A[r, idx(*q)] -= w
                elif bc == P_:
                    q[d] %= n[d]
                    A[r, idx(*q)] -= w
                elif bc == N_:
                    A[r, r] -= w  # ghost = first cell
                else:
                    A[r, r] += w  # ghost = minus the first cell
    return A


CASES = [
    ((6, 4, 5), (0.5, 0.7, 0.3), (P_, N_, D_), (P_, N_, N_), 2.0, 0.3),
    ((4, 4, 4), (1.0, 0.5, 0.25), (N_, N_, N_), (N_, N_, N_), 1.0, 0.05),
    ((8, 6, 4), (0.3, 0.3, 0.2), (P_, P_, D_), (P_, P_, D_), 0.0, 1.0),
]


@pytest.mark.parametrize("n,h,lobc,hibc,alpha,beta", CASES)
def test_abeclap_solve_equals_dense(n, h, lobc, hibc, alpha, beta):
    A = _dense(n, h, lobc, hibc, alpha, beta)
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    x = np.linalg.solve(A, b)
    phi, info = oracle.abeclap_solve(n, h, lobc, hibc, alpha, beta, b, 1e-11, 0.0)
    assert info["status"] == 0, info
    assert np.abs(phi - x).max() <= 1e-9 * np.abs(x).max()
    y = oracle.abeclap_apply(n, h, lobc, hibc, alpha, beta, x)
    assert np.abs(y - A @ x).max() <= 1e-12 * np.abs(A @ x).max()


def test_abeclap_fourier_mode():
    n, h = (16, 8, 12), (0.25, 0.5, 1.0 / 3.0)
    alpha, beta = 3.0, 0.7
    k = (3, 1, 2)
    ii = [np.arange(n[d]) for d in range(3)]
    Z, Y, X = np.meshgrid(ii[2], ii[1], ii[0], indexing="ij")
    mode = np.cos(2 * np.pi * (k[0] * X / n[0] + k[1] * Y / n[1] + k[2] * Z / n[2]))
    lam = alpha + beta * sum(4 * np.sin(np.pi * k[d] / n[d]) ** 2 / h[d] ** 2 for d in range(3))
    phi, info = oracle.abeclap_solve(n, h, (P_,) * 3, (P_,) * 3, alpha, beta, mode.ravel(), 1e-12, 0.0)
    assert info["status"] == 0
    assert np.abs(phi - mode.ravel() / lam).max() <= 1e-11


def test_abeclap_abl_aspect_converges():
    n, h = (32, 32, 16), (20.0, 20.0, 4.0)
    rho, mu, dt = 1.2, 50.0, 0.5  # alpha a = rho, beta b = dt mu
    b = np.random.default_rng(9).standard_normal(int(np.prod(n)))
    opts = oracle.default_opts(semicoarsening=1)
    for lobc, hibc in (((P_, P_, D_), (P_, P_, N_)), ((P_, P_, D_), (P_, P_, D_))):
        phi, info = oracle.abeclap_solve(n, h, lobc, hibc, rho, dt * mu, b, 1e-11, 0.0, opts)
        assert info["status"] == 0 and info["iters"] <= 20, info
        y = oracle.abeclap_apply(n, h, lobc, hibc, rho, dt * mu, phi)
        assert np.abs(y - b).max() <= 1e-10 * np.abs(b).max()
