"""Synthetic workloads of the measurements and tests (inputs only; no computation of
the step path happens here)."""
import numpy as np

SEED = 20261203


def synthetic_history(theta, T1=None, t_scale=1e3, seed=SEED, noise=1e-3):
    """Seeded synthetic history of the kernel microbenchmark.

    Rows k = 0..T1-1 at t_k = t_scale * theta_k (theta resampled to T1 points when
    T1 != N); QK_kj = exp(-|theta_k - theta_j| u_k) (1 + noise xi_kj), u_k ~ U[0.5, 2],
    xi ~ U[-1, 1]; QR likewise with independent draws; dQ_k = Q_k - Q_{k-1}
    (dQ_0 = 0); rvec_k = 1 + theta_k; drvec_k = r_k - r_{k-1}; delta_t_ratio as
    io_input.cpp:169-172.  Generator: numpy PCG64(seed).
    """
    theta = np.asarray(theta, np.float64)
    N = len(theta)
    T1 = N if T1 is None else T1
    thk = theta if T1 == N else np.interp(np.linspace(0, N - 1, T1), np.arange(N), theta)
    rng = np.random.Generator(np.random.PCG64(seed))
    t1 = t_scale * thk
    if np.any(np.diff(t1) <= 0):
        raise ValueError("synthetic t1grid is not strictly increasing")

    def field():
        u = rng.uniform(0.5, 2.0, size=T1)
        xi = rng.uniform(-1.0, 1.0, size=(T1, N))
        return np.exp(-np.abs(thk[:, None] - theta[None, :]) * u[:, None]) * (1 + noise * xi)

    QK, QR = field(), field()
    dQK, dQR = np.zeros_like(QK), np.zeros_like(QR)
    dQK[1:] = QK[1:] - QK[:-1]
    dQR[1:] = QR[1:] - QR[:-1]
    r = 1 + thk
    dr = np.zeros(T1)
    dr[1:] = r[1:] - r[:-1]
    rho = np.zeros(T1)
    for i in range(2, T1):
        rho[i] = (t1[i] - t1[i - 1]) / (t1[i - 1] - t1[i - 2])
    return dict(t1grid=t1, dtratio=rho, rvec=r, drvec=dr, QK=QK.ravel(), QR=QR.ravel(), dQK=dQK.ravel(),
                dQR=dQR.ravel())


BATCH_SEED = 20261203


def batch_workload_spec(n=64, seed=BATCH_SEED):
    """Parameter points of the batched-quench workload: bath temperature Gamma ~ U[0.3, 1.0],
    lambda ~ U[0, 1] and the grid length L drawn 50/50 from {256, 512} (a seeded
    permutation), numpy PCG64(seed).  Returns (Gamma[n], lambda[n], L[n])."""
    rng = np.random.default_rng(seed)
    Gam = rng.uniform(0.3, 1.0, n)
    lam = rng.uniform(0.0, 1.0, n)
    Ls = rng.permutation(np.array([256] * (n // 2) + [512] * (n - n // 2)))
    return Gam, lam, Ls
