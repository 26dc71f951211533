"""Seeded synthetic workloads of the DA-GTM benchmarks, and the model set-up.

gmm(N, D, seed)            3-component Gaussian mixture (means ~ N(0, 4I),
                           per-component diagonal std ~ U(0.3, 1.0), equal
                           weights) -- the N=1000, D=12 CPU-runnable case.
fingerprints(N, D, ...)    binary 0/1 fingerprint-like vectors: P prototypes
                           with per-bit p ~ Beta(0.5, 2) and bits ~ Bernoulli(p);
                           each point copies a random prototype and flips
                           every bit with probability 0.05.
latent_basis(L, K, M, s)   latent grid X and RBF basis FI (+bias column) of
                           the reference's dagtm_stp (comm.c:992-1103,
                           1453-1536), written with numpy.
batch_instances(n, seed)   the batched workload: n independent instances of
                           uneven N, D, K (see the function).
"""
import numpy as np


def gmm(N=1000, D=12, seed=42, ncomp=3):
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, 2.0, size=(ncomp, D))
    sds = rng.uniform(0.3, 1.0, size=(ncomp, D))
    comp = rng.integers(0, ncomp, size=N)
    return means[comp] + sds[comp] * rng.standard_normal((N, D))


def fingerprints(N=100_000, D=166, seed=7, nproto=50, flip=0.05, chunk=1 << 18):
    rng = np.random.default_rng(seed)
    p = rng.beta(0.5, 2.0, size=(nproto, D))
    proto = (rng.random((nproto, D)) < p).astype(np.uint8)
    out = np.empty((N, D), dtype=np.uint8)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        pick = rng.integers(0, nproto, size=e - s)
        fl = (rng.random((e - s, D)) < flip).astype(np.uint8)
        out[s:e] = proto[pick] ^ fl
    return out


def rctg(g, L):
    """Grid of g^L points in [-1,1]^L, first coordinate fastest (comm.c:1453-1504)."""
    axes = np.arange(g, dtype=np.float64)
    grids = np.meshgrid(*([axes] * L), indexing="ij")   # grids[0] varies slowest
    X = np.stack([grids[L - 1 - l].ravel() for l in range(L)], axis=1)
    return X * (2.0 / (g - 1)) + (-1.0)


def latent_basis(L, K, M, s=1.0):
    """X (K x L) and FI (K x (M+1)) with RBF width sigma = s * centre spacing."""
    gx = int(round(K ** (1.0 / L)))
    gm = int(round(M ** (1.0 / L)))
    if gx ** L != K or gm ** L != M:
        raise ValueError("K and M must be perfect L-th powers")
    X = rctg(gx, L)
    MU = rctg(gm, L) * (gm / (gm - 1))
    sigma = s * (MU[1, 0] - MU[0, 0])
    d = (X * X).sum(1)[:, None] - 2.0 * X @ MU.T + (MU * MU).sum(1)[None, :]
    FI = np.ones((K, M + 1))
    FI[:, :M] = np.exp(d * (-0.5 / sigma ** 2))
    return X, FI


def pca_init(T, FI, X):
    """W0 mapping the latent grid onto the top-L principal plane of T
    (standard GTM initialisation), and beta0 from the discarded variance /
    centre spacing.  Returns W0, Y0 = FI W0, beta0."""
    T = np.asarray(T, dtype=np.float64)
    L = X.shape[1]
    mu = T.mean(0)
    C = np.cov(T, rowvar=False)
    ev, V = np.linalg.eigh(C)
    order = np.argsort(ev)[::-1]
    ev, V = ev[order], V[:, order]
    target = mu + (X / X.std(0)) @ (V[:, :L] * np.sqrt(np.maximum(ev[:L], 0.0))).T
    W0 = np.linalg.lstsq(FI, target, rcond=None)[0]
    Y0 = FI @ W0
    dd = ((Y0[:, None, :] - Y0[None, :min(len(Y0), 2000), :]) ** 2).sum(-1) if len(Y0) <= 2000 else None
    interdist = np.inf
    if dd is not None:
        np.fill_diagonal(dd, np.inf)
        interdist = dd.min(1).mean() / 2.0
    var = ev[L] if L < len(ev) else ev[-1]
    beta0 = 1.0 / max(min(interdist, var), 1e-12) if np.isfinite(interdist) else 1.0 / max(var, 1e-12)
    return W0, Y0, float(beta0)


def batch_instances(n=64, seed=2024):
    """n independent instances: N log-uniform in [1e3, 2e5]; D from {12, 32,
    64, 166} (Gaussian-mixture data for D <= 64, binary fingerprints for 166);
    (L, K, M) from {(2, 400, 25), (2, 1600, 25), (3, 8000, 125)}; independent
    seeds and start temperatures in [1.5, 4).  Returns a list of (T, FI, Y0,
    beta0, {"start_temp": t0})."""
    rng = np.random.default_rng(seed)
    shapes = [(2, 400, 25), (2, 1600, 25), (3, 8000, 125)]
    out = []
    for _ in range(n):
        N = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        D = int(rng.choice([12, 32, 64, 166]))
        L, K, M = shapes[int(rng.integers(0, 3))]
        s = int(rng.integers(1, 1 << 30))
        T = (fingerprints(N, D, s).astype(np.float64) if D == 166 else gmm(N, D, s))
        X, FI = latent_basis(L, K, M, 1.0)
        W0, Y0, b0 = pca_init(T, FI, X)
        out.append((T, FI, Y0, b0, dict(start_temp=float(rng.uniform(1.5, 4.0)))))
    return out

