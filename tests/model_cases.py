"""Small deterministic instances of every model family, shared by the tests."""
import numpy as np

BERN = dict(N=10, y=np.array([0, 1, 0, 0, 0, 0, 0, 0, 0, 1]))  # examples/bernoulli/bernoulli.data.json
EIGHT = dict(J=8, y=np.array([28.0, 8, -3, 7, -1, 1, 18, 12]), sigma=np.array([15.0, 10, 16, 11, 9, 11, 10, 18]),
             mu_scale=10.0, tau_scale=10.0)


def logistic(seed=0, N=2000, D=10):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    return dict(N=N, D=D, X=X, y=y)


def hier_normal(seed=0, J=12):
    rng = np.random.default_rng(seed)
    sigma = rng.uniform(5, 20, J)
    mu, tau = rng.normal(0, 5), abs(rng.standard_cauchy() * 5) % 20
    theta = mu + tau * rng.standard_normal(J)
    return dict(J=J, y=rng.normal(theta, sigma), sigma=sigma)


def hier_logistic(seed=0, N=300, K=3, G=6):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, K))
    g = rng.integers(1, G + 1, N)
    beta = rng.normal(0.3, 0.5, (G, K))
    eta = np.einsum("nk,nk->n", X, beta[g - 1])
    y = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(np.int64)
    return dict(N=N, K=K, G=G, X=X, group=g, y=y)


def gauss(seed=0, D=5):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((D, D))
    P = A @ A.T / D + np.eye(D) * 0.5
    return dict(D=D, mu=rng.standard_normal(D), P=P)


CASES = {
    "bernoulli": BERN,
    "eight_schools_cp": EIGHT,
    "hier_normal_nc": hier_normal(),
    "hier_logistic_cp": hier_logistic(),
    "logistic_regression": logistic(),
    "gauss": gauss(),
}
