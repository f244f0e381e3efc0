"""FP64 CPU oracle of the affine-coupling normalizing flow (test infrastructure).

Only the tests, the smoke check and the CPU baseline of the benchmark use this
module; the product path runs the HIP kernels in cmdstan_amd/csrc/flow.hip.

The flow has no counterpart in CmdStan: it follows the published algorithms
  * RealNVP affine coupling (Dinh, Sohl-Dickstein, Bengio 2017, "Density
    estimation using Real NVP", sec. 3.2-3.3): v_a = u_a,
    v_b = u_b * exp(s(u_a)) + t(u_a), log|det| = sum s;
  * Adam (Kingma & Ba 2015, Algorithm 1) with bias correction;
  * NUTS on the pulled-back density p_z(z) = p_x(T(z)) |det dT/dz| (the
    "neural transport" reparameterisation, Hoffman et al. 2019), so that
    U(z) = V(T(z)) - log|det dT/dz| and grad U = J^T grad V - grad log|det|.

Layout (shared with the HIP implementation):
  x = loc + scale * f_{L-1}( ... f_0(z))
  layer l: index sets a_l (conditioner) and b_l (transformed);
  h1 = tanh(u_a W1 + b1), h2 = tanh(h1 W2 + b2), o = h2 W3 + b3 (2 Db wide),
  s = S tanh(o[:Db] / S), t = o[Db:], S = SCALE_BOUND.
Parameters are one flat vector: per layer W1 (Da x H, row-major), b1 (H),
W2 (H x H), b2 (H), W3 (H x 2Db), b3 (2Db), layers back to back.

Pins (tests/test_flow_oracle.py): the VJP against complex-step derivatives of
U, log|det| against the determinant of a finite-difference Jacobian, the
inverse against the forward map, the training gradient against complex-step
derivatives of the loss, and Adam against its closed form on one step.
"""
import numpy as np

SCALE_BOUND = 3.0


def coupling_masks(D, n_layers, n_glob=0, group_size=1):
    """Index sets (a_l, b_l) per layer.

    Dimensions [0, n_glob) are global (hyper)parameters; the rest form groups
    of `group_size` consecutive dimensions.  With globals the pattern repeats
    (globals + even groups | odd groups), (globals + odd groups | even groups),
    (all groups | globals), so that group effects are transformed conditioned
    on the hyperparameters (a coupling layer can then express a
    non-centred reparameterisation).  Without globals, groups alternate.
    """
    loc = np.arange(n_glob, D)
    grp = (loc - n_glob) // group_size
    even, odd = loc[grp % 2 == 0], loc[grp % 2 == 1]
    glob = np.arange(n_glob)
    if len(odd) == 0:  # a single group: split it in half
        h = len(loc) // 2
        even, odd = loc[:h], loc[h:]
    out = []
    for l in range(n_layers):
        if n_glob > 0:
            k = l % 3
            if k == 0:
                a, b = np.concatenate([glob, even]), odd
            elif k == 1:
                a, b = np.concatenate([glob, odd]), even
            else:
                a, b = loc, glob
        else:
            a, b = (even, odd) if l % 2 == 0 else (odd, even)
        out.append((np.sort(a).astype(np.int64), np.sort(b).astype(np.int64)))
    return out


def param_layout(masks, H):
    """[(offset, shapes)] per layer: W1, b1, W2, b2, W3, b3."""
    lay, off = [], 0
    for a, b in masks:
        Da, Db = len(a), len(b)
        shapes = [(Da, H), (H,), (H, H), (H,), (H, 2 * Db), (2 * Db,)]
        offs = []
        for s in shapes:
            offs.append(off)
            off += int(np.prod(s))
        lay.append((offs, shapes))
    return lay, off


def init_params(masks, H, seed):
    """Glorot-uniform W1, W2; zero W3, b3 (the flow starts as the identity)."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout(masks, H)
    p = np.zeros(n)
    for (offs, shapes) in lay:
        for j in (0, 2):
            fan_in, fan_out = shapes[j]
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            p[offs[j]:offs[j] + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
    return p


class Flow:
    def __init__(self, D, masks, H, params, loc=None, scale=None):
        self.D, self.masks, self.H = D, masks, H
        self.lay, self.n_params = param_layout(masks, H)
        self.p = np.asarray(params)
        self.loc = np.zeros(D) if loc is None else np.asarray(loc, dtype=float)
        self.scale = np.ones(D) if scale is None else np.asarray(scale, dtype=float)

    def weights(self, l, p=None):
        p = self.p if p is None else p
        offs, shapes = self.lay[l]
        return [p[o:o + int(np.prod(s))].reshape(s) for o, s in zip(offs, shapes)]

    def _mlp(self, l, ua, p=None):
        W1, b1, W2, b2, W3, b3 = self.weights(l, p)
        h1 = np.tanh(ua @ W1 + b1)
        h2 = np.tanh(h1 @ W2 + b2)
        o = h2 @ W3 + b3
        Db = len(self.masks[l][1])
        s = SCALE_BOUND * np.tanh(o[:, :Db] / SCALE_BOUND)
        return h1, h2, s, o[:, Db:]

    # ------------------------------------------------------------- sampling
    def forward(self, z, p=None):
        """x = T(z) and log|det dT/dz| for rows of z."""
        u = np.array(z, dtype=np.result_type(z, self.p if p is None else p, 1.0), copy=True)
        logdet = np.zeros(u.shape[0], dtype=u.dtype)
        for l, (a, b) in enumerate(self.masks):
            _, _, s, t = self._mlp(l, u[:, a], p)
            u[:, b] = u[:, b] * np.exp(s) + t
            logdet = logdet + s.sum(axis=1)
        x = self.loc + self.scale * u
        return x, logdet + np.log(self.scale).sum()

    def pullback_grad(self, z, gx):
        """d/dz of V(T(z)) - log|det| given gx = grad V at x = T(z)."""
        u = np.array(z, dtype=float, copy=True)
        tape = []
        for l, (a, b) in enumerate(self.masks):
            h1, h2, s, t = self._mlp(l, u[:, a])
            tape.append((u.copy(), h1, h2, s))
            u[:, b] = u[:, b] * np.exp(s) + t
        g = np.asarray(gx, dtype=float) * self.scale
        for l in reversed(range(len(self.masks))):
            a, b = self.masks[l]
            uin, h1, h2, s = tape[l]
            W1, b1, W2, b2, W3, b3 = self.weights(l)
            es = np.exp(s)
            gvb = g[:, b]
            gs = gvb * uin[:, b] * es - 1.0
            graw = gs * (1.0 - (s / SCALE_BOUND) ** 2)
            go = np.concatenate([graw, gvb], axis=1)
            gpre2 = (go @ W3.T) * (1.0 - h2 ** 2)
            gpre1 = (gpre2 @ W2.T) * (1.0 - h1 ** 2)
            gua = gpre1 @ W1.T
            gn = g.copy()
            gn[:, b] = gvb * es
            gn[:, a] += gua
            g = gn
        return g

    # ------------------------------------------------------------- training
    def inverse(self, x, p=None):
        """z = T^{-1}(x) and log|det dT^{-1}/dx|."""
        v = (np.asarray(x) - self.loc) / self.scale
        v = np.array(v, dtype=np.result_type(v, self.p if p is None else p, 1.0), copy=True)
        logdet = np.zeros(v.shape[0], dtype=v.dtype) - np.log(self.scale).sum()
        for l in reversed(range(len(self.masks))):
            a, b = self.masks[l]
            _, _, s, t = self._mlp(l, v[:, a], p)
            v[:, b] = (v[:, b] - t) * np.exp(-s)
            logdet = logdet - s.sum(axis=1)
        return v, logdet

    def loss(self, x, p=None):
        """Mean negative log-density of x under the flow, without the
        constant D/2 log(2 pi)."""
        z, ld = self.inverse(x, p)
        return np.mean(0.5 * np.sum(z * z, axis=1) - ld)

    def loss_grad(self, x):
        """Loss and its gradient with respect to the flat parameters."""
        v = (np.asarray(x, dtype=float) - self.loc) / self.scale
        tape = [None] * len(self.masks)
        for l in reversed(range(len(self.masks))):
            a, b = self.masks[l]
            h1, h2, s, t = self._mlp(l, v[:, a])
            tape[l] = (v.copy(), h1, h2, s, t)
            v = v.copy()
            v[:, b] = (v[:, b] - t) * np.exp(-s)
        z = v
        B = z.shape[0]
        grad = np.zeros(self.n_params)
        gu = z / B
        for l in range(len(self.masks)):
            a, b = self.masks[l]
            vin, h1, h2, s, t = tape[l]
            offs, shapes = self.lay[l]
            W1, b1, W2, b2, W3, b3 = self.weights(l)
            ems = np.exp(-s)
            ub = (vin[:, b] - t) * ems  # z-side output of this layer
            gub = gu[:, b]
            gvb = gub * ems
            gt = -gvb
            gs = -gub * ub + 1.0 / B
            graw = gs * (1.0 - (s / SCALE_BOUND) ** 2)
            go = np.concatenate([graw, gt], axis=1)
            gpre2 = (go @ W3.T) * (1.0 - h2 ** 2)
            gpre1 = (gpre2 @ W2.T) * (1.0 - h1 ** 2)
            for j, G in enumerate([vin[:, a].T @ gpre1, gpre1.sum(0), h1.T @ gpre2, gpre2.sum(0), h2.T @ go,
                                   go.sum(0)]):
                grad[offs[j]:offs[j] + G.size] = G.ravel()
            gva = gpre1 @ W1.T
            gn = gu.copy()
            gn[:, b] = gvb
            gn[:, a] += gva
            gu = gn
        return self.loss(x), grad


def adam_step(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Kingma & Ba 2015, Algorithm 1 (step t starts at 1)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
