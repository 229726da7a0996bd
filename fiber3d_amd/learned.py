"""Learned GMRES initial guess: a per-node MLP trained on the solver's own steps.

No reference counterpart (the reference always starts GMRES from zero,
python/tstep/tstep.py:1209-1216).  The plain zero-guess solve stays the
fallback: fiber3d_amd.gmres drops a guess that is not finite or whose residual
exceeds ||b||, and the FP64 residual alone decides convergence.

Network: 3 layers, width 256, ReLU, applied independently to every fiber node
(predicting the correction of the extrapolated position in a local frame) and
every shell node (predicting coefficients of the density correction in a basis
built from local vectors).  Forward and backward products run on MFMA in BF16
with FP32 accumulation (libf3d f3d_gemm_bf16); master weights and the Adam
update are FP32.

Guess assembly per step: the previous solutions extrapolated (positions by a
least-squares quadratic in time, the rest linearly), corrected by the nets
(see LearnedGuess).
"""
import numpy as np
import torch

from . import _lib as L
from .kernels import _stream

EPI_STORE, EPI_BIAS_RELU, EPI_RELU_GRAD, EPI_ATOMIC, EPI_BIAS = 0, 1, 2, 3, 4


def gemm(A, B, C, epi=EPI_STORE, bias=None, Z=None, Zin=None, alpha=1.0, ksplit=1,
         transA=False, transB=False):
    """C = epi(alpha * op(A) op(B)) on MFMA (BF16 in, FP32 accumulate).

    A, B, C: 2-D float32 CUDA tensors; op(X) = X.t() when trans is set."""
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    Kb = B.shape[1] if transB else B.shape[0]
    N = B.shape[0] if transB else B.shape[1]
    assert K == Kb and C.shape == (M, N), (A.shape, B.shape, C.shape, transA, transB)
    for t in (A, B, C, bias, Z, Zin):
        assert t is None or t.dtype == torch.float32
    sam, sak = (A.stride(1), A.stride(0)) if transA else (A.stride(0), A.stride(1))
    sbk, sbn = (B.stride(1), B.stride(0)) if transB else (B.stride(0), B.stride(1))
    rc = L.load().f3d_gemm_bf16(epi, M, N, K, ksplit, L.ptr(A), sam, sak, L.ptr(B), sbk, sbn,
                                L.ptr(C), C.stride(0), C.stride(1), L.ptr(bias), L.ptr(Z),
                                L.ptr(Zin), float(alpha), _stream())
    L.check(rc, "f3d_gemm_bf16")
    return C


class MLP:
    """in -> 256 -> 256 -> out, ReLU; FP32 master weights, MFMA BF16 products."""

    def __init__(self, n_in, n_out, width=256, seed=0, device="cuda"):
        g = torch.Generator().manual_seed(seed)
        def init(fi, fo):
            return (torch.randn(fi, fo, generator=g, dtype=torch.float32) * np.sqrt(2.0 / fi)).to(device)
        self.W = [init(n_in, width), init(width, width), init(width, n_out) * 0.1]
        self.b = [torch.zeros(width, device=device), torch.zeros(width, device=device),
                  torch.zeros(n_out, device=device)]
        self.params = self.W + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.t = 0
        self.n_in, self.n_out, self.width = n_in, n_out, width

    def forward(self, X, keep=False):
        M = X.shape[0]
        dev = X.device
        Z1 = torch.empty((M, self.width), device=dev)
        H1 = torch.empty_like(Z1)
        gemm(X, self.W[0], H1, EPI_BIAS_RELU, bias=self.b[0], Z=Z1)
        Z2 = torch.empty_like(Z1)
        H2 = torch.empty_like(Z1)
        gemm(H1, self.W[1], H2, EPI_BIAS_RELU, bias=self.b[1], Z=Z2)
        Y = torch.empty((M, self.n_out), device=dev)
        gemm(H2, self.W[2], Y, EPI_BIAS, bias=self.b[2])
        if keep:
            self._cache = (X, Z1, H1, Z2, H2)
        return Y

    def backward(self, dY):
        """Gradients of the loss wrt all parameters, given dL/dY."""
        X, Z1, H1, Z2, H2 = self._cache
        M = X.shape[0]
        ks = max(1, min(64, M // 1024))
        grads = [torch.zeros_like(p) for p in self.params]
        gemm(H2, dY, grads[2], EPI_ATOMIC, transA=True, ksplit=ks)
        grads[5] = dY.sum(0)
        dZ2 = torch.empty_like(Z2)
        gemm(dY, self.W[2], dZ2, EPI_RELU_GRAD, Zin=Z2, transB=True)
        gemm(H1, dZ2, grads[1], EPI_ATOMIC, transA=True, ksplit=ks)
        grads[4] = dZ2.sum(0)
        dZ1 = torch.empty_like(Z1)
        gemm(dZ2, self.W[1], dZ1, EPI_RELU_GRAD, Zin=Z1, transB=True)
        gemm(X, dZ1, grads[0], EPI_ATOMIC, transA=True, ksplit=ks)
        grads[3] = dZ1.sum(0)
        return grads

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for p, g, m, v in zip(self.params, grads, self.m, self.v):
            m.mul_(b1).add_(g, alpha=1 - b1)
            v.mul_(b2).addcmul_(g, g, value=1 - b2)
            mh = m / (1 - b1**self.t)
            vh = v / (1 - b2**self.t)
            p.sub_(lr * mh / (vh.sqrt() + eps))

    def train_step(self, X, target, lr=1e-3):
        Y = self.forward(X, keep=True)
        diff = Y - target
        loss = float((diff * diff).mean())
        dY = (2.0 / diff.numel()) * diff
        self.adam(self.backward(dY.contiguous()), lr=lr)
        return loss


class Normalizer:
    def __init__(self):
        self.mu = None
        self.sd = None

    def fit(self, X):
        self.mu = X.mean(0, keepdim=True)
        self.sd = X.std(0, keepdim=True).clamp_min(1e-12)

    def __call__(self, X):
        return ((X - self.mu) / self.sd).contiguous()


def lsq_weights_at(times, t_next, deg):
    """Weights w_k over values at `times` (newest first) that evaluate, at
    t_next, the least-squares polynomial of degree `deg` through them:
    x(t_next) ~ sum_k w_k x(times[k]).  Times are measured in units of the
    newest interval, which keeps the Vandermonde matrix well conditioned."""
    t = np.asarray(times, dtype=float)
    if not (1 <= deg < t.size):
        raise ValueError("need 1 <= deg < len(times)")
    h = t[0] - t[1]
    if not h > 0:
        raise ValueError("times must decrease from newest to oldest")
    tau = (t - t[0]) / h
    A = np.vander(tau, deg + 1, increasing=True)
    e = ((t_next - t[0]) / h) ** np.arange(deg + 1)
    return e @ np.linalg.pinv(A)


def lsq_weights(npts, deg):
    """Weights w_k (k = 0 newest .. npts-1 oldest) that evaluate, one step
    ahead, the least-squares polynomial of degree `deg` through the last
    `npts` equally spaced values: x_next = sum_k w_k x_{n-k}."""
    if not (1 <= deg < npts):
        raise ValueError("need 1 <= deg < npts")
    return lsq_weights_at(-np.arange(npts, dtype=float), 1.0, deg)


class LearnedGuess:
    """Initial-guess predictor for TimeStepper (fiber3d_amd.tstep).

    Baseline guess: every component extrapolated linearly from the last two
    solutions, except the fiber positions X, which come from the least-squares
    polynomial of degree fit[1] through the last fit[0] positions.  Nearly all
    of the guess's residual is in X: putting the converged X into the guess
    leaves no GMRES iteration to do, while the converged tensions or shell
    densities barely change the count.  The linear extrapolation of X
    carries a second difference dt^2 X'' that stays nearly constant from step
    to step.  A quadratic through three points removes it but multiplies the
    solves' own errors by up to 7.  The least-squares quadratic through six
    points removes it and multiplies them by at most 3.6.

    Two MLPs correct the baseline.  One per fiber node predicts the position
    correction X_n - X_fit as its components in an orthonormal local frame
    (tangent, the part of the last displacement normal to it, their cross
    product).  One per shell node predicts coefficients c_k of the density
    correction q_n - q_ext = sum_k c_k u_k in a local basis u = (dq, d2q,
    n (n . dq)) built from the density history.  Inputs are rotation-invariant
    scalars of the local geometry, the explicit flow and the solution history.
    The last layer starts at zero, so an untrained net reproduces the
    baseline.  After training, each net is kept only if it beats the baseline
    on held-out steps (validation MSE below `gate` times the baseline's).
    """

    def __init__(self, train_steps=100, epochs=200, lr=1e-3, batch=8192, seed=0,
                 use_net=True, gate=0.9, val_frac=0.2, fit=(6, 2)):
        self.fit = (int(fit[0]), int(fit[1]))
        lsq_weights(*self.fit)  # validates the shape
        self.keep = max(3, self.fit[0])
        self.train_steps = train_steps
        self.epochs, self.lr, self.batch, self.seed = epochs, lr, batch, seed
        self.use_net = use_net
        self.gate, self.val_frac = gate, val_frac
        self.samples_f = []
        self.samples_s = []
        self.prev = []   # last `keep` solutions, oldest first
        self.times = []  # their times
        self.net_f = self.net_s = None
        self.enabled_f = self.enabled_s = False
        self.trained = False
        self.report = {}

    # ---- features ---------------------------------------------------------
    def _hist(self, ts, idx):
        """Any gathered component at the last three solutions (newest first)."""
        return [self.prev[-1 - k][idx] for k in range(3)]

    def _fiber_terms(self, ts):
        """Per fiber node: 15 rotation-invariant features and the orthonormal
        frame (t, e2, e3) the position correction is expressed in."""
        S = ts.sys
        X1, X2, X3 = (p.view(-1, 3) for p in self._hist(ts, ts.idx_x))
        T1, T2, _ = self._hist(ts, ts.idx_t)
        d1 = X1 - X2
        d2 = X1 - 2.0 * X2 + X3
        t = S.xs / torch.linalg.norm(S.xs, dim=1, keepdim=True).clamp_min(1e-300)
        p = d1 - (d1 * t).sum(1, keepdim=True) * t
        pn = torch.linalg.norm(p, dim=1, keepdim=True)
        # a displacement along the tangent leaves the normal direction free:
        # take the unit axis least aligned with t, made orthogonal to it
        ax = torch.zeros_like(t)
        ax.scatter_(1, t.abs().argmin(1, keepdim=True), 1.0)
        q = ax - (ax * t).sum(1, keepdim=True) * t
        use_p = pn > 1e-12 * torch.linalg.norm(d1, dim=1, keepdim=True).clamp_min(1e-300)
        e2 = torch.where(use_p, p, q)
        e2 = e2 / torch.linalg.norm(e2, dim=1, keepdim=True).clamp_min(1e-300)
        e3 = torch.cross(t, e2, dim=1)
        frame = torch.stack([t, e2, e3], dim=1)  # (npts, 3, 3), rows are the axes
        comp = lambda v: torch.einsum("nij,nj->ni", frame, v)
        alpha = torch.empty(S.npts, dtype=torch.float64, device=S.dev)
        xss = torch.empty_like(S.xs)
        for N, (node, X, xs, xss_k, D1, P, B) in S._ctx.items():
            alpha[node.reshape(-1)] = S.buckets[N]["alpha"][None, :].expand(B, N).reshape(-1)
            xss[node.reshape(-1)] = xss_k.reshape(-1, 3)
        c1, c2, cv, cs = comp(d1), comp(d2), comp(S.v_explicit), comp(xss)
        feats = torch.cat([alpha[:, None], alpha.abs()[:, None], c1[:, :2], c2, cv, cs,
                           T1[:, None], (T1 - T2)[:, None]], dim=1)
        return feats, frame

    def _shell_terms(self, ts):
        S = ts.sys
        q1, q2, q3 = (p[:3 * S.nper].view(S.nper, 3) for p in self._hist(ts, slice(None)))
        dq = q1 - q2
        d2q = dq - (q2 - q3)
        n = S.shell.nrm
        v = -S.rhs[:3 * S.nper].view(S.nper, 3)  # explicit fiber flow on the shell
        u = torch.stack([dq, d2q, n * (n * dq).sum(-1, keepdim=True)], dim=1)  # (nper, 3, 3)
        a, b = dq.norm(dim=-1), d2q.norm(dim=-1)
        inv = torch.stack([a, b, (n * dq).sum(-1), (n * d2q).sum(-1),
                           (dq * d2q).sum(-1) / (a * b + 1e-300), v.norm(dim=-1),
                           (n * v).sum(-1)], dim=-1)
        return u, inv

    def _extrapolated(self, ts):
        """Linear extrapolation in time of every component to the time of the
        coming solution (2 p1 - p2 at equal steps); fiber positions from the
        least-squares polynomial fit once fit[0] solutions exist.  The
        weights use the solutions' actual times, so steps of changing size
        (adaptive stepping) are extrapolated correctly."""
        t_next = ts.time + ts.sys.dt
        t1, t2 = self.times[-1], self.times[-2]
        a = (t_next - t1) / (t1 - t2)
        x0 = (1.0 + a) * self.prev[-1] - a * self.prev[-2]
        npts, deg = self.fit
        if len(self.prev) >= npts:
            w = lsq_weights_at(self.times[::-1][:npts], t_next, deg)
            ix = ts.idx_x.reshape(-1)
            xf = torch.zeros(ix.numel(), dtype=x0.dtype, device=x0.device)
            for k, wk in enumerate(w):
                xf.add_(self.prev[-1 - k][ix], alpha=float(wk))
            x0[ix] = xf
        return x0

    # ---- predictor protocol ------------------------------------------------
    def predict(self, ts):
        S = ts.sys
        self._pending = None
        if len(self.prev) < 2:
            return self.prev[-1].clone() if self.prev else None
        x0 = self._extrapolated(ts)
        if len(self.prev) < self.keep or not self.use_net:
            return x0
        if not self.trained:
            sh = self._shell_terms(ts) if S.nper else None
            self._pending = (*self._fiber_terms(ts), sh, x0)
            return x0
        # trained: only the nets the gate kept are evaluated
        with torch.no_grad():
            if self.enabled_f:
                ff, frame = self._fiber_terms(ts)
                c = self.net_f.forward(self.nf(ff.float())).double() * self.x_scale
                ix = ts.idx_x
                x0[ix] = x0[ix] + torch.einsum("nij,ni->nj", frame, c)
            if S.nper and self.enabled_s:
                u, inv = self._shell_terms(ts)
                c = self.net_s.forward(self.ns(inv.float())).double() * self.c_scale
                x0[:3 * S.nper] += (u * c[:, :, None]).sum(1).reshape(-1)
        return x0

    def observe(self, ts, sol):
        S = ts.sys
        pend = getattr(self, "_pending", None)
        self.prev = (self.prev + [sol.clone()])[-self.keep:]
        self.times = (self.times + [float(ts.time)])[-self.keep:]
        if self.trained or pend is None:
            return
        ff, frame, sh, x_ext = pend
        dX = (sol - x_ext)[ts.idx_x]  # (npts, 3)
        self.samples_f.append((ff, torch.einsum("nij,nj->ni", frame, dX)))
        if sh is not None:
            u, inv = sh
            dq = (sol - x_ext)[:3 * S.nper].view(S.nper, 3)
            # least-squares coefficients of the correction in the basis u
            G = torch.einsum("nai,nbi->nab", u, u)
            G = G + 1e-12 * (G.diagonal(dim1=1, dim2=2).sum(-1) + 1e-300)[:, None, None] \
                * torch.eye(3, device=u.device, dtype=u.dtype)
            rhs = torch.einsum("nai,ni->na", u, dq)
            c = torch.linalg.solve(G, rhs[..., None])[..., 0]
            self.samples_s.append((inv, c))
        if len(self.samples_f) >= self.train_steps:
            self.train()

    # ---- training -----------------------------------------------------------
    def _split(self, samples):
        nval = max(1, int(round(self.val_frac * len(samples))))
        tr, va = samples[:-nval], samples[-nval:]
        cat = lambda ss, i: torch.cat([s[i] for s in ss]).float()
        return cat(tr, 0), cat(tr, 1), cat(va, 0), cat(va, 1)

    def _train_one(self, samples, n_out, seed):
        Xt, Yt, Xv, Yv = self._split(samples)
        norm = Normalizer()
        norm.fit(Xt)
        # per-output scale: the outputs may differ by orders of magnitude
        scale = Yt.abs().amax(0).clamp_min(1e-30)
        net = MLP(Xt.shape[1], n_out, seed=seed)
        net.W[2].zero_()
        self._fit(net, norm(Xt), (Yt / scale).contiguous())
        with torch.no_grad():
            pred = net.forward(norm(Xv)) * scale
            mse_net = float(((pred - Yv) ** 2).mean())
            mse_base = float((Yv**2).mean())
        return net, norm, scale, mse_net, mse_base

    def train(self):
        torch.manual_seed(self.seed)
        self.net_f, self.nf, self.x_scale, mn, mb = self._train_one(self.samples_f, 3, self.seed)
        self.enabled_f = mn < self.gate * mb
        self.report["fiber"] = {"val_mse_net(position)": mn, "val_mse_fit(position)": mb,
                                "enabled": self.enabled_f}
        if self.samples_s:
            self.net_s, self.ns, self.c_scale, mn, mb = self._train_one(self.samples_s, 3,
                                                                        self.seed + 1)
            self.enabled_s = mn < self.gate * mb
            self.report["shell"] = {"val_mse_net(coeff)": mn, "val_mse_extrapolation(coeff)": mb,
                                    "enabled": self.enabled_s}
        self.trained = True
        self.samples_f, self.samples_s = [], []

    def _fit(self, net, X, Y):
        n = X.shape[0]
        g = torch.Generator(device=X.device).manual_seed(self.seed)
        losses = []
        for _ in range(self.epochs):
            perm = torch.randperm(n, device=X.device, generator=g)
            for i in range(0, n, self.batch):
                idx = perm[i:i + self.batch]
                losses.append(net.train_step(X[idx].contiguous(), Y[idx].contiguous(), lr=self.lr))
        return losses
