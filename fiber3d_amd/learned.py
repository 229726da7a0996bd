"""Learned GMRES initial guess: a per-node MLP trained on the solver's own steps.

No reference counterpart (the reference always starts GMRES from zero,
python/tstep/tstep.py:1209-1216).  The plain zero-guess solve stays the
fallback: fiber3d_amd.gmres drops a guess that is not finite or whose residual
exceeds ||b||, and the FP64 residual alone decides convergence.

Network: 3 layers, width 256, ReLU, applied independently to every fiber node
(predicting the tension change T_n - T_{n-1}) and every shell node (predicting
coefficients of the density change in a frame built from local vectors).
Forward and backward products run on MFMA in BF16 with FP32 accumulation
(libf3d f3d_gemm_bf16); master weights and the Adam update are FP32.

Guess assembly per step: positions extrapolated linearly from the last two
steps, tension = T_{n-1} + net(fiber features), shell density =
q_{n-1} + sum_k c_k u_k with the net giving the c_k and u_k = (q_{n-1},
v_n, n (n . v_n)) the previous density, the explicit fiber flow on the shell
node and its normal component.
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


class LearnedGuess:
    """Initial-guess predictor for TimeStepper (fiber3d_amd.tstep)."""

    def __init__(self, train_steps=100, epochs=200, lr=1e-3, batch=8192, seed=0,
                 use_net=True):
        self.train_steps = train_steps
        self.epochs, self.lr, self.batch, self.seed = epochs, lr, batch, seed
        self.use_net = use_net
        self.samples_f = []
        self.samples_s = []
        self.prev = []  # last solutions
        self.net_f = self.net_s = None
        self.trained = False
        self.losses = []

    # ---- features ---------------------------------------------------------
    def _fiber_features(self, ts):
        S = ts.sys
        dev = S.dev
        feats = []
        v = self._explicit_flow(ts)
        for N, (node, X, xs, xss, D1, P, B) in S._ctx.items():
            vk = v[node]
            alpha = S.buckets[N]["alpha"][None, :].expand(B, N)
            dv = torch.matmul(D1, vk)
            T = ts.tension[node]
            f = torch.stack([alpha, alpha.abs(), T, (xs * dv).sum(-1), (xss * vk).sum(-1),
                             (xss * xss).sum(-1), (xs * vk).sum(-1), vk.norm(dim=-1)], dim=-1)
            feats.append((node.reshape(-1), f.reshape(-1, f.shape[-1])))
        out = torch.empty((S.npts, feats[0][1].shape[1]), dtype=torch.float64, device=dev)
        for node, f in feats:
            out[node] = f
        return out

    def _explicit_flow(self, ts):
        S = ts.sys
        # the per-step RHS carries the explicit flows; rebuild v on fiber nodes
        # from the force used by the RHS (FP64 pair sum, same-fiber exclusion)
        from .kernels import pairsum
        return pairsum("oseen", S.X, S.force, S.X, src_w=S.weights, excl=S.excl, eta=S.eta,
                       reg=S.reg, eps=S.eps)

    def _shell_basis(self, ts):
        S = ts.sys
        q = ts.shell_density.view(S.nper, 3)
        v = -S.rhs[:3 * S.nper].view(S.nper, 3)  # explicit fiber flow on the shell
        n = S.shell.nrm
        u = torch.stack([q, v, n * (n * v).sum(-1, keepdim=True)], dim=1)  # (nper, 3, 3)
        inv = torch.stack([(q * q).sum(-1).sqrt(), (v * v).sum(-1).sqrt(), (n * v).sum(-1),
                           (n * q).sum(-1)], dim=-1)
        return u, inv

    # ---- predictor protocol ------------------------------------------------
    def predict(self, ts):
        S = ts.sys
        if len(self.prev) < 2:
            return None
        x0 = self.prev[-1].clone()
        # positions: linear extrapolation of the last two solutions
        x_idx = ts.idx_x.reshape(-1)
        x0[x_idx] = 2.0 * self.prev[-1][x_idx] - self.prev[-2][x_idx]
        if self.trained and self.use_net:
            with torch.no_grad():
                ff = self.nf(self._fiber_features(ts).float())
                dT = self.net_f.forward(ff)[:, 0].double() * self.t_scale
                x0[ts.idx_t] = ts.tension + dT
                if S.nper:
                    u, inv = self._shell_basis(ts)
                    c = self.net_s.forward(self.ns(inv.float())).double() * self.c_scale
                    x0[:3 * S.nper] = (u * c[:, :, None]).sum(1).reshape(-1) + ts.shell_density
        self._pending = (self._fiber_features(ts), self._shell_basis(ts) if S.nper else None,
                         ts.tension.clone(), ts.shell_density.clone()) if not self.trained else None
        return x0

    def observe(self, ts, sol):
        S = ts.sys
        self.prev = (self.prev + [sol.clone()])[-2:]
        if self.trained or getattr(self, "_pending", None) is None:
            return
        ff, sb, T_prev, q_prev = self._pending
        T_new = sol[ts.idx_t]
        self.samples_f.append((ff, (T_new - T_prev)[:, None]))
        if sb is not None:
            u, inv = sb
            dq = sol[:3 * S.nper].view(S.nper, 3) - q_prev.view(S.nper, 3)
            # least-squares coefficients of dq in the basis u (3 vectors)
            G = torch.einsum("nai,nbi->nab", u, u) + 1e-12 * torch.eye(3, device=u.device)
            rhs = torch.einsum("nai,ni->na", u, dq)
            c = torch.linalg.solve(G, rhs[..., None])[..., 0]
            self.samples_s.append((inv, c))
        if len(self.samples_f) >= self.train_steps:
            self.train()

    def train(self):
        torch.manual_seed(self.seed)
        Xf = torch.cat([a for a, _ in self.samples_f]).float()
        Yf = torch.cat([b for _, b in self.samples_f]).float()
        self.nf = Normalizer()
        self.nf.fit(Xf)
        self.t_scale = float(Yf.abs().max().clamp_min(1e-30))
        self.net_f = MLP(Xf.shape[1], 1, seed=self.seed)
        self.losses = self._fit(self.net_f, self.nf(Xf), (Yf / self.t_scale).contiguous())
        if self.samples_s:
            Xs = torch.cat([a for a, _ in self.samples_s]).float()
            Ys = torch.cat([b for _, b in self.samples_s]).float()
            self.ns = Normalizer()
            self.ns.fit(Xs)
            self.c_scale = float(Ys.abs().max().clamp_min(1e-30))
            self.net_s = MLP(Xs.shape[1], 3, seed=self.seed + 1)
            self._fit(self.net_s, self.ns(Xs), (Ys / self.c_scale).contiguous())
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
