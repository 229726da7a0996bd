"""Learned GMRES initial guess: a per-node MLP trained on the solver's own steps.

No reference counterpart (the reference always starts GMRES from zero,
python/tstep/tstep.py:1209-1216).  The plain zero-guess solve stays the
fallback: fiber3d_amd.gmres drops a guess that is not finite or whose residual
exceeds ||b||, and the FP64 residual alone decides convergence.

Network: 3 layers, width 256, ReLU, applied independently to every fiber
node, predicting the correction of the extrapolated position (in a local
frame) and tension from the local geometry, the history and the solver's
residual signal.  Forward and backward products run on MFMA in BF16 with FP32
accumulation (libf3d f3d_gemm_bf16); master weights and the Adam update are
FP32.

Guess assembly per step: the previous solutions extrapolated (positions by a
least-squares quadratic in time, the rest linearly), corrected by the net
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
    polynomial of degree fit[1] through the last fit[0] positions.  The linear
    extrapolation of X carries a second difference dt^2 X'' that stays nearly
    constant from step to step; the least-squares quadratic through six
    points removes it and multiplies the solves' own errors by at most 3.6.

    Residual correction.  After each solve the stepper hands the predictor
    c = P^-1 r, the preconditioner applied to the residual vector the solve
    closed with (|r| <= tol |b|); x + c approximates the exact solution better
    than x by about the preconditioner's quality.  Solutions are accepted
    just below the tolerance, and extrapolating them passes their errors on:
    the plain guess's residual sits near the tolerance.  correct=True
    extrapolates the corrected history x_k + c_k instead.  How much the
    correction helps depends on how closely the preconditioner resembles the
    operator: with the block-diagonal shell block of a large periphery it
    can make the guess worse.

    The net (use_net=True): one MLP per fiber node predicts the correction
    of the baseline guess.  The baseline is the corrected extrapolation when
    the correction has lowered the residual it was applied to, on average
    over the last calib_steps sampling steps (the stepper measures the gain
    |r - A P^-1 r| / |r| in FP64 there), else the plain one.  Its
    target is the corrected solution x + P^-1 r minus that guess: the
    position part in an orthonormal local frame (tangent, the part of the
    last displacement normal to it, their cross product) and the tension
    part.  Its features are rotation-invariant: the local geometry, the
    explicit flow, the solution history and the solver's residual signal
    (the corrected-minus-plain extrapolation and the last correction, in the
    same frame).  Its correction is projected, fiber by
    fiber, onto the first `modes` Legendre polynomials of the fiber
    parameter: the fiber operator's fourth derivatives amplify node-to-node
    noise by orders of magnitude, so only smooth corrections help.  Its last
    layer starts at zero, so an untrained net reproduces the baseline.  After
    training it must beat the baseline twice on held-out data: its validation
    MSE below `gate` times the baseline's on the last val_frac of the
    training steps, then its guesses' FP64 residuals below `gate` times the
    baseline guesses' over the next `shadow_steps`
    steps, during which the baseline still drives the solves.  Other
    components keep the plain extrapolation.
    """

    def __init__(self, train_steps=100, epochs=200, lr=1e-3, batch=8192, seed=0,
                 use_net=True, gate=0.9, val_frac=0.2, fit=(6, 2), correct=False,
                 max_samples=None, modes=8, shadow_steps=4, calib_steps=4):
        self.fit = (int(fit[0]), int(fit[1]))
        lsq_weights(*self.fit)  # validates the shape
        self.keep = max(3, self.fit[0])
        self.train_steps = train_steps
        self.epochs, self.lr, self.batch, self.seed = epochs, lr, batch, seed
        self.use_net = use_net
        self.correct = bool(correct) and not use_net
        self.gate, self.val_frac = gate, val_frac
        # training samples kept per step (a seeded random subset of the fiber
        # nodes when there are more)
        self.max_samples = max_samples
        self._rng = torch.Generator().manual_seed(seed)
        self.samples_f = []
        # history ring buffers (keep x n): solutions and corrected solutions
        # x + P^-1 r; _rows lists the rows in use, oldest first
        self._H = self._Hc = None
        self._rows = []
        self._cor_ok = []  # per row in use: was a correction given
        self._last_corr = None
        self._wcache = {}
        self._mask_x = None
        self.times = []  # their times
        self.net_f = None
        self.enabled_f = False
        self.modes = modes
        self.shadow_steps = shadow_steps
        # which extrapolation the net corrects: decided at training time by
        # the correction gains measured on the last calib_steps sampling steps
        self.calib_steps = calib_steps
        self.base_mode = None
        self._calib = []
        self.shadow_left = 0
        self.shadow = []
        self.trained = False
        self.report = {}

    @property
    def warm_steps(self):
        """Steps before the predictor's final form: the history, the
        training samples (the baseline is chosen on the last calib_steps of
        them) and the gate steps."""
        if not self.use_net:
            return self.keep + 2
        return self.keep + self.train_steps + self.shadow_steps + 2

    @property
    def wants_correction(self):
        """Whether the stepper should pass P^-1 r to observe() (one
        preconditioner apply per step, used by the corrected and net modes)."""
        return self.correct or self.use_net

    # ---- history and extrapolation -----------------------------------------
    @property
    def prev(self):
        """The last `keep` solutions, oldest first (views of the ring)."""
        return [] if self._H is None else [self._H[r] for r in self._rows]

    def _push(self, sol, corr):
        n = sol.numel()
        if self._H is None or self._H.shape[1] != n:
            self._H = torch.zeros((self.keep, n), dtype=sol.dtype, device=sol.device)
            self._Hc = torch.zeros_like(self._H)
            self._rows, self._cor_ok = [], []
        pos = (self._rows[-1] + 1) % self.keep if self._rows else 0
        self._H[pos].copy_(sol)
        if corr is not None:
            torch.add(sol, corr, out=self._Hc[pos])
        self._rows = (self._rows + [pos])[-self.keep:]
        self._cor_ok = (self._cor_ok + [corr is not None])[-self.keep:]
        self._last_corr = corr.clone() if (corr is not None and self.use_net) else None

    def _weights(self, W):
        """Device copy of the (2, keep) weight matrix, cached by value (the
        weights repeat while the step size does)."""
        key = W.tobytes()
        t = self._wcache.get(key)
        if t is None:
            if len(self._wcache) > 32:
                self._wcache.clear()
            t = torch.from_numpy(W).to(self._H.device)
            self._wcache[key] = t
        return t

    def _extrapolated(self, ts, corrected=False):
        """Linear extrapolation in time of every component of the history
        (corrected: of x + P^-1 r) to the coming solution's time (2 p1 - p2 at
        equal steps); fiber positions from the least-squares polynomial fit
        once fit[0] solutions exist.  The weights use the solutions' actual
        times, so steps of changing size are extrapolated correctly.  Both
        combinations are one (2 x keep) x (keep x n) product over the ring."""
        t_next = ts.time + ts.sys.dt
        t1, t2 = self.times[-1], self.times[-2]
        a = (t_next - t1) / (t1 - t2)
        rows = self._rows
        W = np.zeros((2, self.keep))
        W[0, rows[-1]] = 1.0 + a
        W[0, rows[-2]] = -a
        npts, deg = self.fit
        full = len(rows) >= npts
        if full:
            w = lsq_weights_at(self.times[::-1][:npts], t_next, deg)
            for k, wk in enumerate(w):
                W[1, rows[-1 - k]] = wk
        R = self._weights(W) @ (self._Hc if corrected else self._H)
        if not full:
            return R[0]
        if self._mask_x is None or self._mask_x.numel() != R.shape[1]:
            m = torch.zeros(R.shape[1], dtype=torch.bool, device=R.device)
            m[ts.idx_x.reshape(-1)] = True
            self._mask_x = m
        return torch.where(self._mask_x, R[1], R[0])

    def _have_corrected(self):
        return bool(self._cor_ok) and all(self._cor_ok)

    # ---- features -----------------------------------------------------------
    def _fiber_terms(self, ts, x_ext, x_cor):
        """Per fiber node: 23 rotation-invariant features and the orthonormal
        frame (t, e2, e3) the position correction is expressed in."""
        S = ts.sys
        ix = ts.idx_x
        X1, X2, X3 = (self.prev[-1 - k][ix] for k in range(3))
        T1, T2 = self.prev[-1][ts.idx_t], self.prev[-2][ts.idx_t]
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
        # residual signal: corrected minus plain extrapolation, and the last
        # correction (zero where no correction was given)
        sig = (x_cor - x_ext) if x_cor is not None else torch.zeros_like(x_ext)
        lc = self._last_corr if self._last_corr is not None else torch.zeros_like(x_ext)
        feats = torch.cat([alpha[:, None], alpha.abs()[:, None], c1[:, :2], c2, cv, cs,
                           T1[:, None], (T1 - T2)[:, None],
                           comp(sig[ix]), sig[ts.idx_t][:, None],
                           comp(lc[ix]), lc[ts.idx_t][:, None]], dim=1)
        return feats, frame

    # ---- predictor protocol -------------------------------------------------
    def predict(self, ts):
        self._pending = None
        if len(self._rows) < 2:
            return self._H[self._rows[-1]].clone() if self._rows else None
        have_cor = (self.correct or self.use_net) and self._have_corrected()
        if not self.use_net:
            return self._extrapolated(ts, corrected=have_cor and self.correct)
        x_ext = self._extrapolated(ts)
        x_cor = self._extrapolated(ts, corrected=True) if have_cor else None
        if len(self._rows) < self.keep:
            return x_cor if x_cor is not None else x_ext
        if not self.trained:
            # sampling: both extrapolations are kept, the corrected one (when
            # there is one) drives the solve
            self._pending = (*self._fiber_terms(ts, x_ext, x_cor), x_ext, x_cor)
            return x_cor if x_cor is not None else x_ext
        # the net corrects the better of the two extrapolations (the plain one
        # when the stepper hands over no corrections)
        base = x_cor if (x_cor is not None and self.base_mode == "corrected") else x_ext
        if self.shadow_left > 0:
            # held-out check of the trained net on the quantity that matters:
            # its guess's FP64 residual against the baseline's, while the
            # baseline guess still drives the solve
            x_net = self._net_guess(ts, base, x_ext, x_cor)
            S = ts.sys
            rn = lambda x: float(torch.linalg.norm(S.rhs - S.matvec(x)))
            self.shadow.append((rn(x_net), rn(base)))
            self.shadow_left -= 1
            if self.shadow_left == 0:
                net, bas = (float(np.mean(v)) for v in zip(*self.shadow))
                self.enabled_f = net < self.gate * bas
                self.report["fiber"].update({"held_out_guess_residual_net": net,
                                             "held_out_guess_residual_baseline": bas,
                                             "held_out_steps": len(self.shadow),
                                             "enabled": self.enabled_f})
            return base
        if self.enabled_f:
            return self._net_guess(ts, base, x_ext, x_cor)
        return base

    def _smoother(self, ts):
        """Per node count N: the fibers' node indices (B, N) and the
        least-squares projection onto the first `modes` Legendre polynomials
        in the fiber parameter (N x N).  The fiber operator's fourth
        derivatives amplify rough (node-to-node) errors by orders of
        magnitude, so the net's correction is kept to smooth modes."""
        if getattr(self, "_sm", None) is None:
            S = ts.sys
            self._sm = []
            for N, bk in S.buckets.items():
                a = bk["alpha"].cpu().numpy()
                K = min(self.modes, N)
                V = np.polynomial.legendre.legvander(a, K - 1)
                Pm = V @ np.linalg.pinv(V)
                nodes = bk["start"][:, None] + torch.arange(N, device=S.dev)[None, :]
                self._sm.append((nodes, torch.from_numpy(Pm).to(S.dev)))
        return self._sm

    def _net_guess(self, ts, base, x_ext, x_cor):
        with torch.no_grad():
            ff, frame = self._fiber_terms(ts, x_ext, x_cor)
            c = self.net_f.forward(self.nf(ff.float())).double() * self.f_scale
            d = torch.cat([torch.einsum("nij,ni->nj", frame, c[:, :3]), c[:, 3:4]], dim=1)
            for nodes, Pm in self._smoother(ts):
                d[nodes] = torch.einsum("mn,bnk->bmk", Pm, d[nodes])
            x0 = base.clone()
            ix = ts.idx_x
            x0[ix] = x0[ix] + d[:, :3]
            x0[ts.idx_t] = x0[ts.idx_t] + d[:, 3]
        return x0

    @property
    def needs_gain(self):
        """Whether the stepper should measure the correction's gain
        |r - A P^-1 r| / |r| on this step (the last calib_steps sampling
        steps of the net mode)."""
        return (self.use_net and not self.trained and self._have_corrected()
                and len(self._rows) >= self.keep
                and len(self.samples_f) >= self.train_steps - self.calib_steps)

    def observe(self, ts, sol, corr=None, gain=None):
        """Record an accepted solution, its residual correction P^-1 r and,
        when measured, the correction's gain |r - A P^-1 r| / |r| (below 1:
        x + P^-1 r has the smaller residual)."""
        if gain is not None:
            self._calib.append(float(gain))
        pend = getattr(self, "_pending", None)
        self._push(sol, corr)
        self.times = (self.times + [float(ts.time)])[-self.keep:]
        if self.trained or pend is None:
            return
        ff, frame, x_ext, x_cor = pend
        # targets: the corrected solution x + P^-1 r where given (closer to
        # the exact solution than x) minus each candidate baseline (the
        # baseline is chosen at training time)
        target = sol if corr is None else sol + corr

        def local(base):
            d = target - base
            return torch.cat([torch.einsum("nij,nj->ni", frame, d[ts.idx_x]),
                              d[ts.idx_t][:, None]], dim=1)
        t_ext = local(x_ext)
        t_cor = local(x_cor) if x_cor is not None else t_ext
        if self.max_samples is not None and ff.shape[0] > self.max_samples:
            pick = torch.randperm(ff.shape[0], generator=self._rng)[:self.max_samples]
            pick = pick.to(ff.device)
            ff, t_ext, t_cor = ff[pick], t_ext[pick], t_cor[pick]
        self.samples_f.append((ff, t_ext, t_cor))
        if len(self.samples_f) >= self.train_steps:
            self.train()

    # ---- training -------------------------------------------------------------
    def _split(self, samples):
        nval = max(1, int(round(self.val_frac * len(samples))))
        tr, va = samples[:-nval], samples[-nval:]
        cat = lambda ss, i: torch.cat([s[i] for s in ss]).float()
        return cat(tr, 0), cat(tr, 1), cat(va, 0), cat(va, 1)

    def _train_one(self, samples, n_out, seed):
        Xt, Yt, Xv, Yv = self._split(samples)
        norm = Normalizer()
        norm.fit(Xt)
        # per-output scale: the outputs differ by orders of magnitude
        scale = Yt.abs().amax(0).clamp_min(1e-30)
        net = MLP(Xt.shape[1], n_out, seed=seed)
        net.W[2].zero_()
        self._fit(net, norm(Xt), (Yt / scale).contiguous())
        with torch.no_grad():
            pred = net.forward(norm(Xv)) * scale
            mse_net = float(((pred - Yv) ** 2).mean())
            mse_base = float((Yv**2).mean())
        return net, norm, scale.double(), mse_net, mse_base

    def train(self):
        torch.manual_seed(self.seed)
        if self._calib:
            g = float(np.mean(self._calib))
            self.base_mode = "corrected" if g < 1.0 else "plain"
            self.report["baseline"] = {"mode": self.base_mode, "correction_gain": g,
                                       "steps": len(self._calib)}
        else:
            self.base_mode = "plain"
        col = 2 if self.base_mode == "corrected" else 1
        samples = [(s[0], s[col]) for s in self.samples_f]
        self.net_f, self.nf, self.f_scale, mn, mb = self._train_one(samples, 4, self.seed)
        # first gate: the held-out steps' MSE; the second is the held-out
        # guess residual over the next `shadow_steps` steps (predict)
        self.enabled_f = False
        self.shadow_left = self.shadow_steps if mn < self.gate * mb else 0
        self.report["fiber"] = {"val_mse_net": mn, "val_mse_baseline": mb,
                                "enabled": False, "training_steps": len(self.samples_f)}
        self.trained = True
        self.samples_f = []

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
