"""HIP-graph capture of the GMRES Arnoldi operator.

One Arnoldi step applies v -> A P^-1 v: the preconditioner (shell / body /
batched fiber LU solves) and the matvec (block GEMVs and the libf3d pair
sums), about a hundred kernel launches at the large workload and the whole
cost of an iteration at small ones. `CapturedOperator` records that sequence
once into a torch CUDA graph (a hipGraph on ROCm) on static input/output
buffers. Each later call copies v in, replays the graph and copies the result
out.

The recorded kernels hold the device addresses of everything the operator
reads: positions, weights, assembled blocks, LU factors and the pair-sum
workspace. A graph is therefore valid only until those tensors are replaced.
The time stepper drops its graphs whenever it re-assembles the system.
Recording runs the operator twice on a side stream first. That warms caches
and grows the shared pair-sum workspace to its final size, so nothing is
allocated inside the recording.
"""
import torch


class CapturedOperator:
    def __init__(self, fn, n, device, dtype=torch.float64, warmup=2, keep=()):
        """fn(x, out) computes out = op(x) for (n,) tensors; it is recorded
        once, on static buffers self.x -> self.y.  `keep` lists tensors the
        recording addresses that nothing else keeps alive (scratch buffers
        of fn): the graph writes them on every replay, so freeing one would
        let the allocator hand its memory to another tensor."""
        self.fn = fn
        self.keep = tuple(keep)
        self.x = torch.zeros(n, dtype=dtype, device=device)
        self.y = torch.zeros(n, dtype=dtype, device=device)
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn(self.x, self.y)
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn(self.x, self.y)
        self.replays = 0

    def __call__(self, v, out=None):
        self.x.copy_(v)
        self.graph.replay()
        self.replays += 1
        if out is None:
            return self.y.clone()
        out.copy_(self.y)
        return out


class ArnoldiGraphs:
    """Lazily recorded A P^-1 operators of one assembled system, one per
    operator precision; `reset()` after every re-assembly."""

    def __init__(self, system):
        self.sys = system
        self._ops = {}
        self.captures = 0

    def reset(self):
        self._ops = {}

    def get(self, precision):
        op = self._ops.get(precision)
        if op is None:
            S = self.sys
            z = torch.empty(S.size, dtype=torch.float64, device=S.dev)

            def fn(x, out):
                S.precond(x, out=z)
                S.matvec(z, out=out, precision=precision)

            op = CapturedOperator(fn, S.size, S.dev, keep=(z,))
            self._ops[precision] = op
            self.captures += 1
        return op
