"""MFMA BF16 GEMM and the MLP forward/backward against FP64 references."""
import numpy as np
import pytest
import torch

from conftest import rel2
from fiber3d_amd import learned as LG
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _bf16(a):
    return torch.from_numpy(np.asarray(a, np.float32)).bfloat16().float().double().numpy()


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 70, 37), (1000, 256, 256), (3, 5, 7)])
def test_gemm_bf16_matches_fp64_of_rounded_inputs(M, N, K):
    print(f"case gemm {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    ref = _bf16(A) @ _bf16(B)
    C = torch.empty((M, N), device="cuda")
    LG.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C)
    torch.cuda.synchronize()
    assert rel2(C.cpu().numpy(), ref) < 1e-6  # FP32 accumulation of exact BF16 products
    # transposed operands and split-K atomic accumulation
    At = torch.from_numpy(A.T.copy()).cuda()
    C2 = torch.zeros((M, N), device="cuda")
    LG.gemm(At, torch.from_numpy(B).cuda(), C2, LG.EPI_ATOMIC, transA=True, ksplit=3)
    torch.cuda.synchronize()
    assert rel2(C2.cpu().numpy(), ref) < 1e-6


def test_mlp_forward_backward_against_oracle():
    print("case mlp fwd/bwd", flush=True)
    torch.manual_seed(0)
    net = LG.MLP(8, 3, width=256, seed=3)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(2000, 8)).astype(np.float32)
    T = rng.normal(size=(2000, 3)).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    Y = net.forward(Xd, keep=True)
    dY = (2.0 / Y.numel()) * (Y - torch.from_numpy(T).cuda())
    g = net.backward(dY.contiguous())
    torch.cuda.synchronize()
    params = [p.double().cpu().numpy() for p in net.params]
    Yr, _ = OM.forward(params, X.astype(np.float64))
    gr = OM.grads(params, X.astype(np.float64), T.astype(np.float64))
    # BF16 operands (8-bit mantissa): a few 1e-3 relative on products of
    # width-256 layers
    assert rel2(Y.cpu().numpy(), Yr) < 2e-2
    for a, b in zip(g, gr):
        assert rel2(a.cpu().numpy(), b) < 5e-2


def test_mlp_training_reduces_loss():
    print("case mlp train", flush=True)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(4096, 6)).astype(np.float32)
    T = (np.sin(X[:, :1]) + 0.5 * X[:, 1:2] * X[:, 2:3]).astype(np.float32)
    net = LG.MLP(6, 1, seed=0)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    first = net.train_step(Xd, Td, lr=1e-3)
    for _ in range(300):
        last = net.train_step(Xd, Td, lr=1e-3)
    assert last < 0.1 * first


@pytest.mark.parametrize("N", [128, 256])
@pytest.mark.parametrize("M,K", [(4099, 256), (1000, 48), (257, 160), (4096, 256), (32832, 64)])
@pytest.mark.parametrize("epi", ["store", "bias_relu", "bias", "relu_grad"])
def test_per_node_layer_kernels_agree(N, M, K, epi):
    """The per-node layer kernels (resident-weight, streaming, tiled; chosen
    per N through f3d_gemm_set_stream) on one problem with each epilogue,
    partial last row blocks included, and whole row blocks only (M = 4096:
    fewer row blocks than workgroups; M = 32832: 1026 or 513 row blocks, an
    odd count for some workgroups): each meets the FP64 product of the
    BF16-rounded inputs, and the Z output of the bias + ReLU epilogue holds
    the pre-activation."""
    from fiber3d_amd import _lib as L
    print(f"case layer kernels {M}x{N}x{K} {epi}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    zin = rng.normal(size=(M, N)).astype(np.float32)
    pre = _bf16(A) @ _bf16(B)
    code = {"store": LG.EPI_STORE, "bias_relu": LG.EPI_BIAS_RELU, "bias": LG.EPI_BIAS,
            "relu_grad": LG.EPI_RELU_GRAD}[epi]
    if epi == "store":
        ref = pre
    elif epi == "bias":
        ref = pre + bias
    elif epi == "bias_relu":
        ref = np.maximum(pre + bias, 0.0)
    else:
        ref = np.where(zin > 0, pre, 0.0)
    lib = L.load()
    try:
        for mode in (2, 1, 0):
            L.check(lib.f3d_gemm_set_stream(N, mode), "f3d_gemm_set_stream")
            C = torch.full((M, N), 7.0, device="cuda")
            Z = torch.full((M, N), 7.0, device="cuda") if epi == "bias_relu" else None
            LG.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C, code,
                    bias=torch.from_numpy(bias).cuda() if epi in ("bias", "bias_relu") else None,
                    Z=Z, Zin=torch.from_numpy(zin).cuda() if epi == "relu_grad" else None)
            torch.cuda.synchronize()
            assert rel2(C.cpu().numpy(), ref) < 1e-6, mode
            if Z is not None:
                assert rel2(Z.cpu().numpy(), pre + bias) < 1e-6, mode
    finally:
        from fiber3d_amd import autotune
        autotune.apply()
