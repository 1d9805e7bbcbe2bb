"""libqgcnn's BF16 MFMA kernels against a torch-CPU FP32 reference of the same
operations on the same BF16-rounded tensors (periodic 5x5 cross-correlation,
NHWC activations, weights [cout][25][cin])."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cnn(gpu):
    import torch  # noqa: F401
    from pyqg_amd.closure import cnn
    return cnn


def ref_conv(x_nhwc, w, bias=None):
    """torch CPU FP32: x (B,H,W,Ci), w (Co,25,Ci) -> (B,H,W,Co)"""
    import torch
    import torch.nn.functional as F
    co, _, ci = w.shape
    wt = w.reshape(co, 5, 5, ci).permute(0, 3, 1, 2)
    x = x_nhwc.permute(0, 3, 1, 2)
    y = F.conv2d(F.pad(x, (2, 2, 2, 2), mode="circular"), wt, bias)
    return y.permute(0, 2, 3, 1)


def rnd_bf16(shape, scale, gen):
    import torch
    return (torch.randn(shape, generator=gen) * scale).to(torch.bfloat16).float()


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("cinp,coutp,relu", [(16, 32, True), (32, 128, True), (128, 64, False),
                                             (64, 16, False)])
def test_conv_forward(cnn, cinp, coutp, relu):
    import torch
    g = torch.Generator().manual_seed(cinp + coutp)
    B, H, W = 2, 8, 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    b = torch.randn(coutp, generator=g) * 0.1
    y_bf, y32 = cnn.conv_fwd(x.to(torch.bfloat16).cuda(), w.to(torch.bfloat16).cuda(), b.cuda(),
                             coutp, relu=relu, out_f32=True)
    ref = ref_conv(x, w, b)
    if relu:
        ref = torch.relu(ref)
    assert rel(y32.cpu(), ref) < 1e-5
    assert rel(y_bf.float().cpu(), ref) < 1e-2  # BF16 output rounding
    # the BF16 output is the FP32 result rounded to nearest even
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("cinp,coutp,H,relu,masked", [(32, 32, 8, True, False),
                                                     (64, 32, 12, True, False),
                                                     (128, 64, 64, False, False),
                                                     (32, 64, 6, True, True),
                                                     (96, 32, 16, False, True)])
def test_conv_forward_narrow_layers(cnn, cinp, coutp, H, relu, masked):
    """Layers with 32 or 64 output channels against the float reference, at row
    counts that fill partial blocks, with and without the ReLU mask of the
    backward pass."""
    import torch
    g = torch.Generator().manual_seed(cinp * 7 + coutp + H)
    B, W = 2, 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    b = torch.randn(coutp, generator=g) * 0.1
    mk = torch.relu(rnd_bf16((B, H, W, coutp), 1.0, g)) if masked else None
    y_bf, y32 = cnn.conv_fwd(x.to(torch.bfloat16).cuda(), w.to(torch.bfloat16).cuda(), b.cuda(),
                             coutp, relu=relu, out_f32=True,
                             mask=None if mk is None else mk.to(torch.bfloat16).cuda())
    ref = ref_conv(x, w, b)
    if relu:
        ref = torch.relu(ref)
    if masked:
        ref = ref * (mk > 0)
    assert rel(y32.cpu(), ref) < 1e-5
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("cinp,coutp,B", [(64, 64, 36), (32, 128, 20), (64, 32, 72),
                                           (32, 32, 70), (96, 32, 66)])
def test_conv_forward_blocks_with_several_items(cnn, cinp, coutp, B):
    """Batches whose (sample, 8-row group, channel group) items outnumber the
    resident blocks of the overlapped-staging kernel (one per CU for 64-channel
    groups, two for 32-channel ones), so that each takes several items in turn
    (accumulators cleared between items, staging buffers reused across items
    and input chunks)."""
    import torch
    g = torch.Generator().manual_seed(cinp + 3 * coutp + B)
    H, W = 64, 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    b = torch.randn(coutp, generator=g) * 0.1
    y_bf, y32 = cnn.conv_fwd(x.to(torch.bfloat16).cuda(), w.to(torch.bfloat16).cuda(), b.cuda(),
                             coutp, relu=True, out_f32=True)
    ref = torch.relu(ref_conv(x, w, b))
    assert rel(y32.cpu(), ref) < 1e-5
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("cinp,coutp,B", [(64, 128, 20), (32, 32, 70)])
def test_masked_convolution_several_items_per_block(cnn, cinp, coutp, B):
    """The backward-data form (no ReLU, the previous layer's ReLU mask, no
    bias) at batch sizes where the overlapped-staging kernel's blocks take
    several items each."""
    import torch
    g = torch.Generator().manual_seed(cinp + coutp + B)
    H = W = 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    mk = torch.relu(rnd_bf16((B, H, W, coutp), 1.0, g))
    y_bf, y32 = cnn.conv_fwd(x.to(torch.bfloat16).cuda(), w.to(torch.bfloat16).cuda(), None,
                             coutp, relu=False, out_f32=True, mask=mk.to(torch.bfloat16).cuda())
    ref = ref_conv(x, w) * (mk > 0)
    assert rel(y32.cpu(), ref) < 1e-5
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("B", [48, 80])
def test_first_layer_kernel_several_items_per_block(cnn, B):
    """The first layer at the closure's shape (6 -> 128 channels, 64 x 64) with
    more (sample, row group, channel group) items than resident blocks: each
    block keeps one channel group's weights and takes several items."""
    import torch
    g = torch.Generator().manual_seed(B)
    H = W = 64
    cin, coutp = 6, 128
    x = torch.zeros((B, H, W, 16))
    x[..., :cin] = rnd_bf16((B, H, W, cin), 1.0, g)
    w = torch.zeros((coutp, 25, 16))
    w[..., :cin] = rnd_bf16((coutp, 25, cin), 0.1, g)
    b = torch.randn(coutp, generator=g) * 0.1
    w8 = cnn.pack_first(w.to(torch.bfloat16).cuda())
    y_bf, y32 = cnn.conv_first(x.to(torch.bfloat16).cuda(), w8, b.cuda(), coutp, relu=True,
                               out_f32=True)
    ref = torch.relu(ref_conv(x, w, b))
    assert rel(y32.cpu(), ref) < 1e-5
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("cin,coutp,H,relu", [(6, 128, 8, True), (8, 64, 16, True),
                                              (3, 32, 12, False), (6, 16, 5, True)])
def test_first_layer_kernel(cnn, cin, coutp, H, relu):
    """The first-layer kernel (25 taps x 8 channels packed into K) against the
    float reference, the input stored with 16 channels of which cin are real
    and the rest zero (the network input's layout)."""
    import torch
    g = torch.Generator().manual_seed(cin * 100 + coutp + H)
    B, W = 2, 64
    x = torch.zeros((B, H, W, 16))
    x[..., :cin] = rnd_bf16((B, H, W, cin), 1.0, g)
    w = torch.zeros((coutp, 25, 16))
    w[..., :cin] = rnd_bf16((coutp, 25, cin), 0.1, g)
    b = torch.randn(coutp, generator=g) * 0.1
    w8 = cnn.pack_first(w.to(torch.bfloat16).cuda())
    y_bf, y32 = cnn.conv_first(x.to(torch.bfloat16).cuda(), w8, b.cuda(), coutp, relu=relu,
                               out_f32=True)
    ref = ref_conv(x, w, b)
    if relu:
        ref = torch.relu(ref)
    assert rel(y32.cpu(), ref) < 1e-5
    assert rel(y_bf.float().cpu(), ref) < 1e-2
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))


@pytest.mark.parametrize("cinp,cout,H,coutp", [(32, 2, 8, 16), (64, 3, 12, 16), (32, 1, 5, 16),
                                               (128, 2, 64, 16), (32, 2, 8, 4), (64, 3, 16, 4)])
def test_last_layer_kernel(cnn, cinp, cout, H, coutp):
    """The last-layer kernel (horizontal taps as matrix rows, the horizontal sum
    through LDS) against the float reference; the padded output channels carry
    their bias.  coutp = 4: the narrow output of the closed step."""
    import torch
    g = torch.Generator().manual_seed(cinp + 10 * cout + H)
    B, W = 3, 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = torch.zeros((16, 25, cinp))     # the layer's weights: 16 padded output channels
    w[:cout] = rnd_bf16((cout, 25, cinp), 0.1, g)
    b = torch.zeros(16)
    b[:cout] = torch.randn(cout, generator=g) * 0.1
    wl = cnn.pack_last(w.to(torch.bfloat16).cuda(), cout)
    y_bf, y32 = cnn.conv_last(x.to(torch.bfloat16).cuda(), wl, b.cuda(), coutp, cout,
                              out_bf16=True, out_f32=True)
    ref = ref_conv(x, w, b)[..., :coutp]
    assert rel(y32.cpu(), ref) < 1e-5
    assert rel(y_bf.float().cpu(), ref) < 1e-2
    assert torch.equal(y_bf.cpu(), y32.cpu().to(torch.bfloat16))
    assert not y32[..., cout:].any()


@pytest.mark.parametrize("cinp,B", [(32, 40), (64, 36)])
def test_last_layer_kernel_several_items_per_block(cnn, cinp, B):
    """The last layer at the closure's shape (2 outputs, the 4-channel FP32
    output of the closed step, 64 x 64) with more (sample, row group) items
    than resident blocks, so that each block takes several in turn."""
    import torch
    g = torch.Generator().manual_seed(cinp + B)
    H = W = 64
    cout, coutp = 2, 4
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = torch.zeros((16, 25, cinp))
    w[:cout] = rnd_bf16((cout, 25, cinp), 0.1, g)
    b = torch.zeros(16)
    b[:cout] = torch.randn(cout, generator=g) * 0.1
    wl = cnn.pack_last(w.to(torch.bfloat16).cuda(), cout)
    _, y32 = cnn.conv_last(x.to(torch.bfloat16).cuda(), wl, b.cuda(), coutp, cout,
                           out_bf16=False, out_f32=True)
    ref = ref_conv(x, w, b)[..., :coutp]
    assert rel(y32.cpu(), ref) < 1e-5
    assert not y32[..., cout:].any()


def test_backward_data_with_relu_mask(cnn):
    import torch
    g = torch.Generator().manual_seed(1)
    B, H, W, cinp, coutp = 2, 8, 64, 32, 64
    x = torch.relu(rnd_bf16((B, H, W, cinp), 1.0, g))   # previous layer's ReLU output
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    dy = rnd_bf16((B, H, W, coutp), 1.0, g)
    wt = torch.empty((cinp, 25, coutp), dtype=torch.bfloat16, device="cuda")
    cnn.flip_transpose(w.to(torch.bfloat16).cuda(), wt)
    _, dx32 = cnn.conv_fwd(dy.to(torch.bfloat16).cuda(), wt, None, cinp, relu=False,
                           mask=x.to(torch.bfloat16).cuda(), out_bf16=False, out_f32=True)
    xr = x.clone().requires_grad_(True)
    (ref_conv(xr, w) * dy).sum().backward()
    ref = xr.grad * (x > 0)
    assert rel(dx32.cpu(), ref) < 1e-5


@pytest.mark.parametrize("cinp,coutp", [(16, 128), (128, 64), (32, 32), (32, 16)])
def test_weight_and_bias_gradient(cnn, cinp, coutp):
    import torch
    g = torch.Generator().manual_seed(7 + cinp)
    B, H, W = 2, 16, 64
    x = rnd_bf16((B, H, W, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    dy = rnd_bf16((B, H, W, coutp), 1.0, g)
    dw = torch.zeros((coutp, 25, cinp), device="cuda")
    db = torch.zeros(coutp, device="cuda")
    cnn.conv_wgrad(dy.to(torch.bfloat16).cuda(), x.to(torch.bfloat16).cuda(), dw)
    cnn.bias_grad(dy.to(torch.bfloat16).cuda(), db)
    wr = w.clone().requires_grad_(True)
    br = torch.zeros(coutp, requires_grad=True)
    (ref_conv(x, wr, br) * dy).sum().backward()
    assert rel(dw.cpu(), wr.grad) < 1e-5
    assert rel(db.cpu(), br.grad) < 1e-5


@pytest.mark.parametrize("B,H,cinp,coutp", [(3, 10, 16, 128), (2, 12, 128, 64),
                                            (1, 64, 32, 32), (2, 8, 32, 16), (2, 16, 64, 16),
                                            (1, 8, 32, 48), (1, 8, 48, 32), (2, 8, 96, 80)])
def test_blocked_weight_gradient_matches_atomic(cnn, B, H, cinp, coutp):
    """The matrix-core weight gradient (workspace, K split over rows) equals the
    atomic kernel and the float reference on odd batch and row counts."""
    import torch
    g = torch.Generator().manual_seed(11 * cinp + coutp + H)
    x = rnd_bf16((B, H, 64, cinp), 1.0, g)
    w = rnd_bf16((coutp, 25, cinp), 0.1, g)
    dy = rnd_bf16((B, H, 64, coutp), 1.0, g)
    xb, dyb = x.to(torch.bfloat16).cuda(), dy.to(torch.bfloat16).cuda()
    d1 = torch.zeros((coutp, 25, cinp), device="cuda")
    d2 = torch.zeros((coutp, 25, cinp), device="cuda")
    cnn.conv_wgrad(dyb, xb, d1, blocked=True)
    cnn.conv_wgrad(dyb, xb, d2, blocked=False)
    assert rel(d1, d2) < 1e-5
    # accumulates into dw
    cnn.conv_wgrad(dyb, xb, d1, blocked=True)
    assert rel(d1, 2 * d2) < 1e-5
    wr = w.clone().requires_grad_(True)
    (ref_conv(x, wr) * dy).sum().backward()
    assert rel(d2.cpu(), wr.grad) < 1e-5


@pytest.mark.parametrize("coutp,nch", [(16, 2), (4, 2), (24, 10)])
def test_mse_loss_and_gradient(cnn, coutp, nch):
    """Loss and gradient over the first nch of coutp channels (16-byte path
    for coutp a multiple of 8, with real channels across a piece boundary;
    per-channel path for 4)."""
    import torch
    g = torch.Generator().manual_seed(3 + coutp)
    pred = torch.randn((2, 8, 64, coutp), generator=g)
    target = torch.randn((2, 8, 64, nch), generator=g)
    grad = torch.empty(pred.shape, dtype=torch.bfloat16, device="cuda")
    loss = cnn.mse(pred.cuda(), target.cuda(), grad)
    ref = ((pred[..., :nch] - target) ** 2).mean()
    assert abs(float(loss) - float(ref)) < 1e-5 * float(ref)
    gref = torch.zeros_like(pred)
    gref[..., :nch] = 2 * (pred[..., :nch] - target) / target.numel()
    assert rel(grad.float().cpu(), gref) < 1e-2
    if nch < coutp:
        assert grad[..., nch:].float().abs().max().item() == 0.0


def test_network_gradients_match_torch(cnn):
    """Whole small network: loss and every layer's gradient against torch
    autograd, with activations and activation-gradients rounded to BF16 at
    the same points as the kernels store them."""
    import torch
    net = cnn.CNN(channels=(6, 32, 16, 2), seed=4)
    g = torch.Generator().manual_seed(11)
    x = torch.randn((2, 8, 64, 6), generator=g)
    target = torch.randn((2, 8, 64, 2), generator=g) * 0.1
    xp = net.pad_input(x.cuda())
    out, acts = net.forward(xp, keep=True)
    grad = torch.empty(out.shape, dtype=torch.bfloat16, device="cuda")
    loss = cnn.mse(out, target.cuda(), grad)
    dws, dbs = net.backward(acts, grad)

    class RoundGrad(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            return a.to(torch.bfloat16).float()

        @staticmethod
        def backward(ctx, ga):
            return ga.to(torch.bfloat16).float()

    class RoundBack(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            return a.clone()

        @staticmethod
        def backward(ctx, ga):
            return ga.to(torch.bfloat16).float()

    ws = [w.cpu()[: net.channels[l + 1], :, : net.channels[l]].clone().requires_grad_(True)
          for l, w in enumerate(net.wb)]
    ws_f = [w.float().detach().requires_grad_(True) for w in ws]
    bs = [b.cpu()[: net.channels[l + 1]].clone().requires_grad_(True)
          for l, b in enumerate(net.b32)]
    h = x.to(torch.bfloat16).float()
    for l in range(net.nlayers):
        z = ref_conv(h, ws_f[l], bs[l])
        if l < net.nlayers - 1:
            h = RoundGrad.apply(torch.relu(z))
        else:
            h = RoundBack.apply(z)
    ref_loss = ((h - target) ** 2).mean()
    ref_loss.backward()
    assert abs(float(loss) - float(ref_loss)) < 2e-3 * float(ref_loss)
    for l in range(net.nlayers):
        co, ci = net.channels[l + 1], net.channels[l]
        got = dws[l].cpu()[:co, :, :ci]
        assert rel(got, ws_f[l].grad) < 3e-2, l
        assert rel(dbs[l].cpu()[:co], bs[l].grad) < 3e-2, l


def test_training_reduces_loss(cnn):
    import torch
    net = cnn.CNN(channels=(6, 32, 32, 2), seed=5)
    g = torch.Generator().manual_seed(2)
    x = torch.randn((4, 16, 64, 6), generator=g)
    # a target the network can represent: a fixed smoothing of two input channels
    target = (x[..., :2] + torch.roll(x[..., :2], 1, dims=2)) * 0.5
    xp = net.pad_input(x.cuda())
    t = target.cuda()
    first = float(net.train_step(xp, t, lr=3e-3))
    for _ in range(150):
        last = float(net.train_step(xp, t, lr=3e-3))
    assert last < 0.2 * first


def test_training_is_bit_reproducible(cnn):
    """Two networks from the same seed, trained on the same batches, end with
    identical weights: every gradient is summed in a fixed order."""
    import torch
    g = torch.Generator().manual_seed(3)
    x = torch.randn((8, 32, 64, 6), generator=g).cuda()
    t = torch.randn((8, 32, 64, 2), generator=g).cuda()
    nets = [cnn.CNN(channels=(6, 32, 32, 2), seed=9) for _ in range(2)]
    for net in nets:
        xp = net.pad_input(x)
        for _ in range(5):
            net.train_step(xp, t, lr=1e-3, clip=10.0)
    torch.cuda.synchronize()
    for l in range(nets[0].nlayers):
        assert torch.equal(nets[0].w32[l], nets[1].w32[l]), l
        assert torch.equal(nets[0].b32[l], nets[1].b32[l]), l
