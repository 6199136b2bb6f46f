"""FP64 CPU reference of the denoiser UNet. TEST INFRASTRUCTURE ONLY.

Restates the network of ml/train.py:37-75 (ConvBlock = two 3x3 convs with padding 1 + bias +
activation; MaxPool2d(2); 2x upsample; channel concat of [upsampled, skip]; 1x1 tail added
to the input) in the configuration the project trains: ReLU, bilinear upsampling with
align_corners=False, encoder widths 32/64/128/256, 512-wide bottleneck. Built from the
library routines torch.nn.functional.conv2d / max_pool2d / interpolate / relu / cat in
float64 on the CPU, so it checks the hand-written kernels against independent code.

Weights are the project's layout [cout, 9, cin] (tap = 3*ky + kx); the first convolution's
input channels beyond 3 are zero padding and are dropped here.
"""
import torch
import torch.nn.functional as F

ENC = (32, 64, 128, 256)


def torch_weight(w):
    """[cout, 9, cin] -> conv2d weight [cout, cin, 3, 3]."""
    cout, taps, cin = w.shape
    k = int(round(taps ** 0.5))
    return w.reshape(cout, k, k, cin).permute(0, 3, 1, 2).contiguous()


def conv(x, w, b, relu=True):
    y = F.conv2d(x, torch_weight(w), b, padding=1)
    return F.relu(y) if relu else y


def up(x):
    return F.interpolate(x, scale_factor=2, mode="bilinear", align_corners=False)


def bf16_round(h):
    """Round the value to BF16 and pass the gradient straight through."""
    return h + (h.to(torch.bfloat16).to(torch.float64) - h).detach()


def forward(params, x_nhwc, taps=None, bf16_activations=False):
    """params: dict name -> tensor (weights/biases as in lgi.unet.UNet.state_dict()).
    x_nhwc: [N, H, W, 3]. Returns [N, H, W, 3]; per-level activations (NHWC) into taps.
    bf16_activations rounds every stored activation to BF16 (straight-through gradient), as the
    GPU network stores them, so that ReLU masks and pooling choices follow the same values."""
    # .to() keeps the autograd graph, so float64 CPU parameters that require grad get gradients
    P = {k: v.to("cpu", torch.float64) for k, v in params.items()}
    x = x_nhwc.to("cpu", torch.float64).permute(0, 3, 1, 2)
    rnd = bf16_round if bf16_activations else (lambda h: h)

    def block(prefix, h):
        w0 = P[f"weights.{prefix}a"]
        if prefix == "enc0":
            w0 = w0[:, :, :3]
        h = rnd(conv(h, w0, P[f"biases.{prefix}a"]))
        return rnd(conv(h, P[f"weights.{prefix}b"], P[f"biases.{prefix}b"]))

    def tap(name, h):
        if taps is not None:
            taps[name] = h.permute(0, 2, 3, 1)

    h = rnd(x)
    skips = []
    for i in range(len(ENC)):
        h = block(f"enc{i}", h)
        skips.append(h)
        tap(f"enc{i}", h)
        h = F.max_pool2d(h, 2)
    h = block("mid_", h)
    tap("mid", h)
    for i in reversed(range(len(ENC))):
        h = block(f"dec{i}", torch.cat([rnd(up(h)), skips[i]], 1))
        tap(f"dec{i}", h)
    out = x + F.conv2d(h, P["tail_w"].reshape(3, 32, 1, 1), P["tail_b"])
    return out.permute(0, 2, 3, 1)
