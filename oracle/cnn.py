"""CNN emulator forward (CPU restatement of ``emulate/neural_network.py:9-66``).

Architecture (default cfg, conf/processes/iceflow.yaml:61-72): nb_layers x
[Conv2D k x k, nb_out_filter filters, "same" zero padding, bias] ->
LeakyReLU(alpha=0.01) (Dropout(0) is the identity), then a 1x1 Conv2D to
2*Nz channels with no activation.  Keras kernels are HWIO; torch's conv2d is
a cross-correlation like Keras', so HWIO -> OIHW is permute(3, 2, 0, 1).

Weights are a list of (kernel HWIO, bias) pairs.  Pins: torch's own
conv2d/leaky_relu are the library routines the layers reduce to; the
parameter count 140,852 for 16 x 32 filters with 5 inputs and 20 outputs is
checked in the tests.
"""
import math

import torch
import torch.nn.functional as F


def glorot_uniform(shape, generator):
    """Keras GlorotUniform for an HWIO kernel: U(-l, l), l = sqrt(6/(fan_in+fan_out))."""
    kh, kw, ci, co = shape
    fan_in, fan_out = kh * kw * ci, kh * kw * co
    lim = math.sqrt(6.0 / (fan_in + fan_out))
    return (torch.rand(shape, generator=generator, dtype=torch.float64) * 2 - 1) * lim


def init_weights(nb_inputs=5, nb_outputs=20, nb_layers=16, nb_filters=32, ksize=3, seed=0):
    """neural_network.py:23-64 with GlorotUniform kernels and zero biases."""
    g = torch.Generator().manual_seed(seed)
    ws = []
    cin = nb_inputs
    for _ in range(nb_layers):
        ws.append((glorot_uniform((ksize, ksize, cin, nb_filters), g), torch.zeros(nb_filters, dtype=torch.float64)))
        cin = nb_filters
    ws.append((glorot_uniform((1, 1, cin, nb_outputs), g), torch.zeros(nb_outputs, dtype=torch.float64)))
    return ws


def count_params(weights):
    return sum(k.numel() + b.numel() for k, b in weights)


def forward(X, weights, dtype=torch.float32, return_activations=False):
    """X: (N, H, W, Cin) NHWC -> Y (N, H, W, Cout) NHWC."""
    h = torch.as_tensor(X, dtype=dtype).permute(0, 3, 1, 2)
    acts = []
    for li, (k, b) in enumerate(weights):
        kt = torch.as_tensor(k, dtype=dtype).permute(3, 2, 0, 1)
        pad = kt.shape[-1] // 2
        h = F.conv2d(h, kt, torch.as_tensor(b, dtype=dtype), padding=pad)
        if li < len(weights) - 1:
            h = F.leaky_relu(h, 0.01)
            acts.append(h)
    Y = h.permute(0, 2, 3, 1)
    if return_activations:
        return Y, acts
    return Y
