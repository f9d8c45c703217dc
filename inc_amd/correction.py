"""Correction network s(u) used by the INC workloads.

A small periodic 1D CNN: ``depth`` circular Conv1d layers of ``width``
channels and kernel ``k`` with GELU in between, He-normal weights and zero
biases (the correction networks of the reference take (u, L) and return
[B, N], models/models_1d/resnet.py:22-52; this one keeps that call shape).
It runs in PyTorch (MIOpen convolutions); the solver step is the HIP part.
"""
import torch
from torch import nn


class CorrectionCNN(nn.Module):
    def __init__(self, depth=4, width=32, k=5, generator=None):
        super().__init__()
        layers = []
        cin = 1
        for i in range(depth):
            cout = 1 if i == depth - 1 else width
            conv = nn.Conv1d(cin, cout, k, padding=k // 2, padding_mode="circular")
            with torch.no_grad():
                std = (2.0 / (cin * k)) ** 0.5
                conv.weight.copy_(torch.randn(conv.weight.shape, generator=generator) * std)
                conv.bias.zero_()
            layers.append(conv)
            cin = cout
        self.convs = nn.ModuleList(layers)
        self.act = nn.GELU()

    def forward(self, u, L=None):
        x = u.unsqueeze(1)
        for i, conv in enumerate(self.convs):
            x = conv(x)
            if i < len(self.convs) - 1:
                x = self.act(x)
        return x.squeeze(1)
