"""The reference's learned rank predictors, run on the device.

``RankPredictor`` is the network of tools_qih/ai_rank_predictor.py:14-89
(64 -> 512 -> 512 -> 256 -> 128 -> 1; LayerNorm + ReLU on the first three
hidden layers, ReLU on the fourth, sigmoid output; dropout is the identity at
inference).  Its parameters live in a torch module on the host whose
state-dict keys match the reference's checkpoints (``net.0.weight`` ...), so
``models/rank_predictor.pt`` files load unchanged; the forward pass runs in
the library's dense-stack kernel (``aq_dense_forward``) on inputs computed
from the spectrum on the device (``aq_spectrum_features``), and ``predict``
costs one device->host transfer, of the fraction and the non-zero count.

``DenseStack`` is the shared machinery: it walks an ``nn.Sequential`` of
Linear / LayerNorm / ReLU / GELU / Sigmoid / Dropout modules and keeps an
FP32 copy of its parameters on the device.
"""
from __future__ import annotations

from pathlib import Path
from typing import List

import torch
import torch.nn as nn

from .. import kernels as K
from ..mps_pytorch import hip_device


class DenseStack:
    """Device copy of an ``nn.Sequential`` for ``aq_dense_forward``."""

    def __init__(self, net: nn.Sequential, device):
        self.net = net
        self.device = hip_device(device)
        self.layers: List[dict] = []
        self.refresh()

    def refresh(self) -> None:
        """Upload the module's current parameters."""
        layers = []
        for mod in self.net:
            if isinstance(mod, nn.Linear):
                layers.append({
                    "W": mod.weight.detach().to(self.device, torch.float32).contiguous(),
                    "b": None if mod.bias is None else
                    mod.bias.detach().to(self.device, torch.float32).contiguous(),
                    "act": "none"})
            elif isinstance(mod, nn.LayerNorm):
                L = layers[-1]
                if L["act"] != "none" or "ln_g" in L:
                    raise ValueError("DenseStack: LayerNorm must follow a Linear directly")
                L["ln_g"] = mod.weight.detach().to(self.device, torch.float32).contiguous()
                L["ln_b"] = mod.bias.detach().to(self.device, torch.float32).contiguous()
                L["ln_eps"] = float(mod.eps)
            elif isinstance(mod, (nn.ReLU, nn.GELU, nn.Sigmoid)):
                name = {nn.ReLU: "relu", nn.GELU: "gelu", nn.Sigmoid: "sigmoid"}[type(mod)]
                if isinstance(mod, nn.GELU) and mod.approximate != "none":
                    raise ValueError("DenseStack: only the erf GELU is supported")
                if not layers or layers[-1]["act"] != "none":
                    raise ValueError("DenseStack: activation without a preceding Linear")
                layers[-1]["act"] = name
            elif isinstance(mod, nn.Dropout):
                continue                       # identity at inference
            else:
                raise ValueError(f"DenseStack: unsupported module {type(mod).__name__}")
        self.layers = layers

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return K.dense_forward(self.layers, x.to(self.device))


def _reference_net(input_dim: int, hidden_dim: int) -> nn.Sequential:
    h = hidden_dim
    return nn.Sequential(
        nn.Linear(input_dim, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
        nn.Linear(h, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
        nn.Linear(h, h // 2), nn.LayerNorm(h // 2), nn.ReLU(), nn.Dropout(0.1),
        nn.Linear(h // 2, h // 4), nn.ReLU(),
        nn.Linear(h // 4, 1), nn.Sigmoid())


class RankPredictor(nn.Module):
    """ai_rank_predictor.py:14-89: predicts the fraction of the non-zero
    spectrum to keep from log1p(S[:64] / S[0])."""

    def __init__(self, input_dim: int = 64, hidden_dim: int = 512, device="cuda"):
        super().__init__()
        if input_dim != 64:
            raise ValueError("RankPredictor: the spectrum features have 64 inputs")
        self.input_dim = input_dim
        self.net = _reference_net(input_dim, hidden_dim).eval()
        self._stack = DenseStack(self.net, device)

    def load_state_dict(self, state_dict, strict: bool = True, assign: bool = False):
        res = super().load_state_dict(state_dict, strict=strict, assign=assign)
        self._stack.refresh()
        return res

    def forward(self, sigmas: torch.Tensor, max_ranks):
        """Fractions for a batch of spectra [B, R] (or one spectrum [R]) with
        their non-zero counts; inputs log1p(s[:n] / s[0]), n = min(max_rank, 64)."""
        single = sigmas.dim() == 1
        if single:
            sigmas = sigmas.unsqueeze(0)
            max_ranks = torch.tensor([max_ranks])
        sig = sigmas.to(self._stack.device)
        feats = [K.spectrum_features(sig[b], "log1p64", limit=int(max_ranks[b]))[0]
                 for b in range(sig.shape[0])]
        y = self._stack(torch.cat(feats, 0)).squeeze(-1)
        return y.squeeze(0) if single else y

    def _fraction_and_count(self, sigmas: torch.Tensor):
        sig = sigmas.to(self._stack.device)
        feat, nsv = K.spectrum_features(sig, "log1p64")
        y = self._stack(feat)
        host = torch.cat([y.reshape(-1).double(), nsv.double()]).cpu()
        return float(host[0]), int(host[1])

    def predict_fraction(self, sigmas: torch.Tensor) -> float:
        return self._fraction_and_count(sigmas)[0]

    @torch.no_grad()
    def predict(self, sigmas: torch.Tensor) -> int:
        """ai_rank_predictor.py:71-89: int(fraction * #(S > 1e-10)), clamped to
        [1, #(S > 1e-10)]; 1 for an all-zero spectrum."""
        frac, actual = self._fraction_and_count(sigmas)
        if actual == 0:
            return 1
        return max(1, min(int(frac * actual), actual))


class RankPredictorWrapper:
    """ai_rank_predictor.py:92-115: load / save / predict around RankPredictor.
    Checkpoints are read with ``torch.load(weights_only=True)``."""

    def __init__(self, model_path=None, device="cuda"):
        self.device = device
        self.model = RankPredictor(device=device)
        if model_path and Path(model_path).exists():
            self.load(model_path)

    def load(self, path):
        self.model.load_state_dict(torch.load(path, map_location="cpu", weights_only=True))
        self.model.eval()
        print(f"Loaded rank predictor from {path}")

    def save(self, path):
        torch.save(self.model.state_dict(), path)
        print(f"Saved rank predictor to {path}")

    def predict(self, sigmas) -> int:
        return self.model.predict(sigmas)

    def predict_fraction(self, sigmas) -> float:
        return self.model.predict_fraction(sigmas)

    def train(self, *args, **kwargs):
        raise NotImplementedError(
            "training the reference's 64-512-512-256-128-1 network is not provided by this "
            "build; atlas_q.tools_qih.spectral_predictor.SpectralRankPredictor trains on the "
            "device with the library's MLP kernels")
