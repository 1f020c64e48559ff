"""The reference's fine-tuned rank predictor, run on the device
(tools_qih/finetuned_predictor.py:8-81).

``FinetunedMLP`` is 135 -> 256 -> 256 -> 1 with ReLU and a sigmoid output,
state-dict compatible with ``models/rank_predictor_ft.pt``.  Its 135 inputs
(128 shifted log singular values, spectral entropy, cumulative mass at 8, 16,
32, 64, 96, 128) are computed from S on the device (``aq_spectrum_features``)
instead of in NumPy on the host.  TorchScript archives are never loaded: for
a ``*_compiled*`` path the wrapper loads the plain state-dict file next to it,
the fallback the reference itself takes when TorchScript loading fails.
"""
from __future__ import annotations

from pathlib import Path

import torch
import torch.nn as nn

from .. import kernels as K
from .ai_rank_predictor import DenseStack


class FinetunedMLP(nn.Module):
    def __init__(self, d_in: int = 135, d_h: int = 256, device="cuda"):
        super().__init__()
        self.net = nn.Sequential(nn.Linear(d_in, d_h), nn.ReLU(), nn.Linear(d_h, d_h), nn.ReLU(),
                                 nn.Linear(d_h, 1), nn.Sigmoid()).eval()
        self._stack = DenseStack(self.net, device)

    def load_state_dict(self, state_dict, strict: bool = True, assign: bool = False):
        res = super().load_state_dict(state_dict, strict=strict, assign=assign)
        self._stack.refresh()
        return res

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return self._stack(x).squeeze(-1)


class FinetunedPredictorWrapper:
    def __init__(self, model_path, device="cuda"):
        self.device = device
        path = str(model_path)
        if "compiled" in path:
            plain = path.replace("_compiled", "")
            if plain == path or not Path(plain).exists():
                raise ValueError(f"{path}: TorchScript archives are not loaded by this build; "
                                 "pass the state-dict checkpoint")
            path = plain
        self.model = FinetunedMLP(device=device)
        self.model.load_state_dict(torch.load(path, map_location="cpu", weights_only=True))
        self.model.eval()

    def _features(self, sigmas: torch.Tensor) -> torch.Tensor:
        return K.spectrum_features(sigmas.to(self.model._stack.device), "ft135")[0]

    def _fraction_and_count(self, sigmas: torch.Tensor):
        feat, nsv = K.spectrum_features(sigmas.to(self.model._stack.device), "ft135")
        y = self.model(feat)
        host = torch.cat([y.reshape(-1).double(), nsv.double()]).cpu()
        return float(host[0]), int(host[1])

    def predict_fraction(self, sigmas: torch.Tensor) -> float:
        return self._fraction_and_count(sigmas)[0]

    @torch.no_grad()
    def predict(self, sigmas: torch.Tensor) -> int:
        """finetuned_predictor.py:77-81: int(fraction * n_sv) clamped to [1, n_sv],
        n_sv = #(S > 1e-10) or 1."""
        frac, n_sv = self._fraction_and_count(sigmas)
        n_sv = n_sv or 1
        return max(1, min(int(frac * n_sv), n_sv))
