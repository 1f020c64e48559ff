"""Training sets for the rank predictors from SVD-event archives.

Produces what the reference's dataset script writes
(scripts/ai_tools/build_ft_dataset.py:24-110): for every ``svd_spectrum``
event, the 135 inputs of the reference's fine-tuned predictor
(finetuned_predictor.py:51-75: 128 log singular values shifted so the largest
is 0, the spectral entropy of the normalised weights, the weight fraction of
the leading 8, 16, 32, 64, 96, 128 values) and the target ``adaptive_keep``
over the non-zero count, capped at 1; ``dataset_real_sv.npz`` (X, y, split)
plus ``file_split.json`` with a 90/10 split by archive (archives sorted,
shuffled by ``numpy.random.default_rng(seed)``).

The events are first gathered into one table (``EventTable``: a [N, 128]
float32 spectrum matrix, kept ranks, archive index per row) and the features
are then computed for all rows at once with array operations.  Unreadable
lines and archives are counted, not fatal.  Offline host tooling: the same
features are computed on the device by ``aq_spectrum_features``
(AQ_FEAT_FT135) for the update path.
"""
from __future__ import annotations

import glob
import gzip
import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import List

import numpy as np

WIDTH = 128
MASS_CUTS = np.array([8, 16, 32, 64, 96, 128])
SPECTRUM_KIND = "svd_spectrum"


@dataclass
class EventTable:
    """Spectrum events of a set of archives, one row per event."""
    spectra: np.ndarray                    # [N, 128] float32, zero-padded
    keep: np.ndarray                       # [N] int64 adaptive_keep (-1: absent)
    archive: np.ndarray                    # [N] int64 index into ``files``
    files: List[str] = field(default_factory=list)
    unreadable: int = 0
    events: list = field(default_factory=list)   # the decoded records, in row order

    def __len__(self):
        return int(self.keep.shape[0])


def _fit_width(values) -> np.ndarray:
    row = np.zeros(WIDTH, np.float32)
    v = np.asarray(values, np.float32).reshape(-1)[:WIDTH]
    row[:v.size] = v
    return row


def read_events(paths) -> EventTable:
    """Decode every ``svd_spectrum`` record of the given archives."""
    files = [str(p) for p in paths]
    rows, keep, arch, events = [], [], [], []
    bad = 0
    for a, name in enumerate(files):
        try:
            with gzip.open(name, "rt", encoding="utf-8") as fh:
                lines = fh.readlines()
        except (OSError, EOFError, UnicodeDecodeError):
            bad += 1
            continue
        for line in lines:
            try:
                rec = json.loads(line)
            except ValueError:
                bad += 1
                continue
            if not isinstance(rec, dict) or rec.get("kind") != SPECTRUM_KIND:
                continue
            try:
                spec = _fit_width(rec.get("S_top128", ()))
            except (TypeError, ValueError):
                bad += 1
                continue
            rows.append(spec)
            k = rec.get("adaptive_keep")
            keep.append(-1 if k is None else int(k))
            arch.append(a)
            events.append(rec)
    spectra = np.stack(rows) if rows else np.zeros((0, WIDTH), np.float32)
    return EventTable(spectra, np.asarray(keep, np.int64), np.asarray(arch, np.int64),
                      files, bad, events)


def ft135_features(spectra: np.ndarray) -> np.ndarray:
    """The fine-tuned predictor's 135 inputs for each row of a [N, 128]
    float32 spectrum matrix (arithmetic in float32, as the predictor's)."""
    s = np.asarray(spectra, np.float32)
    if s.ndim == 1:
        s = s[None, :]
    w = s * s
    mass = w / (w.sum(axis=1, keepdims=True) + np.float32(1e-12))
    with np.errstate(divide="ignore", invalid="ignore"):
        plogp = np.where(mass > 0, mass * np.log(mass + np.float32(1e-12)), np.float32(0))
    entropy = -plogp.sum(axis=1, keepdims=True)
    logs = np.log(s + np.float32(1e-12))
    logs = logs - logs.max(axis=1, keepdims=True)
    cum = np.cumsum(mass, axis=1)[:, MASS_CUTS - 1]
    return np.concatenate([logs, entropy, cum], axis=1).astype(np.float32)


def kept_fraction(spectra: np.ndarray, keep: np.ndarray) -> np.ndarray:
    """Target: adaptive_keep / #(S > 0), at most 1 (the non-zero count when
    the event has no adaptive_keep, i.e. a target of 1)."""
    nz = np.maximum(1, (np.asarray(spectra) > 0).sum(axis=1))
    k = np.where(np.asarray(keep) < 0, nz, keep)
    return np.minimum(1.0, k / nz.astype(np.float64)).astype(np.float32)


def archive_split(files, seed: int = 123):
    """Per-archive split: the sorted archive names shuffled by
    default_rng(seed), the first 90 % for training."""
    order = sorted(set(files))
    np.random.default_rng(seed).shuffle(order)
    n_train = int(0.9 * len(order))
    return set(order[:n_train]), set(order[n_train:])


def build_dataset(paths, seed: int = 123):
    """(X [N, 135], y [N], split [N] int8 (0 train / 1 validation),
    {"train": [...], "val": [...]}, unreadable count)."""
    t = read_events(paths)
    if len(t) == 0:
        raise ValueError("no svd_spectrum events found")
    X = ft135_features(t.spectra)
    y = kept_fraction(t.spectra, t.keep)
    names = [t.files[a] for a in t.archive]
    train, val = archive_split(names, seed)
    split = np.fromiter((0 if n in train else 1 for n in names), np.int8, len(names))
    return X, y, split, {"train": sorted(train), "val": sorted(val)}, t.unreadable


def main(log_glob: str = "runs/svd_logs/svd_events_*.jsonl.gz", out_dir: str = "runs/ftdata"):
    dest = Path(out_dir)
    dest.mkdir(parents=True, exist_ok=True)
    X, y, split, files, bad = build_dataset(sorted(glob.glob(log_glob)))
    np.savez_compressed(dest / "dataset_real_sv.npz", X=X, y=y, split=split)
    (dest / "file_split.json").write_text(json.dumps(files, indent=2))
    print(f"{len(y)} events ({bad} unreadable): {(split == 0).sum()} train / "
          f"{(split == 1).sum()} validation -> {dest / 'dataset_real_sv.npz'}")
    return X, y, split


if __name__ == "__main__":
    import sys
    main(*sys.argv[1:])
