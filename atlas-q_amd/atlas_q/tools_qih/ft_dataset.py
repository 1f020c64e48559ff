"""Training-set builder for the rank predictors from SVD-event logs, the
reference's scripts/ai_tools/build_ft_dataset.py:24-110 as a library module.

* ``safe_load_events(paths)``: ``kind == "svd_spectrum"`` records of
  ``svd_events_*.jsonl.gz`` files (the format of ``svd_logger``), skipping
  unreadable lines / files; returns ((path, event) rows, bad count).
* ``sv_to_features_and_target(ev)``: the reference's 135 features of the
  event's ``S_top128`` (as ``FinetunedPredictorWrapper`` computes them) and the
  target ``adaptive_keep / #nonzero`` clipped to 1.
* ``build_dataset(paths, seed=123)``: stacked X [N, 135], y [N] and a
  file-level 90/10 split (files shuffled by ``default_rng(seed)``).
* ``main(log_glob, out_dir)``: writes ``dataset_real_sv.npz`` and
  ``file_split.json`` as the reference's script does.

Offline host tooling (NumPy): it reads logs, it is not on the update path.
"""
from __future__ import annotations

import glob
import gzip
import json
from pathlib import Path

import numpy as np

CUTS = (8, 16, 32, 64, 96, 128)


def safe_load_events(paths):
    rows, bad = [], 0
    for p in paths:
        try:
            with gzip.open(p, "rt") as f:
                for line in f:
                    try:
                        ev = json.loads(line)
                    except Exception:  # noqa: BLE001
                        bad += 1
                        continue
                    if ev.get("kind") == "svd_spectrum":
                        rows.append((p, ev))
        except Exception:  # noqa: BLE001
            bad += 1
    return rows, bad


def sv_to_features_and_target(ev):
    sv = np.array(ev.get("S_top128", []), dtype=np.float32)
    sv = np.pad(sv, (0, 128 - sv.size)) if sv.size < 128 else sv[:128]
    n_sv = max(1, int((sv > 0).sum()))
    sv2 = sv ** 2
    mass = sv2 / (sv2.sum() + 1e-12)
    nz = mass > 0
    H = float(-(mass[nz] * np.log(mass[nz] + 1e-12)).sum())
    logs = np.log(sv + 1e-12)
    logs -= logs.max()
    cm = [float(mass[:k].sum()) for k in CUTS]
    x = np.concatenate([logs, np.array([H], np.float32), np.array(cm, np.float32)])
    frac = min(1.0, int(ev.get("adaptive_keep", n_sv)) / float(n_sv))
    return x.astype(np.float32), np.float32(frac)


def build_dataset(paths, seed: int = 123):
    rows, bad = safe_load_events(paths)
    X, y, files = [], [], []
    for p, ev in rows:
        try:
            xi, yi = sv_to_features_and_target(ev)
        except Exception:  # noqa: BLE001
            continue
        X.append(xi)
        y.append(yi)
        files.append(str(p))
    if not X:
        raise ValueError("no svd_spectrum events found")
    uniq = sorted(set(files))
    np.random.default_rng(seed).shuffle(uniq)
    cut = int(0.9 * len(uniq))
    train = set(uniq[:cut])
    split = np.array([0 if f in train else 1 for f in files], dtype=np.int8)
    return np.stack(X), np.stack(y), split, {"train": sorted(train),
                                             "val": sorted(set(uniq[cut:]))}, bad


def main(log_glob: str = "runs/svd_logs/svd_events_*.jsonl.gz", out_dir: str = "runs/ftdata"):
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    X, y, split, files, bad = build_dataset(sorted(glob.glob(log_glob)))
    np.savez_compressed(out / "dataset_real_sv.npz", X=X, y=y, split=split)
    (out / "file_split.json").write_text(json.dumps(files, indent=2))
    print(f"{len(y)} events ({bad} unreadable), train {(split == 0).sum()}, "
          f"val {(split == 1).sum()} -> {out / 'dataset_real_sv.npz'}")
    return X, y, split


if __name__ == "__main__":
    import sys
    main(*sys.argv[1:])
