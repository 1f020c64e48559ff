"""Host logic of the launch-shape autotuner (no GPU: no timing trials)."""
import json

from atlas_q import autotune
from atlas_q.autotune import Autotuner


def test_shape_heuristic_and_persisted_choice(tmp_path):
    path = tmp_path / "tune.json"
    t = Autotuner(path=path, tune=False)
    # untuned: 64-row tiles and the GEMM heuristic (every tile, up to 4 per CU)
    assert t.shape("gram", {64: 100, 32: 200}) == (64, 100)
    assert t.shape("gram", {64: 5000, 32: 10000}) == (64, 4 * autotune.CUS)
    # a persisted (tile rows, grid) choice is served for its size bucket
    path.write_text(json.dumps({f"gram:shape:{autotune._bucket(5000)}": [32, 1536]}))
    t2 = Autotuner(path=path, tune=False)
    assert t2.shape("gram", {64: 5000, 32: 10000}) == (32, 1536)
    # other buckets keep the heuristic
    assert t2.shape("gram", {64: 20, 32: 40}) == (64, 20)
