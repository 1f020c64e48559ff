"""SVD-event log of the reference (tools_qih/svd_logger.py:14-41).

With ``QIH_SVD_LOG_DIR`` set, ``get_logger()`` returns a logger appending one
JSON object per line to ``svd_events_<YYYYmmdd_HHMMSS>_<pid>.jsonl.gz`` in that
directory, each record carrying ``host``, ``pid`` and ``run_ts`` plus the
caller's fields (``tn_core.mps_apply_2q`` logs ``kind="svd_spectrum"`` events
with ``S_top128`` and ``adaptive_keep``); otherwise a logger that drops
everything.  The file format is the reference's, so logs written here feed
its dataset tools and logs written there feed ``ft_dataset`` here.
"""
from __future__ import annotations

import gzip
import json
import os
import threading
import time
from pathlib import Path


class _NullLogger:
    active = False

    def log(self, **kwargs):
        pass


class _JSONLGZLogger:
    active = True

    def __init__(self, out_dir: str):
        self.dir = Path(out_dir)
        self.dir.mkdir(parents=True, exist_ok=True)
        run = time.strftime("%Y%m%d_%H%M%S")
        pid = os.getpid()
        self.path = self.dir / f"svd_events_{run}_{pid}.jsonl.gz"
        self._fh = gzip.open(self.path, "at", encoding="utf-8")
        self._lock = threading.Lock()
        host = os.uname().nodename if hasattr(os, "uname") else "unknown"
        self.meta = {"host": host, "pid": pid, "run_ts": run}

    def log(self, **kwargs):
        line = json.dumps({**self.meta, **kwargs}, separators=(",", ":"))
        with self._lock:
            self._fh.write(line + "\n")
            self._fh.flush()


def get_logger():
    out = os.getenv("QIH_SVD_LOG_DIR")
    return _JSONLGZLogger(out) if out else _NullLogger()
