"""SVD-event log in the reference's on-disk format (tools_qih/svd_logger.py:14-41).

Format (the interface, kept so logs interoperate in both directions): with the
environment variable ``QIH_SVD_LOG_DIR`` naming a directory, events go to
``<dir>/svd_events_<YYYYmmdd_HHMMSS>_<pid>.jsonl.gz``, one compact JSON object
per line, each carrying ``host``, ``pid`` and ``run_ts`` (the file's time
stamp) before the caller's fields.  ``tn_core.mps_apply_2q`` logs
``kind="svd_spectrum"`` events (``S_top128``, ``adaptive_keep``, ...).

``EventLog`` is one object whatever the setting: without a directory it is
inactive and ``log`` returns at once.  The file is created on the first event
(an idle process leaves no empty archive), every line is flushed so a crashed
run keeps what it logged, and writers on several threads are serialised.
"""
from __future__ import annotations

import gzip
import json
import os
import socket
import threading
import time
from pathlib import Path
from typing import Optional

ENV_DIR = "QIH_SVD_LOG_DIR"


class EventLog:
    """Appends JSON-line events to a gzip archive (inactive without a directory)."""

    def __init__(self, directory: Optional[str]):
        self.directory = Path(directory) if directory else None
        stamp = time.strftime("%Y%m%d_%H%M%S")
        pid = os.getpid()
        self.header = {"host": socket.gethostname(), "pid": pid, "run_ts": stamp}
        self.path = (self.directory / f"svd_events_{stamp}_{pid}.jsonl.gz"
                     if self.directory is not None else None)
        self._stream = None
        self._guard = threading.Lock()

    @property
    def active(self) -> bool:
        return self.path is not None

    def _open(self):
        self.directory.mkdir(parents=True, exist_ok=True)
        return gzip.open(self.path, mode="at", encoding="utf-8")

    def log(self, **fields) -> None:
        if self.path is None:
            return
        record = dict(self.header)
        record.update(fields)
        text = json.dumps(record, separators=(",", ":")) + "\n"
        with self._guard:
            if self._stream is None:
                self._stream = self._open()
            self._stream.write(text)
            self._stream.flush()

    def close(self) -> None:
        """Finish the archive (further events reopen it in append mode)."""
        with self._guard:
            if self._stream is not None:
                self._stream.close()
                self._stream = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def get_logger() -> EventLog:
    """The process's event log as configured by ``QIH_SVD_LOG_DIR``."""
    return EventLog(os.environ.get(ENV_DIR) or None)
