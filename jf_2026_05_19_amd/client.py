"""Client of the JSONL worker, with the interface of the reference's
``JFEM/python/jfem_client.py:30-206`` (``write_batch_manifest``,
``JFEMWorker``, ``load_summary``), so an optimisation script changes only the
import: the worker process it starts is ``python -m jf_2026_05_19_amd.worker``
(device-resident) instead of the Julia one.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
from pathlib import Path
from typing import Any, Dict, Iterable, Mapping, Optional, Union

from .worker import FAST_DEFAULT_FLAGS

DEFAULT_FAST_FLAGS: Dict[str, str] = dict(FAST_DEFAULT_FLAGS)
REPO_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_batch_manifest(path: Union[str, Path], cases: Iterable[Mapping[str, Any]], output_root: Union[str, Path], *,
                         batch_id: Optional[str] = None, flags: Optional[Mapping[str, Any]] = None,
                         output_options: Optional[Mapping[str, Any]] = None, gc_between: bool = True,
                         stop_on_error: bool = False) -> Path:
    """Writes a batch manifest (the schema of jfem_client.py:30-77) and
    returns its path.  Each case needs at least ``input``."""
    manifest_path = Path(path)
    manifest_path.parent.mkdir(parents=True, exist_ok=True)
    merged = dict(DEFAULT_FAST_FLAGS)
    if flags:
        merged.update({str(k): str(v) for k, v in flags.items()})
    manifest: Dict[str, Any] = {
        "output_root": str(output_root),
        "defaults": {"flags": merged, "output_options": dict(output_options or {"binary": False, "json": True}),
                     "gc_between": gc_between, "stop_on_error": stop_on_error},
        "cases": [dict(c) for c in cases],
    }
    if batch_id is not None:
        manifest["batch_id"] = str(batch_id)
    with manifest_path.open("w", encoding="utf-8") as f:
        json.dump(manifest, f, indent=2)
        f.write("\n")
    return manifest_path


class JFEMWorker:
    """Persistent JSONL worker client (jfem_client.py:110-198): starts the
    worker, checks its ready line, sends one command per line and reads one
    response per line."""

    def __init__(self, *, flags: Optional[Mapping[str, Any]] = None, python: str = sys.executable,
                 stderr: Any = None, env: Optional[Mapping[str, str]] = None) -> None:
        merged = dict(DEFAULT_FAST_FLAGS)
        if flags:
            merged.update({str(k): str(v) for k, v in flags.items()})
        flags_raw = ",".join("%s=%s" % (k, v) for k, v in sorted(merged.items()))
        e = dict(os.environ if env is None else env)
        e["PYTHONPATH"] = REPO_ROOT + (os.pathsep + e["PYTHONPATH"] if e.get("PYTHONPATH") else "")
        self.process = subprocess.Popen([python, "-m", "jf_2026_05_19_amd.worker", flags_raw], cwd=REPO_ROOT,
                                        stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                                        stderr=stderr if stderr is not None else sys.stderr, text=True, bufsize=1,
                                        env=e)
        ready = self._read_response()
        if ready.get("status") != "ready":
            self.close()
            raise RuntimeError("worker did not start cleanly: %s" % ready)
        self.ready = ready

    def _read_response(self) -> Dict[str, Any]:
        line = self.process.stdout.readline()
        if not line:
            raise RuntimeError("worker closed unexpectedly, returncode=%s" % self.process.poll())
        return json.loads(line)

    def request(self, payload: Mapping[str, Any]) -> Dict[str, Any]:
        self.process.stdin.write(json.dumps(dict(payload)) + "\n")
        self.process.stdin.flush()
        response = self._read_response()
        if response.get("status") == "error":
            raise RuntimeError(response.get("error", "worker command failed"))
        return response

    def status(self) -> Dict[str, Any]:
        return self.request({"command": "status"})

    def gc(self) -> Dict[str, Any]:
        return self.request({"command": "gc"})

    def run_batch(self, manifest: Union[str, Path], *, request_id: Optional[str] = None) -> Dict[str, Any]:
        payload: Dict[str, Any] = {"command": "run_batch", "manifest": str(manifest)}
        if request_id is not None:
            payload["id"] = request_id
        return self.request(payload)

    def close(self) -> None:
        if self.process.poll() is not None:
            return
        try:
            self.process.stdin.write(json.dumps({"command": "quit"}) + "\n")
            self.process.stdin.flush()
            self._read_response()
        finally:
            try:
                self.process.wait(timeout=30)
            except subprocess.TimeoutExpired:
                self.process.terminate()
                self.process.wait(timeout=10)

    def __enter__(self) -> "JFEMWorker":
        return self

    def __exit__(self, *exc: Any) -> None:
        self.close()


def load_summary(path: Union[str, Path]) -> Dict[str, Any]:
    with Path(path).open("r", encoding="utf-8") as f:
        return json.load(f)
