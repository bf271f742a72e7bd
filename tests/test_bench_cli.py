"""bench.py's command-line contract, checked without a GPU: both error exits happen before
the library is loaded."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_unknown_section_is_refused():
    r = _bench("--sections", "widths,fold")
    assert r.returncode == 2
    assert "unknown sections ['fold']" in r.stderr
    assert r.stdout == ""
