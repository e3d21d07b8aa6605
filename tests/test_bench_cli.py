"""bench.py's command-line contract that needs no GPU: more than one GPU and
unknown configurations are refused with exit status 2 and a message."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    for extra in ((), ("--config", "4")):
        r = _bench("--gpus", "2", *extra)
        assert r.returncode == 2, r.stderr
        assert "one GPU" in r.stderr
        assert r.stdout.strip() == ""


def test_unknown_configuration_is_refused():
    r = _bench("--config", "9")
    assert r.returncode == 2
    assert "unknown configuration" in r.stderr
