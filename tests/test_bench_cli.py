"""bench.py's command line, without a GPU: more than one GPU is refused with
a message before anything loads, and the help lists every section."""
import os
import subprocess
import sys

from tests.conftest import ROOT


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_sections():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--only", "--no-batch", "--no-serve", "--no-variants", "--variants", "--calib-mib"):
        assert flag in r.stdout, flag
