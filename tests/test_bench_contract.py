"""Host-side contract of bench.py and of the package boundary (no GPU)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_package_does_not_load_the_oracle():
    """Importing the product package and every module of it leaves the CPU
    oracle unloaded: nothing on the product path can route through it."""
    code = ("import sys, os, importlib, random_amd\n"
            "for f in sorted(os.listdir(random_amd.__path__[0])):\n"
            "    if f.endswith('.py') and f != '__init__.py':\n"
            "        importlib.import_module('random_amd.' + f[:-3])\n"
            "bad = [n for n in sys.modules if n == 'oracle' or n.startswith('oracle.')]\n"
            "print(','.join(bad))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == ""


def test_missing_library_fails_loudly(monkeypatch):
    """Without libktm.so the package raises instead of falling back."""
    import pytest

    from random_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libktm.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()
