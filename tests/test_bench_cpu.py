"""bench.py's command line without a GPU: the script compiles, lists its
options, and refuses more than one GPU with a JSON line (this build runs on
one GPU)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_bench_help():
    r = bench("--help")
    assert r.returncode == 0, r.stderr
    for opt in ("--steps", "--warmup", "--gpus", "--slice-batch", "--queue", "--variants"):
        assert opt in r.stdout


def test_bench_refuses_more_than_one_gpu():
    r = bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
