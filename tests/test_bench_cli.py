"""bench.py's command line on the CPU: the one-GPU guard, the help text, and the JSON contract
keys of its output line as the source builds them (no GPU needed)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_help_lists_the_legs():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--sweeps-only", "--traffic-probe", "--forms-table", "--gs-solves"):
        assert flag in r.stdout


def test_output_line_carries_the_contract_keys():
    src = open(os.path.join(ROOT, "bench.py")).read()
    for key in ('"metric"', '"value"', '"unit"', '"n_gpus"', '"steps"', '"warmup"', '"ms_per_step"',
                '"higher_is_better"', '"scaling"', '"vs_baseline"', '"dtype"', '"data"', '"config"', '"roofline"',
                '"cpu_baseline"'):
        assert key in src, key
