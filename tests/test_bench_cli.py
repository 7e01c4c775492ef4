"""bench.py's command line (no GPU needed): more than one GPU is refused with a
message and a non-zero status; --help lists the workload options."""
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def test_help_lists_workloads():
    p = run("--help")
    assert p.returncode == 0
    for opt in ("--workload", "--steps", "--warmup", "--precision-report", "--pair-shape"):
        assert opt in p.stdout
