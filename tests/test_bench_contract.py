"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2; every mode parses its flags."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_flags_parse():
    sys.path.insert(0, ROOT)
    import bench
    old = sys.argv
    try:
        for argv in (["bench.py"], ["bench.py", "--spin", "--steps", "3"], ["bench.py", "--mixed", "--warmup", "0"],
                     ["bench.py", "--batch", "8", "--nside", "512", "--no-cpu", "--copy-calibration"],
                     ["bench.py", "--configs", "--no-cpu"], ["bench.py", "--cpu-full"],
                     ["bench.py", "--mixed", "--mixed-workload", "groups16", "--mixed-mode", "persistent"],
                     ["bench.py", "--mixed", "--mixed-workload", "groups16", "--mixed-nside", "512"]):
            sys.argv = argv
            a = bench.parse()
            assert a.gpus == 1 and a.steps >= 1
        sys.argv = ["bench.py"]
        assert bench.parse().batch == 32
    finally:
        sys.argv = old

