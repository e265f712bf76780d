"""bench.py's command-line contract that needs no GPU: it measures one GPU,
so a request for more (--gpus N > 1, or a launcher's WORLD_SIZE > 1) exits
with status 2 and says why, before it touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, world=None):
    env = dict(os.environ)
    if world is not None:
        env["WORLD_SIZE"] = str(world)
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    for args, world in ((["--gpus", "2"], None), (["--gpus", "8", "--steps", "1"], None), ([], 2)):
        r = _run(args, world)
        assert r.returncode == 2, (args, world, r.stdout, r.stderr)
        assert "measures one GPU" in r.stderr
        assert r.stdout == ""
