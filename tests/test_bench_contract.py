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


def test_product_raises_without_its_library():
    """No CPU fallback: with librkf_march.so absent, the package raises as
    soon as anything needs the library (a context, a scene build)."""
    code = (
        "import sys\n"
        "from rkifield_amd import march as M\n"
        "M.LIB_PATH = '/nonexistent/librkf_march.so'\n"
        "for f in (M.lib, lambda: M.Context(0)):\n"
        "    try:\n"
        "        f()\n"
        "    except RuntimeError as e:\n"
        "        assert 'is missing' in str(e), e\n"
        "    else:\n"
        "        sys.exit('no error')\n"
        "print('raised')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert r.stdout.strip() == "raised"
