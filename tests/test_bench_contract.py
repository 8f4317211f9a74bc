"""CPU checks of bench.py's command-line contract: one GPU only, one JSON line
on every exit path, and the build / smoke entry points."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT, has_gpu

BENCH = os.path.join(ROOT, "bench.py")


def _run(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=300, cwd=ROOT, env=env)
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    return r.returncode, lines


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], None), (["--gpus", "8"], None),
                                      ([], {"WORLD_SIZE": "2"})])
def test_more_than_one_gpu_exits_with_a_message(args, env):
    rc, lines = _run(args, env)
    assert rc == 2
    assert len(lines) == 1
    msg = json.loads(lines[0])
    assert "one GPU" in msg["error"] and "--gpus must be 1" in msg["error"]


@pytest.mark.skipif(has_gpu(), reason="checks the no-GPU exit path")
def test_no_gpu_exits_with_a_json_error():
    rc, lines = _run(["--steps", "1", "--warmup", "0"])
    assert rc == 1
    assert json.loads(lines[-1]) == {"error": "no GPU"}


def test_graft_entry_exposes_build_and_smoke():
    sys.path.insert(0, ROOT)
    try:
        import __graft_entry__ as g
    finally:
        sys.path.remove(ROOT)
    assert callable(g.build) and callable(g.smoke)
