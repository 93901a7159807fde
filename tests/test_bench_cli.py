"""bench.py measures one GPU: asked for more (--gpus N > 1, or launched as one rank of several) it exits with status 2
and says why, before touching any device."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}), ([], {"WORLD_SIZE": "4"})])
def test_more_than_one_gpu_is_refused(argv, env):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), *argv], cwd=ROOT, capture_output=True, text=True,
                       timeout=300, env={**os.environ, **env})
    assert r.returncode == 2, r.stderr[-2000:]
    assert "single MI355X" in r.stderr
    assert r.stdout.strip() == ""
