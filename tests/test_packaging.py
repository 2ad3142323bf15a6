"""Packaging (reference src/setup.py:99-124, README.md:74-86): ``pip install`` of the repository
builds libparament.so with hipcc and installs it inside the package, and ``pytest --pyargs
parament_amd`` then runs the shipped tests from the installed copy.  CPU only: the GPU-marked
tests are deselected, and hipcc cross-compiles without a device."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs hipcc")
def test_pip_install_then_pytest_pyargs(tmp_path):
    src = tmp_path / "src"
    src.mkdir()
    shutil.copy(os.path.join(ROOT, "setup.py"), src)
    shutil.copytree(os.path.join(ROOT, "include"), src / "include")
    shutil.copytree(os.path.join(ROOT, "parament_amd"), src / "parament_amd",
                    ignore=shutil.ignore_patterns("*.so", "_obj", "__pycache__", "*.tmp"))
    site = tmp_path / "site"
    env = dict(os.environ, PIP_NO_INDEX="1", PIP_DISABLE_PIP_VERSION_CHECK="1")
    res = subprocess.run([sys.executable, "-m", "pip", "install", "--no-build-isolation", "--no-deps", "--no-index",
                          "--target", str(site), str(src)], capture_output=True, text=True, env=env, timeout=900)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert (site / "parament_amd" / "libparament.so").exists()
    assert (site / "parament_amd" / "test" / "test_error.py").exists()
    # run from an empty directory, so that only the installed copy is importable
    run = tmp_path / "run"
    run.mkdir()
    env = dict(os.environ, PYTHONPATH=str(site))
    res = subprocess.run([sys.executable, "-c", "import parament_amd, sys; print(parament_amd.__file__)"],
                         capture_output=True, text=True, env=env, cwd=run)
    assert res.returncode == 0 and res.stdout.strip().startswith(str(site)), res.stdout + res.stderr
    res = subprocess.run([sys.executable, "-m", "pytest", "--pyargs", "parament_amd", "-m", "not gpu", "-q",
                          "-p", "no:cacheprovider"], capture_output=True, text=True, env=env, cwd=run, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert " passed" in res.stdout and "failed" not in res.stdout
