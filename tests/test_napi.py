"""The Node N-API binding: tests/node/check_napi.js run by node (CPU checks
here, GPU checks marked gpu)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "tests", "node", "check_napi.js")
ADDON = os.path.join(ROOT, "radix_amd", "napi", "radix_napi.node")


def run_node(mode: str):
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    assert os.path.exists(ADDON), "radix_napi.node is not built"
    r = subprocess.run([node, SCRIPT, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_napi_cpu_checks():
    run_node("cpu")


@pytest.mark.gpu
def test_napi_gpu_checks():
    run_node("gpu")
