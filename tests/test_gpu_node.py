"""The Node.js drop-in (N-API addon + JS shim of NBodyGPU.js / mortonTree.js)
against the golden vectors of the reference's mortonTree.js, on the GPU."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed on this machine")
def test_node_shim_parity():
    addon = os.path.join(ROOT, "galaxy-sim_amd", "node", "galaxy_hip.node")
    assert os.path.exists(addon), "N-API addon not built"
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "shim_parity.mjs")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node shim parity ok" in r.stdout


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed on this machine")
def test_node_pool_parity():
    """GalaxyPool.js slots equal NBodyGPU.js run alone, bit for bit."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "pool_parity.mjs")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node pool parity ok" in r.stdout


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed on this machine")
def test_node_stellar_flow():
    """stellarBuffer follows applyPermutation; encodeStarFormation forms stars."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "stellar_flow.mjs")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node stellar flow ok" in r.stdout
