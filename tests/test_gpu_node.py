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
def test_node_arena_parity():
    """GalaxyPool.js GalaxyArena: placed, removed and re-placed galaxies equal
    NBodyGPU.js run alone, bit for bit; a full arena throws GS_E_FULL."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "arena_parity.mjs")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node arena parity ok" in r.stdout


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed on this machine")
def test_node_stellar_flow():
    """stellarBuffer follows applyPermutation; encodeStarFormation forms stars."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "stellar_flow.mjs")],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node stellar flow ok" in r.stdout


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed on this machine")
def test_node_step_is_the_reference_step(tmp_path):
    """NBodyGPU.step() walks without refreshing (NBodyGPU.js:321-329): after
    updateCentralMass and no refresh, the JS step equals the oracle's walk of
    the unrefreshed tree, bit for bit (EXACT force path)."""
    import numpy as np

    from oracle import oracle

    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node", "step_semantics.mjs"),
                        str(tmp_path)], capture_output=True, text=True, timeout=240, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "node step semantics ok" in r.stdout
    before = np.fromfile(tmp_path / "before.f32", np.float32)
    nodes = np.fromfile(tmp_path / "nodes.u32", np.uint32)
    after = np.fromfile(tmp_path / "after.f32", np.float32)
    n = before.size // 8
    op = oracle.pack_params(n, dt=0.005, G=1.0, epsilon=0.02, theta=0.6, gas_drag=0.25)
    want, _, _, _, _ = oracle.walk(before, nodes, op)
    assert np.array_equal(after.view(np.uint32), want.view(np.uint32))
    # the central mass changed, so a refreshed tree would give other forces
    fresh = oracle.refresh_f64(before, nodes)
    assert not np.array_equal(fresh, nodes)
