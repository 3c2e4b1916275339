"""Host unit check of the look-back workspace registry
(tacuda_amd/csrc/ws_registry.hpp): built with g++ and run, no GPU."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_workspace_registry_rules():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "ws_registry_check")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe,
                        os.path.join(HERE, "ws_registry_check.cpp")], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0, out.stderr
        assert "ws_registry ok" in out.stdout
