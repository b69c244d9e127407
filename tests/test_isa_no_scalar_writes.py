"""The device code never writes through the scalar data cache.

Compiles every kernel source to gfx950 assembly and checks that no scalar
memory store, scalar atomic, or scalar-cache write-back/discard appears;
the batched kernel's work counter must use a vector (global) atomic.
Skipped where hipcc is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "cvortex_amd", "csrc")
FORBIDDEN = re.compile(r"^\s*s_(store|atomic|buffer_store|buffer_atomic|dcache_wb|dcache_discard|"
                       r"scratch_store)\w*", re.M)


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
@pytest.mark.parametrize("name", ["p3d", "p2d", "f3d", "batched", "order", "vort", "redist"])
def test_no_scalar_memory_writes(tmp_path, name):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = tmp_path / f"{name}.s"
    subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", SRC,
                    "--offload-device-only", "-S", os.path.join(SRC, f"{name}.hip"), "-o", str(out)],
                   check=True, capture_output=True)
    asm = out.read_text()
    hits = FORBIDDEN.findall(asm)
    assert not hits, f"{name}: scalar memory writes {sorted(set(hits))}"
    if name == "batched":
        assert "global_atomic_add" in asm
