"""The reference-side C++ shim compiles against the reference's own headers
(syntax check with g++; the reference tree is only present in the build
container)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_shim_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "use_shim.cpp"
    src.write_text(
        '#include "cfdxhip/hip_amg_preconditioner.hpp"\n'
        "int main() {\n"
        "  cfdxhip_shim::HipAmgPreconditioner pc;\n"
        "  cfdx::core::Preconditioner& base = pc;\n"
        "  (void)base;\n"
        "  return 0;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{REF}", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
