"""The C++ binding in INTEGRATION.md section 1 compiles against the reference's own
headers (include/gpu/compression/compression_workflow_gpu.hpp,
include/gpu/decompression/decompression_workflow_gpu.hpp) and this library's
include/gfaz_amd.h: g++ -fsyntax-only, no GPU. Skipped where the reference sources
are absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"


def first_cpp_block(md: str) -> str:
    m = re.search(r"```cpp\n(.*?)```", md, flags=re.S)
    assert m, "INTEGRATION.md has no C++ block"
    return m.group(1)


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or not shutil.which("g++"),
                    reason="reference headers or g++ absent")
def test_integration_binding_compiles(tmp_path):
    src = tmp_path / "compression_workflow_amd.cpp"
    src.write_text(first_cpp_block(open(os.path.join(ROOT, "INTEGRATION.md")).read()))
    cmd = ["g++", "-std=c++17", "-fsyntax-only", f"-I{REF_INCLUDE}",
           f"-I{os.path.join(ROOT, 'include')}", "-I/opt/conda/include", str(src)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
