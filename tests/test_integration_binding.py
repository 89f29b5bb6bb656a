"""The reference-side binding shown in INTEGRATION.md (a RenderContextImpl
subclass forwarding to include/rive_hip.h) compiles against the reference's
own headers: the C++ block is extracted from the document and checked with
g++ -fsyntax-only.  Skipped where the reference tree is absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "renderer", "include")),
                    reason="reference tree not present")
def test_binding_compiles_against_reference_headers(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    section = text[text.index("## The reference-side binding"):]
    block = re.search(r"```cpp\n(.*?)```", section, re.S).group(1)
    assert "class RenderContextHIPImpl : public RenderContextImpl" in block
    src = tmp_path / "render_context_hip_impl.cpp"
    src.write_text(block)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-attributes",
                        "-I", os.path.join(REFERENCE, "include"),
                        "-I", os.path.join(REFERENCE, "renderer", "include"),
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # every pure virtual of RenderContextImpl is overridden (an abstract class
    # could not be constructed by make_unique / unique_ptr<...>(new ...))
    probe = tmp_path / "probe.cpp"
    probe.write_text(block + "\nvoid probe() { auto c = rive::gpu::RenderContextHIPImpl::MakeContext(0); (void)c; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-attributes",
                        "-I", os.path.join(REFERENCE, "include"),
                        "-I", os.path.join(REFERENCE, "renderer", "include"),
                        "-I", os.path.join(ROOT, "include"), str(probe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
