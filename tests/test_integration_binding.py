"""CPU: the reference-side binding shown in the integration notes compiles against
the reference's own public headers (include/fyzzxyzzy/world.h) and this
repo's C-ABI (include/fyzz_hip.h): every `WorldBackend` pure virtual is
overridden, every override matches a reference signature, and the record
layouts the spans reinterpret are checked by static_assert. Syntax only
(g++ -std=c++23 -fsyntax-only): nothing is linked or run. Needs the
reference tree, which only the build container has."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF_INCLUDE = Path("/root/reference/include")


def _blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"<!-- compiled: hip_world_backend\.h -->\s*```cpp\n(.*?)```", text, re.S)
    assert len(blocks) >= 2, "the integration notes lost their compiled backend blocks"
    return blocks


@pytest.mark.skipif(not (REF_INCLUDE / "fyzzxyzzy" / "world.h").exists() or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_backend_binding_compiles_against_reference_headers(tmp_path):
    # the header as a maintainer would include it, then a translation unit
    (tmp_path / "hip_world_backend.h").write_text("\n".join(_blocks()))
    src = tmp_path / "hip_world_backend.cpp"
    body = '#include "hip_world_backend.h"\n'
    # instantiate the class so every override is checked, and make World
    # constructible from it
    body += ("\nnamespace fyzz { [[maybe_unused]] static World (*const make)(HipDevice, WorldOptions)"
             " = &create_native_gpu_world; }\n"
             "static_assert(!std::is_abstract_v<fyzz::HipWorldBackend>);\n")
    src.write_text("#include <type_traits>\n" + body)
    r = subprocess.run(["g++", "-std=c++23", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-function",
                        f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
