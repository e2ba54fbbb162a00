"""Build of tests/dropin/dropin_check (test-only host program, g++ against libkt_hip.so).

The code block of INTEGRATION.md section 2 (the USE_HIP branch of render_image) is copied
verbatim into _build/snippet.inc, minus its `#elif` and `#include` lines, so the documented
integration is what gets compiled.

    python tests/dropin/build.py [OUT_DIR]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "kerrtracecpp2_amd")


def snippet(text=None):
    """The USE_HIP code block of INTEGRATION.md, without the preprocessor lines."""
    if text is None:
        text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    code = [b for b in blocks if "#elif defined(USE_HIP)" in b]
    if len(code) != 1:
        raise RuntimeError("INTEGRATION.md: expected one cpp block with the USE_HIP branch")
    lines = [ln for ln in code[0].splitlines() if not ln.lstrip().startswith(("#elif", "#include"))]
    return "\n".join(lines) + "\n"


def build(out_dir=None):
    out_dir = out_dir or os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "snippet.inc"), "w") as f:
        f.write(snippet())
    exe = os.path.join(out_dir, "dropin_check")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-Wno-unused-parameter", "-o", exe,
                           os.path.join(HERE, "dropin_check.cpp"), f"-I{out_dir}", f"-I{ROOT}/include",
                           f"-I{PKG}/csrc", f"-L{PKG}", "-lkt_hip", f"-Wl,-rpath,{PKG}"])
    return exe


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else None))
