"""Build recipe for the HIP library (gfx950) and the host CLI.

    python -m kerrtracecpp2_amd.build          # libkt_hip.so + bin/kerr_tracer_hip
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for all device code (-packed-fp32-ops)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-ffp-contract=on",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # machine-level loop-invariant code motion hoists every 64-bit
             # constant of the integrator into a VGPR pair; without it the
             # binary64 trace kernel drops from 236 to 179 VGPRs and the bundle
             # and adjoint-forward kernels stop spilling
             "-mllvm", "-disable-machine-licm"]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    src = os.path.join(PKG, "csrc", "kt_hip.hip")
    out = os.path.join(PKG, "libkt_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, src, "-lz"])
    os.replace(tmp, out)
    return out


def build_cli():
    src = os.path.join(PKG, "csrc", "kerr_tracer_hip.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    out = os.path.join(PKG, "bin", "kerr_tracer_hip")
    run(["g++", "-std=c++17", "-O2", "-Wall", "-o", out, src, f"-I{ROOT}/include", f"-L{PKG}", "-lkt_hip",
         "-Wl,-rpath,$ORIGIN/..", "-lz", "-lpthread"])
    return out


def main():
    build_lib()
    build_cli()
    return 0


if __name__ == "__main__":
    sys.exit(main())
