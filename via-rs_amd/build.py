"""Build recipe for libvia_gpu.so (gfx950) and the oracle checker library.

    python via-rs_amd/build.py            # HIP library only
    python via-rs_amd/build.py --oracle   # also the CPU restatement (make -C oracle)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/ntt.hip", "csrc/kernels.hip", "csrc/first_dim.hip", "csrc/repack.hip", "csrc/prg.hip", "csrc/noalloc.hip", "csrc/cascade_mac.hip", "csrc/gate_fx.hip", "csrc/via_gpu.cpp",
           "csrc/serve.cpp", "csrc/wire.cpp", "csrc/client.cpp",
           "csrc/client_keys.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # exact fp64 residue arithmetic relies on unfused products
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def build_hip(out=None):
    out = out or os.path.join(HERE, "libvia_gpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(HERE, s) for s in SOURCES]
    subprocess.run(cmd, check=True, cwd=HERE)
    os.replace(tmp, out)
    return out


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


if __name__ == "__main__":
    print(build_hip())
    if "--oracle" in sys.argv:
        build_oracle()
