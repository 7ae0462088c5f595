"""Build the test helpers in tests/cpp (user-weight callback library)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
FLAGS = ["-O2", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(HERE, "libuserweights.so")
    subprocess.run([hipcc] + FLAGS + ["-shared", os.path.join(HERE, "user_weights.hip"), "-o", out + ".tmp"],
                   check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
