"""Build the test helpers in tests/cpp (user-weight callback library)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
FLAGS = ["-O2", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


ROOT = os.path.dirname(os.path.dirname(HERE))
SHELL_SYS = ["N=10", "m=3", "w=0.05", "eps=1e-4", "strength=0", "exponent=0"]


def build():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(HERE, "libuserweights.so")
    subprocess.run([hipcc] + FLAGS + ["-shared", os.path.join(HERE, "user_weights.hip"), "-o", out + ".tmp"],
                   check=True)
    os.replace(out + ".tmp", out)
    # sys_params.h of the shell driver (N=10, m=3, V = 0.025 r^2 + 1e-4 r^4), then the driver
    gen = os.path.join(HERE, "sysgen")
    subprocess.run(["g++", "-O2", "-o", gen, os.path.join(ROOT, "laughlinedge_amd", "system_parameters.cpp")],
                   check=True)
    subprocess.run([gen, os.path.join(HERE, "sys_params.h")] + SHELL_SYS, check=True)
    os.remove(gen)
    drv = os.path.join(HERE, "shell_driver")
    libdir = os.path.join(ROOT, "laughlinedge_amd")
    subprocess.run([hipcc] + FLAGS + ["-x", "hip", "-I", HERE, "-I", os.path.join(ROOT, "include"),
                                      os.path.join(HERE, "shell_driver.cpp"), "-o", drv + ".tmp",
                                      "-L", libdir, "-llaughlinedge", "-Wl,-rpath,$ORIGIN/../../laughlinedge_amd"],
                   check=True)
    os.replace(drv + ".tmp", drv)
    # failure-path driver (no GPU call on any of its paths)
    fdrv = os.path.join(HERE, "failure_driver")
    subprocess.run([hipcc] + FLAGS + ["-x", "hip", "-I", HERE, "-I", os.path.join(ROOT, "include"),
                                      os.path.join(HERE, "failure_driver.cpp"), "-o", fdrv + ".tmp",
                                      "-L", libdir, "-llaughlinedge", "-Wl,-rpath,$ORIGIN/../../laughlinedge_amd"],
                   check=True)
    os.replace(fdrv + ".tmp", fdrv)
    return out


if __name__ == "__main__":
    print(build())
