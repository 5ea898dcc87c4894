"""Builds the in-tree native parts of dbscan_torch for gfx950.

  libdbscan_hip.so   hipcc, the HIP kernels and the C-ABI (include/dbscan_hip.h)
  _C.<abi>.so        g++, the pybind/torch binding that calls the C-ABI

Both land next to dbscan_torch/__init__.py.  Usage: python dbscan-torch_amd/build.py
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG_ROOT = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(PKG_ROOT)
PKG_DIR = os.path.join(PKG_ROOT, "dbscan_torch")
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO_ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("DBSCAN_HIP_ARCH", "gfx950")

HIP_LIB = os.path.join(PKG_DIR, "libdbscan_hip.so")
EXT_LIB = os.path.join(PKG_DIR, "_C" + sysconfig.get_config_var("EXT_SUFFIX"))

# -ffp-contract=off: the squared distance stays two multiplies and an add, as
# in the reference's x86 build.  Denormals are kept (no flush-to-zero), and
# packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that
    # line is harmless and is dropped from the log.
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, file=sys.stderr)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def hip_command() -> list[str]:
    src = os.path.join(CSRC, "hip", "dbscan_hip.hip")
    return [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, f"-I{INCLUDE}", "-o", HIP_LIB, src]


def ext_command() -> list[str]:
    import torch

    tdir = os.path.dirname(torch.__file__)
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    import pybind11

    return [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
        "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=_C",
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        f"-I{INCLUDE}", f"-I{tdir}/include", f"-I{tdir}/include/torch/csrc/api/include",
        f"-I{ROCM}/include", f"-I{sysconfig.get_paths()['include']}", f"-I{pybind11.get_include()}",
        os.path.join(CSRC, "extension.cpp"),
        "-o", EXT_LIB,
        f"-L{PKG_DIR}", "-ldbscan_hip",
        f"-L{tdir}/lib", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip",
        "-ltorch_python",
        "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tdir}/lib",
    ]


def write_build_info() -> None:
    import torch

    with open(os.path.join(PKG_DIR, "_build_info.py"), "w") as f:
        f.write('"""Written by build.py at build time."""\n')
        f.write(f'BUILD_TORCH_VERSION = "{torch.__version__}"\n')
        f.write(f'BUILD_ARCH = "{ARCH}"\n')


def build() -> None:
    """Compile both libraries (the binding links against the kernel library)."""
    _run(hip_command())
    _run(ext_command())
    write_build_info()


def build_parallel(extra: list[list[str]] | None = None) -> None:
    """Like build(), overlapping the slow binding compile with other commands."""
    obj = EXT_LIB + ".o"
    ext = ext_command()
    # compile the binding to an object first (independent of the HIP library)
    i_out = ext.index("-o")
    compile_cmd = [a for a in ext[: i_out] if a != "-shared"] + ["-c", "-o", obj]
    link_cmd = (["g++", "-shared", "-o", EXT_LIB, obj] + ext[i_out + 2:])
    jobs = [hip_command(), compile_cmd] + list(extra or [])
    with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
        for f in [pool.submit(_run, j) for j in jobs]:
            f.result()
    _run(link_cmd)
    os.remove(obj)
    write_build_info()


if __name__ == "__main__":
    build_parallel()
    print(HIP_LIB)
    print(EXT_LIB)
