"""In-tree build of the native libraries (gfx950).

    python -m cicy_geom_amd.build          # both libraries
    libcicy_dea.so  <- csrc/dea_solver.hip csrc/dea_mixed.hip ... (hipcc, gfx950)
    libcicy_net.so  <- csrc/netgen.cpp csrc/gnp_geom.cpp csrc/netio.cpp csrc/percolation.cpp
                     (g++, host only)

Every HIP source is compiled with packed-FP32 instruction generation disabled
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["dea_solver.hip", "dea_setup.hip", "dea_mixed.hip", "dea_batched.hip", "dea_mpnn.hip",
               "dea_resident.hip", "dea_condensed.hip", "dea_core.hip", "dea_backbone.hip",
               "dea_contacts.hip", "dea_assembly.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_net(out_dir: str = HERE) -> str:
    out = os.path.join(out_dir, "libcicy_net.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          f"-I{INCLUDE}", os.path.join(CSRC, "netgen.cpp"), os.path.join(CSRC, "gnp_geom.cpp"),
          os.path.join(CSRC, "netio.cpp"), os.path.join(CSRC, "percolation.cpp"), "-o", out])
    return out


def build_dea(out_dir: str = HERE) -> str:
    """Each HIP source is compiled to an object on its own process (in parallel), then the
    objects are linked into the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    import tempfile
    out = os.path.join(out_dir, "libcicy_dea.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="cicy_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(lambda so: _run([HIPCC, *cflags, "-c", so[0], "-o", so[1]]),
                        zip(srcs, objs)))
        tmp = out + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-fPIC", "-shared", *objs, "-o", tmp])
        os.replace(tmp, out)
    return out


def build_cpp_dropin_test() -> str:
    """The C++ drop-in check program (tests/cpp/dropin_chain.cpp) linked against
    libcicy_dea.so with an rpath to the package directory."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_chain.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "dropin_chain")
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", src, f"-L{HERE}", "-lcicy_dea",
          "-Wl,-rpath,$ORIGIN/../../cicy_geom_amd", "-L/opt/rocm/lib", "-lamdhip64",
          "-o", out])
    return out


def build_all() -> None:
    build_net()
    build_dea()
    build_cpp_dropin_test()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "net" in what:
        build_net()
    if "all" in what or "dea" in what:
        build_dea()
