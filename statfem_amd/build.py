"""Builds libstatfem_hip.so (gfx950) from statfem_amd/csrc/*.hip with hipcc.

Each translation unit is compiled in parallel to an object file under
statfem_amd/_build/, then linked into statfem_amd/libstatfem_hip.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstatfem_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-DNDEBUG",
         # packed FP32 VALU ops are disabled for all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith(".hip"))


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(os.path.basename(f) for f in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cpp_client()
    return LIB


CPP_CLIENT = os.path.join(OUT, "system_solve_hip")
CPP_PIPELINE = os.path.join(OUT, "statfem_pipeline")


def build_cpp_client():
    """The C++ drop-in test clients (tests/cpp/system_solve_hip.cpp: the
    SystemSolve concept; tests/cpp/statfem_pipeline.cpp: AssemblerHip and the
    statX classes) against the header-only classes and the library, with the
    host compiler."""
    for src, exe in (("system_solve_hip.cpp", CPP_CLIENT), ("statfem_pipeline.cpp", CPP_PIPELINE)):
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(HERE, "..", "tests", "cpp", src), "-L", HERE, "-lstatfem_hip",
               "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        subprocess.run(cmd, check=True)
    return CPP_CLIENT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
