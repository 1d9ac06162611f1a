"""
Build ``libxentica_hip.so`` in-tree with hipcc for gfx950.

Run ``python -m xentica_amd.build`` (or ``__graft_entry__.build()``).
Every HIP translation unit under ``csrc/`` goes into the one library; the
flags disable packed FP32 instructions and FP contraction.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "--offload-arch=%s" % ARCH, "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    """All translation units of the library."""
    files = sorted(glob.glob(os.path.join(HERE, "csrc", "*.cpp")))
    files += sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    return files


def build(verbose=False):
    """Compile every unit and link ``xentica_amd/libxentica_hip.so``."""
    out_dir = os.path.join(ROOT, "_build")
    os.makedirs(out_dir, exist_ok=True)
    jobs = []
    for src in sources():
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        cmd = [HIPCC] + FLAGS + [
            "-I", os.path.join(ROOT, "include"),
            "-I", os.path.join(HERE, "csrc"),
            "-c", src, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd))
        jobs.append((cmd, obj))
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for proc in pool.map(lambda job: subprocess.run(
                job[0], capture_output=True, text=True), jobs):
            if proc.returncode:
                sys.stderr.write(proc.stdout + proc.stderr)
                raise subprocess.CalledProcessError(proc.returncode, proc.args)
    objects = [obj for _, obj in jobs]
    target = os.path.join(HERE, "libxentica_hip.so")
    tmp = target + ".tmp%d" % os.getpid()
    cmd = [HIPCC, "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp]
    cmd += objects + ["-lhiprtc"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
