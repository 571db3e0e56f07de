"""Build the gfx950 engine library ``dtmm_amd/libdtmmx.so`` with hipcc.

Usage: ``python -m dtmm_amd.build`` (or ``dtmm_amd.build.build()``).
"""
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "dtmmx.hip")
OUT = os.path.join(HERE, "libdtmmx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # never emit packed fp32 arithmetic (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", os.path.join(HERE, "csrc"),
]


# Translation units compiled in parallel (host + ABI, complex64, complex128).
PARTS = (1, 2, 3)


def _source_hash():
    # _lib's hash of the sources, loaded by path so that building needs no
    # package import
    import importlib.util
    spec = importlib.util.spec_from_file_location("_dtmmx_lib", os.path.join(HERE, "_lib.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.source_hash()


def build(verbose=True, defines=(), out=None):
    """Compile the library (always recompiles).  `defines` adds -D flags and
    `out` another output path (for build variants)."""
    out = out or OUT
    t = time.time()
    srchash = _source_hash()
    tmp = tempfile.mkdtemp(prefix="dtmmx_build_")
    try:
        objs, procs = [], []
        for part in PARTS:
            obj = os.path.join(tmp, "part%d.o" % part)
            cmd = ([HIPCC] + FLAGS + ["-c", "-DDTMMX_PART=%d" % part] +
                   ["-D" + d for d in defines] + [SRC, "-o", obj])
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
            objs.append(obj)
        codes = [p.wait() for p in procs]
        if any(codes):
            raise RuntimeError("hipcc failed (exit codes %s)" % (codes,))
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    os.replace(out + ".tmp", out)
    # record of the sources compiled (checked when the library is loaded)
    with open(out + ".srchash", "w") as fh:
        fh.write("%s %s\n" % (srchash, " ".join(defines)))
    if verbose:
        print("built %s in %.1f s" % (out, time.time() - t), flush=True)
    return out


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
