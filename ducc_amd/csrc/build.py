"""Build libducc_amd.so (gfx950) in-tree.

Steps: generate _gen/kernel_db.inc from ducc_amd/data/kernel_db.csv, compile
every source with hipcc in parallel into _build/, link the shared library next
to the Python package.  Run: python ducc_amd/csrc/build.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libducc_amd.so")
BUILD = os.path.join(HERE, "_build")
GEN = os.path.join(HERE, "_gen")

SOURCES = ["params.cpp", "runtime.hip", "api.hip", "sort.hip", "spread.hip",
           "interp.hip", "fft.hip", "grid.hip", "chooser.hip", "mlp.hip",
           "c2c.hip", "noalloc.hip", "spread_det.hip", "records.hip",
           "sht.hip", "wgridder.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DMX_ARCH", "gfx950")
INCLUDE = os.environ.get("DMX_INCLUDE",
                         os.path.join(os.path.dirname(PKG), "include"))
COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", "-I" + HERE,
          "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-flush-denormals-to-zero"]


def gen_kernel_db():
    os.makedirs(GEN, exist_ok=True)
    src = os.path.join(PKG, "data", "kernel_db.csv")
    lines = []
    with open(src) as f:
        for ln in f:
            if ln.startswith("#") or not ln.strip():
                continue
            W, of, eps, beta, e0, nd, sp = ln.strip().split(",")
            lines.append(f"{{{W}, {of}, {eps}, {beta}, {e0}, {nd}, {sp}}},")
    body = "\n".join(lines) + "\n"
    dst = os.path.join(GEN, "kernel_db.inc")
    if not os.path.exists(dst) or open(dst).read() != body:
        with open(dst, "w") as f:
            f.write(body)


def _newest_header():
    hdrs = [os.path.join(HERE, f) for f in os.listdir(HERE)
            if f.endswith(".h")]
    hdrs.append(os.path.join(INCLUDE, "ducc_amd.h"))
    hdrs.append(os.path.join(GEN, "kernel_db.inc"))
    return max(os.path.getmtime(h) for h in hdrs if os.path.exists(h))


def compile_one(src, incremental=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(HERE, src)
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(path) and t > _newest_header():
            return obj, ""
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"compile of {src} failed:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False, incremental=False):
    """Compile every source (incremental: only those older than their
    object or than a header) and link the library."""
    gen_kernel_db()
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, err in ex.map(lambda f: compile_one(f, incremental),
                               SOURCES):
            objs.append(obj)
            if verbose and err.strip():
                print(err, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + \
        objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-i", action="store_true",
                    help="incremental: recompile changed sources only")
    a = ap.parse_args()
    print(build(a.j, a.v, a.i))
