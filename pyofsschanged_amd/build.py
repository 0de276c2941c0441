"""Build libgnlse_hip.so in-tree: every translation unit under csrc/ is
compiled for gfx950 in parallel, then linked into one shared library.

    python -m pyofsschanged_amd.build            # build (always recompiles)
    python -m pyofsschanged_amd.build --jobs 4
"""

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgnlse_hip.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, extra, objdir=OBJ):
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(jobs=None, extra=(), verbose=False, out=LIB, objdir=OBJ):
    """Compile every csrc/*.hip for gfx950 and link ``out``. ``extra`` holds
    additional hipcc flags, e.g. -DGN_T_8192=1024 to pick another workgroup
    size for a grid."""
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 16)
    objs = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(lambda s: _compile(s, list(extra), objdir), srcs):
            # the host pass does not know the packed-fp32 feature; that warning is expected
            lines = [l for l in log.splitlines() if "packed-fp32-ops" not in l]
            if rc != 0:
                sys.stderr.write("\n".join(lines) + "\n")
                raise RuntimeError(f"hipcc failed for {src}")
            if verbose and lines:
                sys.stderr.write("\n".join(lines) + "\n")
            objs.append(obj)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"link of {out} failed")
    os.replace(tmp, out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--resource-usage", action="store_true", help="print per-kernel register/LDS usage")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor define")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("--objdir", default=OBJ)
    a = ap.parse_args(argv)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if a.resource_usage else []
    extra += ["-D" + d for d in a.defines]
    print(build(a.jobs, extra, verbose=a.resource_usage, out=a.out, objdir=a.objdir))


if __name__ == "__main__":
    main()
