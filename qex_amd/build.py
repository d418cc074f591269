"""Build the qexhip HIP library (gfx950) in-tree: qex_amd/libqexhip.so.

Run as `python -m qex_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["dslash.hip", "kernels.hip", "reduce.hip", "gen.hip", "gauge.hip", "batch.hip", "smear.hip", "milcio.hip", "scidac.hip", "qexlayout.hip", "api.hip"]
OUT = os.path.join(HERE, "libqexhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QEXHIP_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=fast",
    # no packed-FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC,
    "-I" + os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=True, jobs=4, out=OUT, defines=()):
    objs = []
    procs = []
    for src in SOURCES:
        obj = out + "." + src.replace(".hip", ".o")
        cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-z,defs", "-o", out + ".tmp"] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def _wait(procs):
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass warns that it does not know the packed-fp32 feature
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            print("\n".join(lines), flush=True)
        if p.returncode != 0:
            failed.append(src)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))


if __name__ == "__main__":
    build()
    sys.exit(0)
