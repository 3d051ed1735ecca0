"""Build libnixis_hip.so (gfx950) in-tree with hipcc.

Usage: python -m nixis_amd.build   (or nixis_amd.build.build() from Python)

Every HIP source is compiled with
  -ffp-contract=off      the FP64 lattice section must keep the reference's
                         separate multiply/add roundings; FP32 fused
                         multiply-adds are written explicitly with fmaf;
  -packed-fp32-ops off   packed FP32 instructions are not generated.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnixis_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NX_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-bitwise-instead-of-logical", "-I", os.path.join(HERE, "..", "include")]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src, objdir):
    obj = os.path.join(objdir, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
                                  "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=False):
    """Compile every source and link libnixis_hip.so (always rebuilds)."""
    with tempfile.TemporaryDirectory(prefix="nx_build_") as objdir:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            results = list(ex.map(lambda s: _compile(s, objdir), sources()))
        for _, warn in results:
            if warn and verbose:
                print(warn, file=sys.stderr)
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp,
               *[o for o, _ in results], "-pthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
