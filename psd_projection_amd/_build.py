"""Builds libpsdp.so (HIP kernels + C ABI) in-tree with hipcc for gfx950."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libpsdp.so"
OBJ = PKG / "_objs"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


# Scratch bytes per lane a kernel may use.  The 256 x 256 FMT_F16MX tile
# keeps 5 values computed before its main loop in scratch and reloads them
# after it (the loop itself holds 250 registers and touches no scratch).
SPILL_ALLOWED = {"_ZN4psdp23sym_prod_dma8_mx_kernelENS_8ProdDescEPKNS_7TileRefE": 32}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 8) -> Path:
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = []

    usage = {}

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        lines = (r.stdout + r.stderr).splitlines()
        msgs = "\n".join(l for l in lines if "packed-fp32-ops" not in l and "remark:" not in l
                         and "kernel-resource-usage" not in l)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{msgs}")
        if verbose and msgs.strip():
            print(msgs, flush=True)
        usage[src.name] = _parse_usage(lines)
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    report = OBJ / "kernel_resources.txt"
    with open(report, "w") as f:
        for src, kernels in sorted(usage.items()):
            for name, info in kernels.items():
                f.write(f"{src} {name} " + " ".join(f"{k}={v}" for k, v in info.items()) + "\n")
    spills = [(src, k) for src, ks in usage.items() for k, v in ks.items()
              if v.get("scratch", 0) > SPILL_ALLOWED.get(k, 0)]
    if spills:
        raise RuntimeError(f"kernels use scratch memory (register spills or private arrays): {spills}")
    return LIB


def _parse_usage(lines):
    """Per-kernel VGPR/AGPR/scratch/LDS figures from -Rpass-analysis remarks."""
    out, cur = {}, None
    keys = {"VGPRs:": "vgpr", "AGPRs:": "agpr", "ScratchSize [bytes/lane]:": "scratch",
            "LDS Size [bytes/block]:": "lds", "Occupancy [waves/SIMD]:": "occupancy"}
    for l in lines:
        if "Function Name:" in l:
            cur = l.split("Function Name:")[1].split("[-Rpass")[0].strip()
            out[cur] = {}
            continue
        if cur is None:
            continue
        for k, name in keys.items():
            if k in l:
                try:
                    out[cur][name] = int(l.split(k)[1].split("[-Rpass")[0].strip())
                except ValueError:
                    pass
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
