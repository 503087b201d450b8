"""In-tree build of the HIP library for gfx950.

``build()`` always recompiles ``ott_amd/csrc/*.hip`` into
``ott_amd/lib/libottx.so`` with hipcc (cross-compiling works without a GPU).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
CSRC = os.path.join(_HERE, "csrc")
LIB_DIR = os.path.join(_HERE, "lib")
LIB = os.path.join(LIB_DIR, "libottx.so")

SOURCES = ["ottx_sinkhorn.hip", "ottx_deepsets.hip"]
ARCH = os.environ.get("OTTX_OFFLOAD_ARCH", "gfx950")
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results stay in architectural VGPRs, so the exp/log-sum-exp
    # epilogue reads them without v_accvgpr_read copies.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def hipcc() -> str:
  path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
  if not os.path.exists(path):
    raise RuntimeError("hipcc not found; a ROCm toolchain is required to build ottx")
  return path


def build(verbose: bool = False) -> str:
  os.makedirs(LIB_DIR, exist_ok=True)
  tmp = LIB + ".tmp"
  cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp,
         *[os.path.join(CSRC, s) for s in SOURCES]]
  if verbose:
    print(" ".join(cmd), flush=True)
  proc = subprocess.run(cmd, capture_output=True, text=True)
  noise = "is not a recognized feature for this target"
  err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
  if proc.returncode != 0:
    raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{err}")
  if verbose and err.strip():
    print(err, file=sys.stderr)
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
