"""Build libchimera_hip.so (gfx950) in-tree with hipcc, and libchimera_h5.so.

Every HIP source under chimera_amd/csrc is compiled for gfx950 with
-ffp-contract=off (no FMA contraction: the results must match the reference
Fortran bit for bit) and with packed-FP32 instruction selection disabled, then
linked into chimera_amd/libchimera_hip.so.  The snapshot writer (csrc/h5io.c) is
plain C over the HDF5 C library and goes into its own chimera_amd/libchimera_h5.so,
so the solver library does not depend on HDF5.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libchimera_hip.so"
LIB_H5 = PKG / "libchimera_h5.so"
HDF5_PREFIX = Path(os.environ.get("CHIMERA_HDF5_PREFIX", "/opt/conda"))
OBJ = PKG / "_obj"

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile all HIP sources and link the shared library; returns its path."""
    OBJ.mkdir(exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "--offload-arch=gfx950", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


def build_h5(verbose: bool = False) -> Path | None:
    """Build the snapshot writer against the HDF5 C library; None when HDF5 is absent."""
    inc, lib = HDF5_PREFIX / "include", HDF5_PREFIX / "lib"
    if not (inc / "hdf5.h").exists():
        if verbose:
            print(f"HDF5 not found under {HDF5_PREFIX}; libchimera_h5.so not built", file=sys.stderr)
        return None
    tmp = LIB_H5.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", f"-I{inc}", "-o", str(tmp),
           str(CSRC / "h5io.c"), f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed on h5io.c:\n{r.stderr}")
    os.replace(tmp, LIB_H5)
    if verbose:
        print(f"built {LIB_H5}", file=sys.stderr)
    return LIB_H5


def build_checks(verbose: bool = False) -> Path | None:
    """Build the GPU check library of the tests (tests/host/libshort_forms_check.so:
    the stage kernel's short division and square root against hipcc's IEEE forms) with
    the library's own flags; None when the tests are not in the tree."""
    src = PKG.parent / "tests" / "host" / "short_forms_check.hip"
    if not src.exists():
        return None
    out = src.with_name("libshort_forms_check.so")
    tmp = out.with_suffix(".so.tmp")
    OBJ.mkdir(exist_ok=True)
    obj = OBJ / (src.stem + ".o")
    # compiled and linked in two steps like the library (a one-step -shared build
    # embeds the device code in a form torch's bundled HIP runtime does not load)
    for cmd in ([hipcc(), *HIPCC_FLAGS, f"-I{CSRC}", "-c", str(src), "-o", str(obj)],
                [hipcc(), "-shared", "--offload-arch=gfx950", "-fPIC", "-o", str(tmp), str(obj)]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_h5(verbose=True)
    build_checks(verbose=True)
