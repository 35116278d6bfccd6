"""Builds libmedusa_amd.so (gfx950) in-tree with hipcc.

Every .hip and .cpp file under csrc/ is compiled to an object in parallel and
linked into medusa_amd/libmedusa_amd.so.  Run as a script or call build().
"""
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
ROOT = HERE.parent
OBJ = HERE / "_build"
LIB = HERE / "libmedusa_amd.so"

ARCH = os.environ.get("MEDUSA_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # FP32 results are defined bit-for-bit: no contraction into FMA except explicit fmaf
    "-ffp-contract=off",
    f"-I{ROOT / 'include'}",
]
# device code: gfx950 only, packed-FP32 instructions disabled
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: pathlib.Path, verbose: bool) -> pathlib.Path:
    obj = OBJ / (src.name + ".o")
    cmd = [hipcc(), *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip", *HIP_FLAGS]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


APPS = HERE / "apps"
APP_BIN = APPS / "bin"


# Extra builds of an application with the reference's compile-time layout choice
# (Algorithm/Configuration.h: one of MEG, AA, HY, ELL is defined): name -> (source app, defines).
LAYOUT_APPS = {"bfs_meg": ("bfs", ["-DMEG"]), "bfs_hy": ("bfs", ["-DHY"]), "bfs_ell": ("bfs", ["-DELL"])}


def _build_app(name: str, verbose: bool) -> pathlib.Path:
    """Compiles apps/<name>/main.cpp (Medusa's user API, header-only) against the library."""
    src, defines = LAYOUT_APPS.get(name, (name, []))
    out = APP_BIN / f"{name}_app"
    cmd = [hipcc(), "-x", "hip", *HIP_FLAGS, *COMMON, *defines, str(APPS / src / "main.cpp"), "-o", str(out),
           f"-L{HERE}", "-lmedusa_amd", "-Wl,-rpath,$ORIGIN/../.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"app build failed: {name}\n{r.stdout}\n{r.stderr}")
    return out


def build_apps(verbose: bool = False):
    APP_BIN.mkdir(exist_ok=True)
    names = sorted(d.name for d in APPS.iterdir() if (d / "main.cpp").exists()) + sorted(LAYOUT_APPS)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        return list(ex.map(lambda n: _build_app(n, verbose), names))


def build(verbose: bool = False, apps: bool = True) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if apps:
        build_apps(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
