"""Build libcuda_toolkit.so in-tree for gfx950 with hipcc.

Every .hip/.cpp file under csrc/ is compiled to an object (in parallel) and
linked into one shared library next to this file.  No torch, no CMake.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libcuda_toolkit.so"
INCLUDE = HERE.parent / "include"
# out-of-process transfer protocol: server (links the library) and client (plain C)
TRANSFER_SERVER = HERE / "cut_transfer_server"
TRANSFER_CLIENT = HERE / "libcut_transfer.so"
ARCH = os.environ.get("CUT_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "-O3", "-fPIC", "-std=c++20", "-fvisibility=hidden", f"--offload-arch={ARCH}",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _compile(src: Path, incremental: bool = False) -> Path:
    obj = OBJ / (src.name + ".o")
    if incremental and obj.exists():
        newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in CSRC.glob("*.hpp")] +
                     [h.stat().st_mtime for h in (HERE.parent / "include").glob("*.h*")])
        if obj.stat().st_mtime > newest:
            return obj
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err:
        sys.stderr.write(err + "\n")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> Path:
    """Compile every source (incremental: only those newer than their object
    or than any header) and link."""
    OBJ.mkdir(exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda p: _compile(p, incremental), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_transfer()
    if verbose:
        print(f"built {LIB} from {len(sources)} sources, {TRANSFER_SERVER.name} and {TRANSFER_CLIENT.name}")
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stderr}")


def build_transfer() -> None:
    """The transfer server (C++, linked against the library, rpath $ORIGIN)
    and the client library (C, gcc, no GPU dependency)."""
    tdir = CSRC / "transfer"
    tmp = TRANSFER_SERVER.with_name(TRANSFER_SERVER.name + ".tmp")
    _run([HIPCC, "-O2", "-std=c++20", "-Wall", f"-I{INCLUDE}", f"-I{tdir}", str(tdir / "transfer_server.cpp"),
          "-o", str(tmp), f"-L{HERE}", "-lcuda_toolkit", "-Wl,-rpath,$ORIGIN"], "transfer server build")
    os.replace(tmp, TRANSFER_SERVER)
    gcc = shutil.which("gcc") or "gcc"
    tmp = TRANSFER_CLIENT.with_name(TRANSFER_CLIENT.name + ".tmp")
    _run([gcc, "-O2", "-std=c11", "-Wall", "-fPIC", "-shared", "-fvisibility=hidden", f"-I{INCLUDE}", f"-I{tdir}",
          str(tdir / "transfer_client.c"), "-o", str(tmp), "-lrt"], "transfer client build")
    os.replace(tmp, TRANSFER_CLIENT)


# ---------------------------------------------------------------------------
# Host-sanitizer build: the host-side sources (.cpp: API layers, validation,
# RF state, DAS constants, batch set-up, synthetic generator, transfer server
# and client) instrumented with AddressSanitizer and UndefinedBehaviorSanitizer
# (host code only: each -fsanitize= follows -Xarch_host, so no GPU code is
# instrumented), linked with the regular kernel objects into SAN_DIR.  The CPU
# suite runs its host-side tests against it (tests/test_host_sanitizers.py).
# ---------------------------------------------------------------------------
SAN_DIR = HERE / "_build_san"
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
CLANG = "/opt/rocm/llvm/bin/clang"


def sanitizer_runtime() -> Path | None:
    """The clang AddressSanitizer runtime the sanitized build links against."""
    hits = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def build_sanitized(verbose: bool = True) -> Path:
    rt = sanitizer_runtime()
    if rt is None:
        raise RuntimeError("no clang AddressSanitizer runtime under /opt/rocm/llvm")
    SAN_DIR.mkdir(exist_ok=True)
    hosts = sorted(CSRC.glob("*.cpp"))

    def comp(src: Path) -> Path:
        obj = SAN_DIR / (src.name + ".o")
        cmd = [HIPCC, "-O1", "-g", "-fPIC", "-std=c++20", "-fvisibility=hidden", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-x", "hip", *SAN_HOST,
               "-c", str(src), "-o", str(obj)]
        _run(cmd, f"sanitizer build of {src.name}")
        return obj

    with cf.ThreadPoolExecutor(min(len(hosts), 8)) as ex:
        objs = list(ex.map(comp, hosts))
    kernels = [OBJ / (p.name + ".o") for p in sorted(CSRC.glob("*.hip"))]
    for k in kernels:
        if not k.exists():
            raise RuntimeError(f"{k} missing: build() first")
    lib = SAN_DIR / "libcuda_toolkit.so"
    rpath = f"-Wl,-rpath,{rt.parent}"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *SAN_HOST, "-shared-libsan", rpath, "-o", str(lib),
          *map(str, objs), *map(str, kernels), "-lpthread"], "sanitizer link")
    tdir = CSRC / "transfer"
    _run([HIPCC, "-O1", "-g", "-std=c++20", "-Wall", f"-I{INCLUDE}", f"-I{tdir}", *SAN_HOST, "-shared-libsan", rpath,
          str(tdir / "transfer_server.cpp"), "-o", str(SAN_DIR / "cut_transfer_server"), f"-L{SAN_DIR}",
          "-lcuda_toolkit", "-Wl,-rpath,$ORIGIN"], "sanitizer build of the transfer server")
    _run([CLANG, "-O1", "-g", "-std=c11", "-Wall", "-fPIC", "-shared", "-fvisibility=hidden", f"-I{INCLUDE}",
          f"-I{tdir}", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-shared-libsan", rpath,
          str(tdir / "transfer_client.c"), "-o", str(SAN_DIR / "libcut_transfer.so"), "-lrt"],
         "sanitizer build of the transfer client")
    if verbose:
        print(f"built host-sanitizer variants in {SAN_DIR}")
    return lib


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    if sanitizer_runtime() is not None and "--no-sanitize" not in sys.argv:
        build_sanitized()
