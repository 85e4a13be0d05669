"""Builds the native libraries of amuse_amd in-tree with hipcc (gfx950 only).

    python -m amuse_amd.build          # libg6.so, libhermite4.so, g6_worker
    python -m amuse_amd.build --isa    # also dump the gfx950 assembly to _build/
    python -m amuse_amd.build --install PREFIX
        # copies the GRAPE-6 library where AMUSE's build looks for the reference
        # ones: PREFIX/lib/libg6.so (g6lib's name) and libsapporo.so
        # (sapporo_light's name, a link to the same file), PREFIX/include/g6lib.h
        # and g6x.h, PREFIX/lib/pkgconfig/g6lib.pc and sapporo_light.pc

The libraries land in amuse_amd/lib/ so they travel with the source tree; no
JIT cache and no site-packages install is involved.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
ARCH = "gfx950"

# Packed FP32 instructions are kept out of every code object (the compiler
# would otherwise form v_pk_* from float2-shaped code).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIBS = {
    "libg6.so": ["g6_kernels.hip", "g6_abi.hip", "g6_clusters.hip", "g6_adjoint.hip",
                 "g6_blocks.hip", "g6_spatial.hip"],
}
# host-only C++ (no device code): built with g++
CPU_LIBS = {
    "libhermite4.so": (["hermite4.cpp"], ["-ldl"]),
}
# host-only executables linked against libg6.so (found next to them at run time)
EXES = {
    "g6_worker": ["g6_worker.cpp"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_lib(name: str, isa: bool = False, defines=(), outdir: str = LIBDIR) -> str:
    os.makedirs(outdir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in LIBS[name]]
    out = os.path.join(outdir, name)
    tmp = out + ".tmp"
    cmd = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
           "-fvisibility=hidden", "-Wno-unused-command-line-argument",
           *DEVICE_FLAGS, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include"),
           *srcs, "-Wl,-Bsymbolic", "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    if isa:
        bdir = os.path.join(ROOT, "_build")
        os.makedirs(bdir, exist_ok=True)
        for s in srcs:
            base = os.path.splitext(os.path.basename(s))[0]
            _run([hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "--cuda-device-only",
                  "-S", *DEVICE_FLAGS, "-I", os.path.join(ROOT, "include"), s,
                  "-o", os.path.join(bdir, base + ".s")])
    return out


def build_cpu_lib(name: str) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs, libs = CPU_LIBS[name]
    out = os.path.join(LIBDIR, name)
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall",
          *[os.path.join(CSRC, s) for s in srcs], *libs, "-o", tmp])
    os.replace(tmp, out)
    return out


def rocm_libdir() -> str:
    return os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "lib")


def build_exe(name: str) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, name)
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          *[os.path.join(CSRC, s) for s in EXES[name]], "-L", LIBDIR, "-lg6",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + rocm_libdir(), "-o", tmp])
    os.replace(tmp, out)
    return out


# pkg-config modules as lib/g6lib/Makefile:12-27 and lib/sapporo_light/Makefile:12-27
# write them at install time; AMUSE's configure probes g6_npipes in -lg6 and
# get_device_count in -lsapporo first and falls back to these modules
# (support/shared/m4/amuse_lib.m4:17-45,126-136).
PKG_MODULES = {
    "g6lib": ("GRAPE-6 ABI on AMD Instinct MI355X (HIP, gfx950)", "g6"),
    "sapporo_light": ("GRAPE-6 ABI on AMD Instinct MI355X, under sapporo_light's name", "sapporo"),
}
HEADERS = ["g6lib.h", "g6x.h"]


def install(prefix: str) -> list[str]:
    """Installs the built libg6.so under the names the reference libraries
    use.  libsapporo.so is a symbolic link: the library has no SONAME, so a
    program linked with -lsapporo records libsapporo.so and finds it through
    the link."""
    src = os.path.join(LIBDIR, "libg6.so")
    if not os.path.exists(src):
        raise RuntimeError(f"{src} is not built; run amuse_amd.build")
    prefix = os.path.abspath(prefix)
    lib, inc, pc = (os.path.join(prefix, d) for d in ("lib", "include", "lib/pkgconfig"))
    for d in (lib, inc, pc):
        os.makedirs(d, exist_ok=True)
    out = []
    dst = os.path.join(lib, "libg6.so")
    shutil.copy2(src, dst + ".tmp")
    os.replace(dst + ".tmp", dst)
    out.append(dst)
    link = os.path.join(lib, "libsapporo.so")
    if os.path.lexists(link):
        os.remove(link)
    os.symlink("libg6.so", link)
    out.append(link)
    for h in HEADERS:
        shutil.copy2(os.path.join(ROOT, "include", h), os.path.join(inc, h))
        out.append(os.path.join(inc, h))
    for name, (desc, libname) in PKG_MODULES.items():
        path = os.path.join(pc, name + ".pc")
        with open(path, "w") as f:
            f.write(f"prefix={prefix}\nincludedir=${{prefix}}/include\nlibdir=${{prefix}}/lib\n\n"
                    f"Name: {name}\nDescription: {desc}\nVersion: 1.0\n"
                    f"Cflags: -I${{includedir}}\nLibs: -L${{libdir}} -l{libname}\n")
        out.append(path)
    return out


def build_all(isa: bool = False) -> list[str]:
    return ([build_lib(n, isa) for n in LIBS] + [build_cpu_lib(n) for n in CPU_LIBS]
            + [build_exe(n) for n in EXES])


if __name__ == "__main__":
    if "--install" in sys.argv:
        at = sys.argv.index("--install")
        if at + 1 >= len(sys.argv):
            sys.exit("usage: python -m amuse_amd.build --install PREFIX")
        for p in install(sys.argv[at + 1]):
            print("installed", p)
    else:
        for p in build_all(isa="--isa" in sys.argv):
            print("built", p)
