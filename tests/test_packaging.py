"""The installed library under the names AMUSE's build looks for.

AMUSE's configure (support/shared/m4/amuse_lib.m4:17-45 AMUSE_LIB, with
AMUSE_LIB_G6LIB :126-128 and AMUSE_LIB_SAPPORO_LIGHT :135-136) first runs
AC_SEARCH_LIBS -- a C program declaring `char f ();` and calling it, linked
with -l<lib> -- and, when that fails, reads the pkg-config module.  Here
`python -m amuse_amd.build --install PREFIX` is followed by both probes for
both modules: g6_npipes in -lg6 / module g6lib (lib/g6lib/Makefile:5-27),
get_device_count in -lsapporo / module sapporo_light
(lib/sapporo_light/Makefile:5-27).  The probe programs are also run: without
a GPU get_device_count answers 0 and g6_npipes the pipeline count; with one,
the device count.  A program linked through the sapporo name records
libsapporo.so and loads it through the installed link.
"""
import os
import re
import subprocess

import pytest

from amuse_amd import build

PROBES = [("g6lib", "g6", "g6_npipes"), ("sapporo_light", "sapporo", "get_device_count")]


def _pc_flags(path):
    """Cflags and Libs of a .pc file, ${var} expanded (no pkg-config here)."""
    var, out = {}, {}
    for line in open(path).read().splitlines():
        m = re.match(r"^([A-Za-z_]+)=(.*)$", line)
        if m:
            var[m.group(1)] = m.group(2)
            continue
        m = re.match(r"^([A-Za-z.]+):\s*(.*)$", line)
        if m:
            out[m.group(1)] = m.group(2)
    expand = lambda s: re.sub(r"\$\{(\w+)\}", lambda k: expand(var[k.group(1)]), s)
    return expand(out.get("Cflags", "")).split(), expand(out["Libs"]).split(), out


@pytest.fixture(scope="module")
def prefix(tmp_path_factory):
    if not os.path.exists(os.path.join(build.LIBDIR, "libg6.so")):
        pytest.fail("amuse_amd/lib/libg6.so is not built")
    p = tmp_path_factory.mktemp("prefix")
    build.install(str(p))
    return str(p)


def _probe_source(tmp, fn):
    c = os.path.join(tmp, f"conftest_{fn}.c")
    with open(c, "w") as f:   # what AC_LANG_CALL writes for a link check
        f.write(f"char {fn} ();\nint main (void) {{ {fn} (); return 0; }}\n")
    return c


@pytest.mark.parametrize("module, lib, fn", PROBES)
def test_search_libs_probe(prefix, tmp_path, module, lib, fn):
    exe = str(tmp_path / "probe")
    r = subprocess.run(["gcc", _probe_source(str(tmp_path), fn), f"-L{prefix}/lib", f"-l{lib}",
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    needed = subprocess.run(["readelf", "-d", exe], capture_output=True, text=True).stdout
    assert f"[lib{lib}.so]" in needed
    r = subprocess.run([exe], env=dict(os.environ, LD_LIBRARY_PATH=f"{prefix}/lib"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("module, lib, fn", PROBES)
def test_pkg_config_module(prefix, tmp_path, module, lib, fn):
    pc = os.path.join(prefix, "lib", "pkgconfig", module + ".pc")
    cflags, libs, fields = _pc_flags(pc)
    assert fields["Name"] == module
    assert libs == [f"-L{prefix}/lib", f"-l{lib}"]
    assert cflags == [f"-I{prefix}/include"]
    src = tmp_path / "uses_header.c"
    # a consumer that includes the installed header and calls through it
    src.write_text('#include "g6lib.h"\n#include "g6x.h"\n'
                   "int main (void) { return g6_npipes_() > 0 ? 0 : 1; }\n")
    exe = str(tmp_path / "uses_header")
    r = subprocess.run(["gcc", str(src), *cflags, *libs, "-o", exe], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], env=dict(os.environ, LD_LIBRARY_PATH=f"{prefix}/lib"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_sapporo_name_is_the_same_library(prefix):
    lib = os.path.join(prefix, "lib")
    assert os.path.islink(os.path.join(lib, "libsapporo.so"))
    assert os.path.realpath(os.path.join(lib, "libsapporo.so")) == \
        os.path.realpath(os.path.join(lib, "libg6.so"))
    dyn = subprocess.run(["readelf", "-d", os.path.join(lib, "libg6.so")], capture_output=True,
                         text=True).stdout
    assert "SONAME" not in dyn   # so -lsapporo records libsapporo.so
