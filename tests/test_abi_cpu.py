"""libg6.so on the CPU: it loads, exports every entry point the headers
declare, reports its configuration, and fails loudly (non-zero return,
message) when no device is present.  No kernel is launched here."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "starlab_amd", "libg6.so")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(g6\w*)\s*\(", text)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libg6.so not built (python -m starlab_amd.build)")
    import starlab_amd
    return starlab_amd.load()


def test_grape6_header_lists_all_19_reference_entry_points():
    names = _declared("grape6.h")
    assert len(names) == 19
    # the reference's full set (include/grape6.h:18-102 of the Starlab tree)
    assert names == sorted([
        "g6_set_tunit_", "g6_set_xunit_", "g6_open_", "g6_close_", "g6_open_special",
        "g6_open_special2", "g6_reset_", "g6_reset_fofpga_", "g6_npipes_", "g6_set_ti_",
        "g6_set_j_particle_", "g6calc_firsthalf_", "g6calc_lasthalf_", "g6calc_lasthalf2_",
        "g6_read_neighbour_list_", "g6_read_neighbour_list_old_", "g6_get_neighbour_list_",
        "g6_initialize_jp_buffer_", "g6_flush_jp_buffer_"])


def test_every_declared_symbol_is_exported(lib):
    exported = _exported()
    for h in ("grape6.h", "g6x.h", "g6ens.h"):
        missing = [n for n in _declared(h) if n not in exported]
        assert not missing, (h, missing)


def test_ctypes_table_covers_headers(lib):
    from starlab_amd import _abi
    assert sorted(_abi.GRAPE6) == _declared("grape6.h")
    assert sorted(_abi.G6X) == _declared("g6x.h")
    assert sorted(_abi.G6E) == _declared("g6ens.h")


def _run(code, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          env=e, cwd=ROOT, timeout=120)


def test_npipes_default_and_env():
    code = "import starlab_amd; L = starlab_amd.load(); print(L.g6_npipes_())"
    r = _run(code)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout.strip()) == 2048
    r = _run(code, {"G6_NPIPES": "512"})
    assert int(r.stdout.strip()) == 512


def test_options_roundtrip(lib):
    from starlab_amd.g6 import GRAPE6
    g = GRAPE6()
    old = g.option("nb_cap")
    g.option("nb_cap", 123)
    assert g.option("nb_cap") == 123
    g.option("nb_cap", old)
    with pytest.raises(KeyError):
        g.option("no_such_option", 1)


def test_fails_loudly_without_device():
    """Without a GPU, open fails and every compute call returns an error
    (there is no host fallback)."""
    code = r"""
import torch, sys
if torch.cuda.is_available():
    print("HAVE_GPU"); sys.exit(0)
import numpy as np
from starlab_amd.g6 import GRAPE6
g = GRAPE6()
rc_open = g.open()
rc_set = g.set_j_particle(0, 0, 0.0, 1.0, 1.0, np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3))
g.calc_firsthalf(1, [0], np.zeros((1, 3)), np.zeros((1, 3)), 0.0, [0.0])
rc, *_ = g.calc_lasthalf2(1, 1)
print(rc_open, rc, g.self_sweep(1, 0, 1, 0.0, 0.0), g.read_neighbour_list())
"""
    r = _run(code)
    assert r.returncode == 0, r.stderr
    if "HAVE_GPU" in r.stdout:
        pytest.skip("a GPU is present")
    rc_open, rc_last, rc_sweep, rc_nl = map(int, r.stdout.split())
    assert rc_open != 0 and rc_last != 0 and rc_sweep != 0 and rc_nl != 0
    assert "libg6:" in r.stderr


def test_missing_library_raises(tmp_path, monkeypatch):
    import starlab_amd
    monkeypatch.setattr(starlab_amd, "LIBG6", str(tmp_path / "nope.so"))
    monkeypatch.setattr(starlab_amd, "_lib", None)
    with pytest.raises(starlab_amd.LibraryMissing):
        starlab_amd.load()


def test_plain_c_caller_links_and_runs(tmp_path):
    """A C program written against include/grape6.h links -lg6 with no other
    library named, exactly as kira's configure probe does (AC_CHECK_LIB(g6,
    g6_open_), configure.ac:204-212 of the Starlab tree)."""
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "grape6.h"\n#include <stdio.h>\n'
        "int main(void){int c=0; printf(\"%d\\n\", g6_npipes_());"
        " int rc = g6_open_(&c); if (rc == 0) g6_close_(&c); printf(\"%d\\n\", rc); return 0;}\n")
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "starlab_amd")
    subprocess.run(["gcc", f"-I{os.path.join(ROOT, 'include')}", str(src), f"-L{libdir}",
                    f"-Wl,-rpath,{libdir}", "-lg6", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    lines = r.stdout.split()
    assert lines[0] == "2048"


def test_abi_bench_is_built_and_fails_loudly_without_device():
    """The compiled per-block ABI timing program links libg6.so and reports a
    failed g6_open_ (exit status 1) when no device is present."""
    exe = os.path.join(ROOT, "starlab_amd", "abi_bench")
    assert os.access(exe, os.X_OK), "build() makes starlab_amd/abi_bench"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, "1024", "2", "16"], capture_output=True, text=True, timeout=120)
    if r.returncode == 0:
        pytest.skip("a device is present")
    assert r.returncode == 1 and "g6_open_ failed" in r.stderr
