"""libg4x.so (the GRAPE-4 ABI) and its oracle on the CPU.

- the library loads, exports every entry point include/grape4.h and
  include/g4x.h declare, and not the GRAPE-6 probe symbol (Starlab's configure
  selects GRAPE-4 only when the named library has no g6_open_,
  configure.ac:204-221);
- a plain C caller links -lg4x alone;
- without a device a GRAPE-4 call prints "libg4x:" and exits (status 1), or,
  with G4_EXIT_ON_ERROR=0, returns and counts the failure;
- the GRAPE-4 oracle restatement (oracle.g4_force) is pinned bit for bit to the
  reference's own IEEE GRAPE-4 simulator (oracle/_ref, compiled from
  src/harpsims/harp3fastsim.c), and its pot (which the simulator does not
  compute) to closed forms.
No kernel is launched here.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBG4 = os.path.join(ROOT, "starlab_amd", "libg4x.so")
LIBG6 = os.path.join(ROOT, "starlab_amd", "libg6.so")

REFERENCE_GRAPE4 = sorted([
    "h3open_", "h3close_", "h3npipe_", "h3setnboards_", "h3getnboards_", "h3wait_",
    "h3setti_", "h3setmode_", "h3setled_", "h3jpmax_", "h3jpdma_indirect_",
    "h3mjpdma_indirect_", "h3mjpdma_start_", "h3mjpdma_flush_", "h3calc_",
    "h3calc_firsthalf_", "h3calc_lasthalf_", "h3nbread_", "h3nblist_", "count_nblist_low",
    "h3setdebuglevel_", "h3checkdmabuffers_", "set_time_check_mode"])


def _declared(header, pattern=r"\b(\w+)\s*\("):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(n for n in re.findall(pattern, text) if n not in ("extern",)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def _run(code, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          env=e, cwd=ROOT, timeout=120)


def test_grape4_header_lists_the_reference_entry_points():
    # the reference's full set, include/grape4.h:22-45 of the Starlab tree
    assert _declared("grape4.h") == REFERENCE_GRAPE4


def test_exports_and_configure_probe_semantics():
    if not os.path.exists(LIBG4):
        pytest.fail("libg4x.so not built (python -m starlab_amd.build)")
    g4 = _exported(LIBG4)
    for h in ("grape4.h", "g4x.h"):
        missing = [n for n in _declared(h) if n not in g4]
        assert not missing, (h, missing)
    # GRAPE-6 and GRAPE-4 live in different libraries, so configure's
    # AC_CHECK_LIB(..., g6_open_) fails on libg4x and succeeds on libg6
    assert "g6_open_" not in g4 and not any(n.startswith("g6") for n in g4)
    g6 = _exported(LIBG6)
    assert "g6_open_" in g6 and "h3open_" not in g6
    # nothing but the declared C entry points is exported as code
    assert g4 == set(_declared("grape4.h")) | set(_declared("g4x.h"))


def test_ctypes_table_covers_headers():
    from starlab_amd import _abi
    assert sorted(_abi.GRAPE4) == _declared("grape4.h")
    assert sorted(_abi.G4X) == _declared("g4x.h")


def test_npipe_jpmax_defaults_and_env():
    code = ("import starlab_amd; L = starlab_amd.load_g4x(); "
            "print(L.h3npipe_(), L.h3jpmax_(), L.h3getnboards_(), L.h3wait_())")
    r = _run(code)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["2048", "4096", "1", "0"]
    r = _run(code, {"G4_NPIPES": "96", "G4_JPMAX": "26"})
    assert r.stdout.split()[:2] == ["96", "26"]
    # the GRAPE-6 knob does not reach the GRAPE-4 library
    r = _run(code, {"G6_NPIPES": "7"})
    assert r.stdout.split()[0] == "2048"


def test_options_roundtrip():
    from starlab_amd.g4 import GRAPE4
    g = GRAPE4(exit_on_error=False)
    old = g.option("nb_cap")
    g.option("nb_cap", 17)
    assert g.option("nb_cap") == 17
    g.option("nb_cap", old)
    assert g.option("exit_on_error") == 0
    with pytest.raises(KeyError):
        g.option("no_such_option", 1)


def test_exits_loudly_without_device():
    code = r"""
import torch, sys
if torch.cuda.is_available():
    print("HAVE_GPU"); sys.exit(0)
import starlab_amd
L = starlab_amd.load_g4x()
L.h3open_()
print("survived")
"""
    r = _run(code)
    if "HAVE_GPU" in r.stdout:
        pytest.skip("a GPU is present")
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "survived" not in r.stdout
    assert "libg4x:" in r.stderr


def test_counts_failures_when_asked_not_to_exit():
    code = r"""
import torch, sys
if torch.cuda.is_available():
    print("HAVE_GPU"); sys.exit(0)
import numpy as np
from starlab_amd.g4 import GRAPE4
g = GRAPE4(exit_on_error=False)
g.open()
g.setti(0.0)
g.calc_firsthalf(4, np.zeros((2, 3)), np.zeros((2, 3)), 0.0, 0.0)
g.calc_lasthalf(2)
g.nbread()
print(g.errors(), g.last_error())
"""
    r = _run(code)
    assert r.returncode == 0, r.stderr
    if "HAVE_GPU" in r.stdout:
        pytest.skip("a GPU is present")
    n = int(r.stdout.split()[0])
    assert n == 4          # open, firsthalf, lasthalf, nbread
    assert "libg4x:" in r.stderr


def test_plain_c_caller_links_and_runs(tmp_path):
    """A C program written against include/grape4.h links -lg4x with no other
    library named, as configure's AC_CHECK_LIB(g4x, h3open_) does."""
    src = tmp_path / "caller.c"
    src.write_text('#include "grape4.h"\n#include <stdio.h>\n'
                   'int main(void){ printf("%d %d\\n", h3npipe_(), h3jpmax_()); return 0; }\n')
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "starlab_amd")
    subprocess.run(["gcc", f"-I{os.path.join(ROOT, 'include')}", str(src), f"-L{libdir}",
                    f"-Wl,-rpath,{libdir}", "-lg4x", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert r.stdout.split() == ["2048", "4096"]


# --- the GRAPE-4 oracle -----------------------------------------------------

@pytest.mark.skipif(not oracle.Harp3Sim.available(), reason="reference tree absent at build time")
@pytest.mark.parametrize("eps2", [0.0, 1e-4])
def test_g4_oracle_bitwise_vs_reference_simulator(eps2):
    """acc, jerk and neighbour lists of oracle.g4_force equal the reference
    simulator's bit for bit, the i-particles themselves in j-memory (their
    r^2 = 0 pair skipped when eps2 = 0, harp3fastsim.c:304)."""
    sim = oracle.Harp3Sim()
    n = 900
    m, x, v = oracle.makeplummer(n, 7)
    z = np.zeros((n, 3))
    sim.set_j(x, v, z, z, m, np.zeros(n))
    for sel in (np.array([10, 11, 12]), np.array([0, 450, 899])):
        h2 = np.array([0.01, 0.03, 0.0])
        # Harp3Sim.calc halves what it is given; the simulator doubles it back
        acc, jerk, lists = sim.calc(0.0, n, x[sel], v[sel], 2 * eps2, 2 * h2)
        a, j, p, c, l = oracle.g4_force(x[sel], v[sel], eps2, h2, x, v, m)
        np.testing.assert_array_equal(a, acc)
        np.testing.assert_array_equal(j, jerk)
        for k in range(3):
            np.testing.assert_array_equal(l[k, :c[k]], lists[k])
        assert c[2] == 0 and c[0] >= 1 and sel[0] in l[0, :c[0]]


def test_g4_oracle_pot_closed_forms():
    """pot (not computed by the simulator): two-body -m/r with the doubled
    softening; with the i in j-memory and eps = 0 its own pair is skipped, so
    pot equals the GRAPE-6 oracle's with index exclusion, and sum m pot / 2
    is the pairwise potential energy."""
    x = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    v = np.zeros((2, 3))
    m = np.array([0.25, 0.75])
    r = np.sqrt(0.3 ** 2 + 0.4 ** 2 + 1.2 ** 2)
    for e in (0.0, 0.01):
        _, _, p, _, _ = oracle.g4_force(x[:1], v[:1], e, 0.0, x, v, m)
        want = -0.75 / np.sqrt(r * r + 2 * e) - (0.25 / np.sqrt(2 * e) if e > 0 else 0.0)
        assert abs(p[0] - want) <= 1e-15 * abs(want)
    n = 500
    m, x, v = oracle.makeplummer(n, 11)
    idx = np.arange(n, dtype=np.int32)
    _, _, p4, _, _ = oracle.g4_force(x, v, 0.0, 0.0, x, v, m)
    _, _, p6, _, _ = oracle.force(x, v, idx, x, v, m, idx, 0.0)
    np.testing.assert_allclose(p4, p6, rtol=1e-13)
    ep, _ = oracle.energies(m, x, v, 0.0)
    assert abs(0.5 * (m * p4).sum() - ep) <= 1e-12 * abs(ep)


def test_g4_oracle_lists_vs_brute_force():
    rng = np.random.default_rng(4)
    n = 300
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    eps2 = rng.uniform(0, 0.01, 20)
    h2 = rng.uniform(0, 0.2, 20)
    _, _, _, c, l = oracle.g4_force(x[:20], v[:20], eps2, h2, x, v, m, cap=n)
    for k in range(20):
        d = x - x[k]
        r2 = 2 * eps2[k] + (d * d).sum(1)
        want = np.nonzero(r2 < 2 * h2[k])[0]
        np.testing.assert_array_equal(l[k, :c[k]], want)
