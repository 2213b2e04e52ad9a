"""Starlab snapshot text format (starlab_amd.snapshot): the layout of
put_node (src/node/util/tree_io.C:417-525) and dyn::print_dyn_story
(src/node/dyn/util/dyn_io.C:223-256), the reader's keyword and bracket
rules (src/std/util_io.C:94-160), round trips, and binary trees."""
import io

import numpy as np

from starlab_amd import ics, snapshot

# A two-body snapshot laid out exactly as put_node writes it (root with log
# and system_time; daughters with index, N, Dynamics, empty Hydro and Star).
TWO_BODY = """(Particle
  N = 2
(Log
  ===>  Wed Dec  3 00:00:00 2026
       makeplummer -n 2 -s 42
)Log
(Dynamics
  system_time  =  0
  m  =  1
  r  =  0 0 0
  v  =  0 0 0
)Dynamics
(Hydro
)Hydro
(Star
)Star
(Particle
  i = 1
  N = 1
(Dynamics
  m  =  0.5
  r  =  -0.5 0 0
  v  =  0 -0.5 0
)Dynamics
(Hydro
)Hydro
(Star
)Star
)Particle
(Particle
  i = 2
  N = 1
(Dynamics
  m  =  0.5
  r  =  0.5 0 0
  v  =  0 0.5 0
)Dynamics
(Hydro
)Hydro
(Star
)Star
)Particle
)Particle
"""


def test_reads_the_put_node_layout():
    root = snapshot.read_snapshot(TWO_BODY)
    assert root.system_time == 0.0 and root.mass == 1.0 and len(root.daughters) == 2
    assert root.log[1].strip() == "makeplummer -n 2 -s 42"
    m, x, v = snapshot.top_level(root)
    np.testing.assert_array_equal(m, [0.5, 0.5])
    np.testing.assert_array_equal(x, [[-0.5, 0, 0], [0.5, 0, 0]])
    np.testing.assert_array_equal(v, [[0, -0.5, 0], [0, 0.5, 0]])
    assert [d.index for d in root.daughters] == [1, 2]


def test_writer_reproduces_the_layout():
    root = snapshot.read_snapshot(TWO_BODY)
    assert snapshot.write_snapshot(None, root) == TWO_BODY


def test_short_brackets_and_unknown_story_lines():
    short = TWO_BODY.replace("(Particle", "(P").replace(")Particle", ")P") \
        .replace("(Dynamics", "(D").replace(")Dynamics", ")D")
    short = short.replace("  m  =  0.5\n  r  =  0.5", "  m  =  0.5\n  t  =  0.125\n  r  =  0.5")
    root = snapshot.read_snapshot(short)
    assert root.daughters[1].dyn_extra == ["  t  =  0.125"]
    m, x, v = snapshot.top_level(root)
    assert x[1, 0] == 0.5


def test_round_trip_is_exact_and_streams_concatenate():
    m, x, v = ics.makeplummer(100, 3, scale=False)
    root = snapshot.from_arrays(m, x, v, system_time=0.375)
    buf = io.StringIO()
    snapshot.write_snapshot(buf, root)
    snapshot.write_snapshot(buf, snapshot.from_arrays(m[:5], x[:5], v[:5]))
    buf.seek(0)
    snaps = list(snapshot.read_snapshots(buf))
    assert len(snaps) == 2
    m2, x2, v2 = snapshot.top_level(snaps[0])
    np.testing.assert_array_equal(m2, m)
    np.testing.assert_array_equal(x2, x)      # %.18g round-trips doubles exactly
    np.testing.assert_array_equal(v2, v)
    assert snaps[0].system_time == 0.375 and snaps[1].n_leaves() == 5


def test_binary_trees_from_the_cluster_pipeline():
    c = ics.king_cluster(400, 5)
    root = snapshot.from_arrays(c["m"], c["x"], c["v"], binaries=c)
    text = snapshot.write_snapshot(None, root)
    back = snapshot.read_snapshot(text)
    assert back.n_leaves() == 400 + len(c["bin_index"])
    mt, xt, vt = snapshot.top_level(back)
    np.testing.assert_array_equal(mt, c["m"])
    ml, xl, vl, ids = snapshot.leaves(back)
    assert abs(ml.sum() - 1) < 1e-12
    # components sit about their centre of mass
    for b, k in enumerate(c["bin_index"][:20]):
        cm = back.daughters[k]
        p, s = cm.daughters
        np.testing.assert_allclose(p.mass * p.pos + s.mass * s.pos, 0, atol=1e-18)
        np.testing.assert_allclose(s.pos - p.pos, c["dx"][b], rtol=1e-12)
        assert cm.name == f"({p.index},{s.index})"
