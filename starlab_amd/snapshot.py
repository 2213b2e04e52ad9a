"""Starlab snapshots: the nested text format written by put_node and read
by get_node (src/node/util/tree_io.C:417-525, src/node/dyn/util/dyn_io.C:148-256,
src/std/util_io.C:94-160).

A snapshot is one ``(Particle`` ... ``)Particle`` tree.  Each node holds,
in order: ``i = index`` (if any), ``name = ...`` (if any), ``N = leaves``,
an optional ``(Log`` story, a ``(Dynamics`` story with ``system_time``
(root only), ``m``, ``r`` and ``v`` (position and velocity relative to the
parent node), optional ``(Hydro`` and ``(Star`` stories, then its daughter
particles.  Numbers are printed with ``%.<precision>g`` (precision 18 by
default, STARLAB_PRECISION).  Brackets may be abbreviated to their first
letter (``(P``, ``)D``), and lines inside a story that are not understood
are kept verbatim.

Binary stars are a centre-of-mass node with two leaf daughters, which is
also the tree kira integrates; the GRAPE j-memory holds the top-level
nodes.
"""
from __future__ import annotations

import io
from dataclasses import dataclass, field
from typing import Iterator, List, Optional, TextIO

import numpy as np

PRECISION = 18


@dataclass
class Node:
    mass: float = 0.0
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))   # relative to parent
    vel: np.ndarray = field(default_factory=lambda: np.zeros(3))   # relative to parent
    index: Optional[int] = None
    name: Optional[str] = None
    system_time: Optional[float] = None                           # root only
    log: List[str] = field(default_factory=list)
    dyn_extra: List[str] = field(default_factory=list)
    hydro: List[str] = field(default_factory=list)
    star: List[str] = field(default_factory=list)
    daughters: List["Node"] = field(default_factory=list)

    def n_leaves(self) -> int:
        return 1 if not self.daughters else sum(d.n_leaves() for d in self.daughters)


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
def _is_open(line: str, token: str) -> bool:
    """matchbracket (util_io.C:94-110): the first two characters decide."""
    s = line.lstrip()
    return len(s) >= 2 and s[0] == "(" and s[1] == token[0]


def _is_close(line: str, token: str) -> bool:
    s = line.lstrip()
    return len(s) >= 2 and s[0] == ")" and s[1] == token[0]


def _keyword(line: str):
    """getequals (util_io.C:111-136): keyword = leading [A-Za-z0-9_]*,
    value = text after the first '='."""
    s = line.lstrip()
    k = 0
    while k < len(s) and (s[k].isalnum() or s[k] == "_"):
        k += 1
    eq = s.find("=", k)
    if eq < 0:
        return s[:k], None
    return s[:k], s[eq + 1:].strip()


def _vector(val: str) -> np.ndarray:
    parts = val.split()
    if len(parts) < 3:
        raise ValueError(f"snapshot: expected 3 components, got {val!r}")
    return np.array([float(p) for p in parts[:3]])


def _read_story(lines, end: str) -> List[str]:
    out = []
    for line in lines:
        if _is_close(line, end):
            return out
        out.append(line)
    raise ValueError(f"snapshot: missing ){end}")


def _read_node(lines) -> Node:
    b = Node()
    last_real_time = False
    for line in lines:
        if not line.strip():
            continue
        if _is_close(line, "Particle"):
            return b
        if _is_open(line, "Particle"):
            b.daughters.append(_read_node(lines))
        elif _is_open(line, "Log"):
            b.log = _read_story(lines, "Log")
        elif _is_open(line, "Dynamics"):
            for dl in lines:
                if _is_close(dl, "Dynamics"):
                    break
                key, val = _keyword(dl)
                if val is None:
                    b.dyn_extra.append(dl)
                elif key == "real_system_time":
                    b.system_time = float(val.split()[0])
                    last_real_time = True
                elif key == "system_time":
                    try:
                        b.system_time = float(val.split()[0])
                    except ValueError:
                        if not last_real_time:
                            raise
                elif key == "m":
                    b.mass = float(val.split()[0])
                elif key == "r":
                    b.pos = _vector(val)
                elif key == "v":
                    b.vel = _vector(val)
                else:
                    b.dyn_extra.append(dl)
            else:
                raise ValueError("snapshot: missing )Dynamics")
        elif _is_open(line, "Hydro"):
            b.hydro = _read_story(lines, "Hydro")
        elif _is_open(line, "Star"):
            b.star = _read_story(lines, "Star")
        else:
            key, val = _keyword(line)
            if key == "i" and val is not None:
                b.index = int(val.split()[0])
            elif key == "name" and val is not None:
                b.name = val
            # N = ... is derived from the tree
    raise ValueError("snapshot: missing )Particle")


def read_snapshots(f: TextIO) -> Iterator[Node]:
    """Every snapshot in a stream (kira and the tools write one after
    another)."""
    lines = (ln.rstrip("\n") for ln in f)
    for line in lines:
        if _is_open(line, "Particle"):
            yield _read_node(lines)
        elif line.strip():
            raise ValueError(f"snapshot: expected (Particle, got {line!r}")


def read_snapshot(f) -> Node:
    """The first snapshot of a stream, a path or a string."""
    if isinstance(f, str) and f.lstrip().startswith("("):
        f = io.StringIO(f)
    if isinstance(f, str):
        with open(f) as fh:
            return next(read_snapshots(fh))
    return next(read_snapshots(f))


# ---------------------------------------------------------------------------
# writing
# ---------------------------------------------------------------------------
def _num(x: float, p: int) -> str:
    return "%.*g" % (p, x)


def _write_node(b: Node, out: List[str], p: int, root: bool):
    out.append("(Particle")
    if b.index is not None and b.index >= 0:
        out.append(f"  i = {b.index}")
    if b.name is not None:
        out.append(f"  name = {b.name}")
    out.append(f"  N = {b.n_leaves()}")
    if root or b.log:
        out.append("(Log")
        out.extend(b.log)
        out.append(")Log")
    out.append("(Dynamics")
    if root:
        out.append("  system_time  =  " + _num(b.system_time or 0.0, p))
    out.append("  m  =  " + _num(b.mass, p))
    out.append("  r  =  " + " ".join(_num(c, p) for c in b.pos))
    out.append("  v  =  " + " ".join(_num(c, p) for c in b.vel))
    out.extend(b.dyn_extra)
    out.append(")Dynamics")
    out.append("(Hydro")
    out.extend(b.hydro)
    out.append(")Hydro")
    out.append("(Star")
    out.extend(b.star)
    out.append(")Star")
    for d in b.daughters:
        _write_node(d, out, p, False)
    out.append(")Particle")


def write_snapshot(f, root: Node, precision: int = PRECISION):
    out: List[str] = []
    _write_node(root, out, precision, True)
    text = "\n".join(out) + "\n"
    if f is None:
        return text
    if isinstance(f, str):
        with open(f, "w") as fh:
            fh.write(text)
    else:
        f.write(text)
    return text


# ---------------------------------------------------------------------------
# arrays <-> trees
# ---------------------------------------------------------------------------
def from_arrays(m, x, v, system_time: float = 0.0, binaries=None, log=None) -> Node:
    """Root node with one daughter per top-level body (indices 1..n).
    binaries: optional dict with bin_index (top-level slots), m1, m2, dx, dv
    (secondary relative to primary): those slots become centre-of-mass nodes
    with two leaf daughters placed about the centre of mass."""
    m = np.asarray(m, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    root = Node(mass=float(m.sum()), system_time=system_time, log=list(log or []))
    n = len(m)
    for k in range(n):
        root.daughters.append(Node(mass=float(m[k]), pos=x[k].copy(), vel=v[k].copy(),
                                   index=k + 1))
    if binaries is not None and len(binaries["bin_index"]):
        nmax = n
        for b, k in enumerate(binaries["bin_index"]):
            cm = root.daughters[k]
            m1, m2 = float(binaries["m1"][b]), float(binaries["m2"][b])
            mt = m1 + m2
            dx, dv = binaries["dx"][b], binaries["dv"][b]
            p = Node(mass=m1, pos=-(m2 / mt) * dx, vel=-(m2 / mt) * dv, index=cm.index)
            s = Node(mass=m2, pos=(m1 / mt) * dx, vel=(m1 / mt) * dv, index=cm.index + nmax)
            cm.index = None
            cm.name = f"({p.index},{s.index})"
            cm.daughters = [p, s]
    return root


def top_level(root: Node):
    """(m, x, v) of the root's daughters in absolute coordinates: the bodies
    kira keeps in GRAPE j-memory."""
    ds = root.daughters
    m = np.array([d.mass for d in ds])
    x = np.array([root.pos + d.pos for d in ds]).reshape(-1, 3)
    v = np.array([root.vel + d.vel for d in ds]).reshape(-1, 3)
    return m, x, v


def leaves(root: Node):
    """(m, x, v, index) of every leaf in absolute coordinates, tree order."""
    ms, xs, vs, ids = [], [], [], []

    def walk(b: Node, x0, v0):
        x = x0 + b.pos
        v = v0 + b.vel
        if not b.daughters:
            ms.append(b.mass); xs.append(x); vs.append(v)
            ids.append(-1 if b.index is None else b.index)
            return
        for d in b.daughters:
            walk(d, x, v)

    for d in root.daughters:
        walk(d, root.pos, root.vel)
    return (np.array(ms), np.array(xs).reshape(-1, 3), np.array(vs).reshape(-1, 3),
            np.array(ids, dtype=np.int64))
