"""Memory-experiment circuit generators: rotated surface codes and
bivariate-bicycle (BB) codes.

Same circuits as the reference's generator (include/GreenPeas/QEC/Codes/
Code.hpp:125-407, python/greenpeas/codes/__init__.py:19-31):

* qubits: data 0..n-1, then X-check ancillas, then Z-check ancillas
  (Code.hpp:93-119);
* head: reset data, X_ERROR(p) on data, DEPOLARIZE1(p/10) on checks;
* round: reset checks; X_ERROR(p) checks + DEPOLARIZE1(p/10) data; H on X
  checks; CNOT layers of the X then Z schedules, each followed by
  DEPOLARIZE2(p) on the pairs and DEPOLARIZE1(p/10) on idle qubits; H on X
  checks; M(p) checks (Code.hpp:298-326);
* detectors: Z checks of the first round alone, then every check against its
  previous round (X checks optional, Code.hpp:333-356);
* tail: M(p) data, Z-check detectors from data parities, one
  OBSERVABLE_INCLUDE per Z logical (Code.hpp:361-394).

A code is a pair of parity-check matrices whose entries are CNOT layers
(``Hx``, ``Hz``) plus logical supports (``Lx``, ``Lz``).  BB codes load the
reference's matrices (``data/codes/bb/dN/*.mtx``); rotated surface codes of
any distance are constructed here (rotated plaquette layout, N-shaped X and
Z-shaped Z CNOT order) and match the reference's d3..d14 matrices exactly.

Extras used by the batched and serving workloads: per-instruction noise
rates drawn from a seeded generator, and seeded "reset-conditioned branches"
that follow a measurement with an extra reset layer on that ancilla.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def fmt_p(p: float) -> str:
    """Probability as stim prints it (shortest round-trip decimal)."""
    r = repr(float(p))
    return r[:-2] if r.endswith(".0") else r


@dataclass
class SparseCode:
    """Check matrices as (row, col, layer) triplets in CNOT-layer order."""

    num_data: int
    hx: list[tuple[int, int, int]]
    hz: list[tuple[int, int, int]]
    num_x: int
    num_z: int
    lx: dict[int, list[int]]
    lz: dict[int, list[int]]


def read_mtx(path: str) -> tuple[int, int, list[tuple[int, int, int]]]:
    """Matrix Market coordinate file -> (rows, cols, [(row, col, value)]),
    0-based, in file order (SparseMatrix::fromMtx, SparseMatrix.hpp:271-307)."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln and not ln.startswith("%")]
    nr, nc, nnz = (int(x) for x in lines[0].split())
    out = []
    for ln in lines[1 : 1 + nnz]:
        r, c, v = ln.split()
        out.append((int(r) - 1, int(c) - 1, int(v)))
    return nr, nc, out


def write_mtx(path: str, nrows: int, ncols: int, entries) -> None:
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n%\n")
        f.write(f"{nrows} {ncols} {len(entries)}\n")
        for r, c, v in entries:
            f.write(f"{r + 1} {c + 1} {v}\n")


def rotated_surface_code(d: int) -> SparseCode:
    """Rotated surface code of distance d (d >= 2).

    Data qubit (r, c) of the d x d grid is index r*d + c.  A plaquette with
    top-left corner (r0, c0), r0, c0 in [-1, d-1], is an X check when
    r0 + c0 is even and a Z check when odd; weight-2 X checks sit on the
    left/right edges, weight-2 Z checks on the top/bottom edges.  CNOT layer
    k touches corner offset X: (1,1),(0,1),(1,0),(0,0); Z: (1,1),(1,0),
    (0,1),(0,0).  Checks are numbered row-major by corner; entries are
    ordered by (layer, check).  Z logical: the main diagonal; X logical: the
    bottom row.
    """
    if d < 2:
        raise ValueError("distance must be >= 2")
    x_order = [(1, 1), (0, 1), (1, 0), (0, 0)]
    z_order = [(1, 1), (1, 0), (0, 1), (0, 0)]
    xs, zs = [], []
    for r0 in range(-1, d):
        for c0 in range(-1, d):
            cells = [(r0 + a, c0 + b) for a in (0, 1) for b in (0, 1)]
            inside = [(r, c) for r, c in cells if 0 <= r < d and 0 <= c < d]
            if (r0 + c0) % 2 == 0:
                # X plaquette: interior, or weight-2 on the left/right edge
                if len(inside) == 4 or (len(inside) == 2 and (c0 == -1 or c0 == d - 1)):
                    xs.append((r0, c0))
            else:
                if len(inside) == 4 or (len(inside) == 2 and (r0 == -1 or r0 == d - 1)):
                    zs.append((r0, c0))

    def entries(corners, order):
        out = []
        for layer, (a, b) in enumerate(order):
            for i, (r0, c0) in enumerate(corners):
                r, c = r0 + a, c0 + b
                if 0 <= r < d and 0 <= c < d:
                    out.append((i, r * d + c, layer))
        return out

    return SparseCode(
        num_data=d * d,
        hx=entries(xs, x_order),
        hz=entries(zs, z_order),
        num_x=len(xs),
        num_z=len(zs),
        lx={0: [(d - 1) * d + c for c in range(d)]},
        lz={0: [i * (d + 1) for i in range(d)]},
    )


def load_code(kind: str, d: int) -> SparseCode:
    """Loads data/codes/<kind>/d<d>/{Hx,Hz,Lx,Lz}.mtx (Code.hpp:225-231)."""
    base = os.path.join(DATA, "codes", kind, f"d{d}")
    if not os.path.isdir(base):
        raise ValueError(f"no {kind} code data for d={d}")
    nx, n, hx = read_mtx(os.path.join(base, "Hx.mtx"))
    nz, _, hz = read_mtx(os.path.join(base, "Hz.mtx"))
    lx: dict[int, list[int]] = {}
    lz: dict[int, list[int]] = {}
    for r, c, _ in read_mtx(os.path.join(base, "Lx.mtx"))[2]:
        lx.setdefault(r, []).append(c)
    for r, c, _ in read_mtx(os.path.join(base, "Lz.mtx"))[2]:
        lz.setdefault(r, []).append(c)
    return SparseCode(n, hx, hz, nx, nz, dict(sorted(lx.items())), dict(sorted(lz.items())))


class _Emitter:
    """Builds stim circuit text; merges nothing (one line per instruction)."""

    def __init__(self):
        self.lines: list[str] = []

    def op(self, name: str, targets, arg: float | None = None) -> None:
        head = name if arg is None else f"{name}({fmt_p(arg)})"
        if targets is None:
            self.lines.append(head)
        else:
            self.lines.append(head + " " + " ".join(str(t) for t in targets))

    def text(self) -> str:
        return "\n".join(self.lines) + "\n"


@dataclass
class Noise:
    """Per-instruction noise rates.  With `rng`, every noise instruction draws
    its own rate log-uniformly from [lo, hi] (its base rate is ignored);
    otherwise the base rates p (two-qubit, X_ERROR, measurement) and p/10
    (single-qubit idle) are used as in the reference."""

    p: float = 1e-3
    rng: np.random.Generator | None = None
    lo: float = 1e-4
    hi: float = 1e-2

    def __call__(self, base: float) -> float:
        if self.rng is None:
            return base
        return float(math.exp(self.rng.uniform(math.log(self.lo), math.log(self.hi))))


@dataclass
class Code:
    """A CSS code with a CNOT schedule (Code.hpp:216-407)."""

    sparse: SparseCode
    name: str = "code"
    x_schedule: list[list[int]] = field(default_factory=list)
    z_schedule: list[list[int]] = field(default_factory=list)
    z_support: list[list[int]] = field(default_factory=list)

    def __post_init__(self):
        s = self.sparse
        self.data = list(range(s.num_data))
        self.x_checks = list(range(s.num_data, s.num_data + s.num_x))
        self.z_checks = list(range(s.num_data + s.num_x, s.num_data + s.num_x + s.num_z))
        self.checks = self.x_checks + self.z_checks
        self.all = self.data + self.checks
        self.x_schedule = self._schedule(s.hx, self.x_checks, x_basis=True)
        self.z_schedule = self._schedule(s.hz, self.z_checks, x_basis=False)
        self.z_support = [[] for _ in range(s.num_z)]
        for r, c, _ in s.hz:
            self.z_support[r].append(c)

    @staticmethod
    def _schedule(entries, ids, x_basis: bool) -> list[list[int]]:
        """CNOT layers (Code.hpp:125-155): X checks control data, data
        controls Z checks."""
        layers = 1 + max((v for _, _, v in entries), default=-1)
        sched: list[list[int]] = [[] for _ in range(layers)]
        for r, c, v in entries:
            sched[v] += [ids[r], c] if x_basis else [c, ids[r]]
        return sched

    # ------------------------------------------------------------------
    def _cnot_layers(self, e: _Emitter, sched, noise: Noise) -> None:
        everyone = set(self.all)
        for active in sched:
            e.op("CX", active)
            e.op("TICK", None)
            e.op("DEPOLARIZE2", active, noise(noise.p))
            idle = sorted(everyone - set(active))
            if idle:
                e.op("DEPOLARIZE1", idle, noise(noise.p / 10))

    def _round(self, e: _Emitter, noise: Noise) -> None:
        e.op("R", self.checks)
        e.op("TICK", None)
        e.op("X_ERROR", self.checks, noise(noise.p))
        e.op("DEPOLARIZE1", self.data, noise(noise.p / 10))
        if self.x_checks:
            e.op("H", self.x_checks)
            e.op("TICK", None)
            e.op("DEPOLARIZE1", self.all, noise(noise.p / 10))
        self._cnot_layers(e, self.x_schedule, noise)
        self._cnot_layers(e, self.z_schedule, noise)
        if self.x_checks:
            e.op("H", self.x_checks)
            e.op("TICK", None)
            e.op("DEPOLARIZE1", self.all, noise(noise.p / 10))
        e.op("M", self.checks, noise(noise.p))

    def memory_text(self, rounds: int, p: float = 1e-3, include_x_detectors: bool = True,
                    noise: Noise | None = None, branches: set[tuple[int, int]] | None = None) -> str:
        """stim text of a `rounds`-round memory experiment (Code.hpp:403-407).

        `branches` holds (round, ancilla position in checks) pairs whose
        measurement is followed by a reset-conditioned branch: two extra
        layers that re-reset that ancilla and apply X_ERROR to it."""
        if rounds < 1:
            raise ValueError("rounds must be >= 1")
        noise = noise or Noise(p)
        if noise.rng is None:
            noise.p = p
        e = _Emitter()
        nd, nc, nzc = len(self.data), len(self.checks), len(self.z_checks)
        nxc = len(self.x_checks)
        # head
        e.op("R", self.data)
        e.op("TICK", None)
        e.op("X_ERROR", self.data, noise(noise.p))
        e.op("DEPOLARIZE1", self.checks, noise(noise.p / 10))
        for r in range(rounds):
            self._round(e, noise)
            if r == 0:
                for i in range(nzc):
                    e.op("DETECTOR", [f"rec[-{nzc - i}]"])
            else:
                lo = 0 if include_x_detectors else nxc
                for i in range(lo, nc):
                    e.op("DETECTOR", [f"rec[-{nc - i}]", f"rec[-{2 * nc - i}]"])
            if branches:
                taken = sorted(a for (rr, a) in branches if rr == r)
                if taken:
                    anc = [self.checks[a] for a in taken]
                    e.op("TICK", None)
                    e.op("R", anc)
                    e.op("TICK", None)
                    e.op("X_ERROR", anc, noise(noise.p))
                    e.op("TICK", None)
        # tail
        e.op("M", self.data, noise(noise.p))
        for k in range(nzc):
            sup = self.z_support[k]
            if not sup:
                continue
            e.op("DETECTOR", [f"rec[-{nd - q}]" for q in sup] + [f"rec[-{nd + nzc - k}]"])
        for logical, sup in sorted(self.sparse.lz.items()):
            e.op(f"OBSERVABLE_INCLUDE({logical})", [f"rec[-{nd - q}]" for q in sup])
        return e.text()

    # API names of the reference's Python wrapper (get_memory returns a
    # stim.Circuit there; here, stim circuit text)
    def get_memory(self, rounds: int, p: float = 1e-3, include_x_detectors: bool = True) -> str:
        return self.memory_text(rounds, p, include_x_detectors)


def SurfaceCode(d: int) -> Code:  # noqa: N802 -- the reference's class name
    """Rotated surface code of distance d (any d >= 2)."""
    return Code(rotated_surface_code(d), name=f"surface{d}")


def BBCode(d: int) -> Code:  # noqa: N802 -- the reference's class name
    """Bivariate-bicycle code from data/codes/bb/d<d> (d in 6, 10, 12, 18)."""
    return Code(load_code("bb", d), name=f"bb{d}")


def count_locations(text: str) -> int:
    """Fault locations: one per DEPOLARIZE1 / X_ERROR / M target, one per
    DEPOLARIZE2 pair."""
    n = 0
    for ln in text.splitlines():
        head, _, rest = ln.partition(" ")
        name = head.split("(")[0]
        k = len(rest.split())
        if name in ("DEPOLARIZE1", "X_ERROR", "M"):
            n += k
        elif name == "DEPOLARIZE2":
            n += k // 2
    return n
