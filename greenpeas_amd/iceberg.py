"""Concatenated Iceberg x surface-code memory experiments, static and adaptive.

The reference's generator (include/GreenPeas/QEC/Codes/ConcatenatedSurfaceCode.hpp,
after Berthusen et al., arXiv:2502.14835) and its adaptive recompile loop
(include/GreenPeas/Cli/Cli.hpp:333-432), written here from their behaviour:

* layout: the 2n data qubits of n/2 inner Iceberg [[4,2,2]] blocks, then the
  inner X-check, inner Z-check, outer X-check and outer Z-check ancillas
  (ConcatenatedSurfaceCode.hpp:100-155);
* outer data qubit (r, c) of the d x d rotated surface code pairs with
  (d - 1 - r, (c + 2) mod d) into one Iceberg block, as its logical 0 and 1
  (:357-378); an outer check's CNOT layer k becomes two physical layers 2k,
  2k + 1 on the block's qubits {4b, 4b + 1 + logical} (X) or
  {4b + 3, 4b + 1 + logical} (Z) (:404-470);
* one QEC round = an inner QED round (every Iceberg block's X and Z check,
  4 CNOT layers each) then an outer round over the flagged outer checks, the
  X and Z schedules interleaved; circuit-level noise (X_ERROR p on resets,
  DEPOLARIZE2 p after CNOT layers, DEPOLARIZE1 p/10 on idle qubits, M(p))
  in `circl`, DEPOLARIZE1 p/10 on data per round in `pheno` (:503-635);
* static: every outer check every round; adaptive: the outer checks of the
  blocks whose inner checks flipped since the previous round, and every
  outer check halfway and in the last round (:737-867);
* detectors and the Z observables as in the reference (:737-928).

The adaptive branch taken depends on the measurement outcomes of a shot.
The reference draws them with stim's tableau simulator; here a Pauli-frame
simulator over the generated instructions (seeded numpy generator) gives,
per inner check, whether its outcome differs from the noiseless one, which
is all the flag differences depend on (a check's noiseless outcome repeats
from its second measurement on).  At p = 0 no check flips and both
strategies are deterministic: the generated circuits equal the reference's
data/circuits/iceberg_surface_{4,6}_{static,adaptive}_{circl,pheno}.stim
line for line (CPU tests).  The shot itself (detector hits) is not produced:
the compile path needs only the circuits.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .codes import SparseCode, fmt_p, rotated_surface_code

_UNFUSED = {"TICK", "DETECTOR", "OBSERVABLE_INCLUDE"}
REC = 1 << 31  # record target bit (a lookback k is REC | k)


class _Circuit:
    """Instructions [name, args, targets] with stim's append rule: an
    instruction with the same name and arguments as the last one (other
    than TICK and annotations) extends it."""

    def __init__(self):
        self.ops: list[list] = []

    def append(self, name: str, targets, args=()) -> None:
        args = tuple(float(a) for a in args)
        targets = list(targets)
        if self.ops and name not in _UNFUSED and self.ops[-1][0] == name and self.ops[-1][1] == args:
            self.ops[-1][2].extend(targets)
        else:
            self.ops.append([name, args, targets])

    def extend(self, other: "_Circuit") -> None:
        for name, args, targets in other.ops:
            self.append(name, targets, args)

    def text(self) -> str:
        lines = []
        for name, args, targets in self.ops:
            head = name if not args else f"{name}({', '.join(fmt_p(a) for a in args)})"
            ts = " ".join(f"rec[-{t & ~REC}]" if t & REC else str(t) for t in targets)
            lines.append(f"{head} {ts}" if ts else head)
        return "\n".join(lines)  # as stim's str(circuit): no final newline


@dataclass
class CircuitPair:
    circl: _Circuit = field(default_factory=_Circuit)
    pheno: _Circuit = field(default_factory=_Circuit)

    def both(self, name, targets, args=()):
        self.circl.append(name, targets, args)
        self.pheno.append(name, targets, args)

    def extend(self, other: "CircuitPair") -> None:
        self.circl.extend(other.circl)
        self.pheno.extend(other.pheno)


class _FrameSim:
    """Pauli-frame simulator: per qubit an X and a Z error bit; each
    measurement records whether its outcome differs from the noiseless
    one."""

    def __init__(self, n: int, rng: np.random.Generator | None):
        self.x = np.zeros(n, bool)
        self.z = np.zeros(n, bool)
        self.rng = rng
        self.record: list[bool] = []

    def _hits(self, k: int, p: float) -> np.ndarray:
        if self.rng is None or p <= 0:
            return np.zeros(k, bool)
        return self.rng.random(k) < p

    def run(self, c: _Circuit) -> None:
        for name, args, targets in c.ops:
            t = np.asarray(targets, np.int64)
            if name == "R":
                self.x[t] = False
                self.z[t] = False
            elif name == "H":
                self.x[t], self.z[t] = self.z[t].copy(), self.x[t].copy()
            elif name == "CX":
                for a, b in zip(targets[0::2], targets[1::2]):
                    self.x[b] ^= self.x[a]
                    self.z[a] ^= self.z[b]
            elif name == "X_ERROR":
                self.x[t] ^= self._hits(len(t), args[0])
            elif name == "DEPOLARIZE1":
                hit = self._hits(len(t), args[0])
                if hit.any():
                    kind = self.rng.integers(1, 4, int(hit.sum()))  # X, Y, Z
                    q = t[hit]
                    self.x[q] ^= (kind & 1).astype(bool)
                    self.z[q] ^= (kind & 2).astype(bool)
            elif name == "DEPOLARIZE2":
                pairs = t.reshape(-1, 2)
                hit = self._hits(len(pairs), args[0])
                if hit.any():
                    kind = self.rng.integers(1, 16, int(hit.sum()))  # 15 non-identity pairs
                    a, b = pairs[hit, 0], pairs[hit, 1]
                    self.x[a] ^= (kind & 1).astype(bool)
                    self.z[a] ^= (kind & 2).astype(bool)
                    self.x[b] ^= (kind & 4).astype(bool)
                    self.z[b] ^= (kind & 8).astype(bool)
            elif name == "M":
                flips = self.x[t] ^ self._hits(len(t), args[0] if args else 0.0)
                self.record.extend(bool(f) for f in flips)


class IcebergSurfaceCode:
    """Iceberg [[4,2,2]] blocks concatenated with the rotated surface code of
    even distance d (ConcatenatedSurfaceCode(d))."""

    def __init__(self, d: int, code: SparseCode | None = None):
        if d < 2 or d % 2:
            raise ValueError("the Iceberg pairing needs an even outer distance d >= 2")
        c = code or rotated_surface_code(d)
        self.d = d
        n_outer = c.num_data
        n_ix = n_outer // 2
        n_iz = n_outer // 2
        self.num_data = 2 * n_outer
        base = self.num_data
        self.inner_x = list(range(base, base + n_ix))
        self.inner_z = list(range(base + n_ix, base + n_ix + n_iz))
        self.inner = self.inner_x + self.inner_z
        base += n_ix + n_iz
        self.outer_x = list(range(base, base + c.num_x))
        self.outer_z = list(range(base + c.num_x, base + c.num_x + c.num_z))
        self.outer = self.outer_x + self.outer_z
        self.all = list(range(base + c.num_x + c.num_z))
        self.data = list(range(self.num_data))
        self.checks = list(range(self.num_data, len(self.all)))

        # outer data qubits paired into Iceberg blocks (block, logical)
        blocks: list[tuple[int, int]] = [(0, 0)] * n_outer
        assigned = [False] * n_outer
        b = 0
        for r in range(d):
            for col in range(d):
                q0 = r * d + col
                if assigned[q0]:
                    continue
                q1 = (d - 1 - r) * d + (col + 2) % d
                blocks[q0] = (b, 0)
                blocks[q1] = (b, 1)
                assigned[q0] = assigned[q1] = True
                b += 1

        self.inner_x_sched = [[] for _ in range(4)]
        for i, anc in enumerate(self.inner_x):
            for k in range(4):
                self.inner_x_sched[k] += [anc, 4 * i + k]
        self.inner_z_sched = [[] for _ in range(4)]
        for i, anc in enumerate(self.inner_z):
            for k in range(4):
                self.inner_z_sched[k] += [4 * i + k, anc]

        def lift(entries, anc_ids, first, x_basis):
            sched: list[list[int]] = []
            support = [[] for _ in anc_ids]
            block_support = [[] for _ in range(n_ix)]
            for row, col, layer in entries:
                blk, logical = blocks[col]
                q0 = 4 * blk + (0 if x_basis else 3)
                q1 = 4 * blk + 1 + logical
                anc = anc_ids[row]
                support[row] += [q0, q1]
                block_support[blk].append(anc)
                while len(sched) < 2 * layer + 2:
                    sched.append([])
                if x_basis:
                    sched[2 * layer] += [anc, q0]
                    sched[2 * layer + 1] += [anc, q1]
                else:
                    sched[2 * layer] += [q0, anc]
                    sched[2 * layer + 1] += [q1, anc]
            return sched, support, block_support

        x_sched, self.outer_x_support, self.block_x_support = lift(c.hx, self.outer_x, 0, True)
        z_sched, self.outer_z_support, self.block_z_support = lift(c.hz, self.outer_z, 0, False)
        self.logical_z = {row: [q for col in cols for q in (4 * blocks[col][0] + 3, 4 * blocks[col][0] + 1 + blocks[col][1])]
                          for row, cols in c.lz.items()}
        n_layers = max(len(x_sched), len(z_sched))
        self.outer_sched = [(x_sched[i] if i < len(x_sched) else []) + (z_sched[i] if i < len(z_sched) else [])
                            for i in range(n_layers)]

    # ------------------------------------------------------------------
    def _idle(self, active) -> list[int]:
        busy = set(active)
        return [q for q in self.all if q not in busy]

    def _cnot_layers(self, c: CircuitPair, sched, p: float, flagged: set | None = None) -> None:
        for active in sched:
            if flagged is not None:
                active = [q for a, b in zip(active[0::2], active[1::2]) if a in flagged or b in flagged
                          for q in (a, b)]
                if not active:
                    continue
            c.both("CX", active)
            c.both("TICK", [])
            c.circl.append("DEPOLARIZE2", active, [p])
            idle = self._idle(active)
            if idle:
                c.circl.append("DEPOLARIZE1", idle, [p / 10])

    def _head(self, p: float) -> CircuitPair:
        c = CircuitPair()
        c.both("R", self.data)
        c.both("TICK", [])
        c.circl.append("X_ERROR", self.data, [p])
        c.circl.append("DEPOLARIZE1", self.checks, [p / 10])
        return c

    def _qed_round(self, p: float) -> CircuitPair:
        c = CircuitPair()
        c.both("R", self.inner)
        c.both("TICK", [])
        c.circl.append("X_ERROR", self.inner, [p])
        idle = self._idle(self.inner)
        if idle:
            c.circl.append("DEPOLARIZE1", idle, [p / 10])
        c.pheno.append("DEPOLARIZE1", self.data, [p / 10])
        if self.inner_x:
            c.both("H", self.inner_x)
            c.both("TICK", [])
            c.circl.append("DEPOLARIZE1", self.all, [p / 10])
        self._cnot_layers(c, self.inner_x_sched, p)
        self._cnot_layers(c, self.inner_z_sched, p)
        if self.inner_x:
            c.both("H", self.inner_x)
            c.both("TICK", [])
            c.circl.append("DEPOLARIZE1", self.all, [p / 10])
        c.both("M", self.inner, [p])
        return c

    def _qec_round(self, p: float, flags) -> CircuitPair:
        c = CircuitPair()
        nx = len(self.outer_x)
        fx = [self.outer_x[i] for i in range(nx) if flags[i]]
        fz = [self.outer_z[i - nx] for i in range(nx, len(self.outer)) if flags[i]]
        flagged = fx + fz
        if not flagged:
            return c
        c.both("R", flagged)
        c.both("TICK", [])
        c.circl.append("X_ERROR", flagged, [p])
        idle = self._idle(flagged)
        if idle:
            c.circl.append("DEPOLARIZE1", idle, [p / 10])
        c.pheno.append("DEPOLARIZE1", self.data, [p / 10])
        if fx:
            c.both("H", fx)
            c.both("TICK", [])
            c.circl.append("DEPOLARIZE1", self.all, [p / 10])
        if len(flagged) == len(self.outer):
            self._cnot_layers(c, self.outer_sched, p)
        else:
            self._cnot_layers(c, self.outer_sched, p, set(flagged))
        if fx:
            c.both("H", fx)
            c.both("TICK", [])
            c.circl.append("DEPOLARIZE1", self.all, [p / 10])
        c.both("M", flagged, [p])
        return c

    def _inner_flags(self, sim: _FrameSim) -> np.ndarray:
        n = len(self.inner)
        return np.array(sim.record[-n:], bool)

    def _outer_flags(self, diff: np.ndarray) -> list[int]:
        flags = [0] * len(self.outer)
        first = self.num_data + len(self.inner)
        nix = len(self.inner_x)
        for i in range(nix):
            if diff[i]:
                for q in self.block_x_support[i]:
                    flags[q - first] = 1
        for i in range(len(self.inner_z)):
            if diff[nix + i]:
                for q in self.block_z_support[i]:
                    flags[q - first] = 1
        return flags

    def memory(self, rounds: int, p: float = 1e-3, adaptive: bool = False, include_x_detectors: bool = True,
               seed: int | None = None) -> tuple[str, str]:
        """(circuit-level, phenomenological) stim text of one memory
        experiment (getMemory(rounds, p, strategy, includeXDetectors)).
        Adaptive circuits follow the flags of one shot drawn with `seed`."""
        rng = np.random.default_rng(seed) if (adaptive and p > 0) else None
        sim = _FrameSim(len(self.all), rng)
        n_in, n_ix, n_iz = len(self.inner), len(self.inner_x), len(self.inner_z)
        n_out, n_ox, n_oz = len(self.outer), len(self.outer_x), len(self.outer_z)

        out = self._head(p)
        sim.run(out.circl)

        outer_flags = [1] * n_out
        total = n_in + n_out
        last = [n_in + i for i in range(n_out)]

        rnd = 0
        qed = self._qed_round(p)
        for i in range(n_iz):
            qed.both("DETECTOR", [REC | (n_iz - i)], [1, self.inner_z[i], rnd])
        sim.run(qed.circl)
        prev = self._inner_flags(sim)
        qec = self._qec_round(p, outer_flags)
        for i in range(n_oz):
            qec.both("DETECTOR", [REC | (n_oz - i)], [1, self.outer_z[i], rnd])
        sim.run(qec.circl)
        out.extend(qed)
        out.extend(qec)
        rnd += 1
        n_flagged = n_out
        for rnd in range(1, rounds):
            qed = self._qed_round(p)
            lo = 0 if include_x_detectors else n_ix
            for i in range(lo, n_in):
                qed.both("DETECTOR", [REC | (n_in - i), REC | (2 * n_in - i + n_flagged)],
                         [1, self.inner[i], rnd])
            sim.run(qed.circl)
            cur = self._inner_flags(sim)
            if not adaptive:
                outer_flags = [1] * n_out
            else:
                outer_flags = self._outer_flags(prev ^ cur)
                prev = cur
                if rnd == rounds - 1 or rnd == rounds // 2:
                    outer_flags = [1] * n_out
            qec = self._qec_round(p, outer_flags)
            n_flagged = sum(outer_flags)
            total += n_in + n_flagged
            k = 0
            for i in range(n_out):
                before = now = last[i]
                if outer_flags[i]:
                    now = total - n_flagged + k
                    last[i] = now
                    k += 1
                if now - before > 0 and (include_x_detectors or i >= n_ox):
                    qec.both("DETECTOR", [REC | (total - before), REC | (total - now)],
                             [1, self.outer[i], rnd])
            sim.run(qec.circl)
            out.extend(qed)
            out.extend(qec)

        tail = CircuitPair()
        nd = self.num_data
        tail.both("M", self.data, [p])
        for i in range(n_iz):
            b = 4 * i
            recs = [REC | (nd - b - k) for k in range(4)] + [REC | (nd + n_out + n_iz - i)]
            tail.both("DETECTOR", recs, [1, self.inner_z[i], rounds])
        for check in range(n_oz):
            support = self.outer_z_support[check]
            if not support:
                continue
            recs = [REC | (nd - q) for q in support] + [REC | (nd + n_oz - check)]
            tail.both("DETECTOR", recs, [1, self.outer_z[check], rounds])
        for logical, support in sorted(self.logical_z.items()):
            tail.both("OBSERVABLE_INCLUDE", [REC | (nd - q) for q in support], [logical])
        out.extend(tail)
        return out.circl.text(), out.pheno.text()


def adaptive_branches(d: int, count: int, p: float = 1e-3, rounds: int | None = None, seed: int = 0,
                      include_x_detectors: bool = True) -> list[str]:
    """The adaptive recompile loop's circuits (Cli.hpp:378-416): `count`
    circuit-level memory experiments of the Iceberg x surface code of
    distance d (rounds = d by default), each following the outer-check
    flags of its own shot (seeds seed, seed + 1, ...)."""
    code = IcebergSurfaceCode(d)
    r = d if rounds is None else rounds
    return [code.memory(r, p, adaptive=True, include_x_detectors=include_x_detectors, seed=seed + i)[0]
            for i in range(count)]
