"""Detector error models as compact CSR arrays, with stim's text format.

The reference returns a ``stim.DetectorErrorModel`` (Stim.pybind.hpp:50-59).
stim is optional here: :class:`DetectorErrorModel` prints and parses the same
text (``error(p) D0 D5 L0`` lines), compares like stim's ``==`` (same errors,
same order, same probabilities) and converts with :meth:`to_stim` when stim
is importable.
"""

from __future__ import annotations

import re
from dataclasses import dataclass

import numpy as np

_ERROR_RE = re.compile(r"^\s*error\(([^)]*)\)(.*)$")


@dataclass
class DetectorErrorModel:
    """Errors in output order: probability, and targets (detector index t <
    num_detectors, else observable t - num_detectors) as CSR.  Targets are
    unsigned integers: 16-bit in a driver's DEM when every target fits, else
    32-bit."""

    probs: np.ndarray
    offsets: np.ndarray
    targets: np.ndarray
    num_detectors: int
    num_observables: int
    hashes: np.ndarray | None = None

    # ------------------------------------------------------------------
    @property
    def num_errors(self) -> int:
        return int(len(self.probs))

    def errors(self):
        """Yields (probability, (target, ...)) in order."""
        off = self.offsets
        tg = self.targets
        for i in range(self.num_errors):
            yield float(self.probs[i]), tuple(int(t) for t in tg[off[i]:off[i + 1]])

    def target_lists(self) -> list[tuple[int, ...]]:
        return [t for _, t in self.errors()]

    def __str__(self) -> str:
        D = self.num_detectors
        lines = []
        for p, ts in self.errors():
            parts = [f"error({p!r})"]
            parts += [f"D{t}" if t < D else f"L{t - D}" for t in ts]
            lines.append(" ".join(parts))
        return "\n".join(lines) + ("\n" if lines else "")

    def __repr__(self) -> str:
        return (f"DetectorErrorModel(num_errors={self.num_errors}, "
                f"num_detectors={self.num_detectors}, num_observables={self.num_observables})")

    def _signed_targets(self) -> np.ndarray:
        """Detectors as d >= 0, observables as -(k + 1)."""
        t = self.targets.astype(np.int64)
        return np.where(t < self.num_detectors, t, -(t - self.num_detectors) - 1)

    def __eq__(self, other) -> bool:
        """Same errors in the same order with identical probabilities (the
        comparison stim's ``DetectorErrorModel.__eq__`` makes)."""
        if not isinstance(other, DetectorErrorModel):
            return NotImplemented
        return (self.num_errors == other.num_errors
                and np.array_equal(self.offsets, other.offsets)
                and np.array_equal(self._signed_targets(), other._signed_targets())
                and np.array_equal(self.probs, other.probs))

    # ------------------------------------------------------------------
    def canonical(self) -> "DetectorErrorModel":
        """Errors sorted by target list (the order-independent form)."""
        keys = self.target_lists()
        order = sorted(range(self.num_errors), key=lambda i: keys[i])
        probs = self.probs[order]
        lens = np.diff(self.offsets)[order] if self.num_errors else np.zeros(0, np.int64)
        offsets = np.zeros(self.num_errors + 1, np.uint32)
        offsets[1:] = np.cumsum(lens)
        targets = np.concatenate([self.targets[self.offsets[i]:self.offsets[i + 1]] for i in order]) \
            if self.num_errors else np.zeros(0, np.uint32)
        return DetectorErrorModel(probs, offsets, targets.astype(np.uint32), self.num_detectors,
                                  self.num_observables)

    def approx_equals(self, other: "DetectorErrorModel", rel_tol: float = 1e-12,
                      canonical: bool = True) -> bool:
        """Same error set (exact target lists) with probabilities within
        `rel_tol` relative error; order-independent when `canonical`."""
        a, b = (self.canonical(), other.canonical()) if canonical else (self, other)
        if a.num_errors != b.num_errors:
            return False
        if not (np.array_equal(a.offsets, b.offsets)
                and np.array_equal(a._signed_targets(), b._signed_targets())):
            return False
        if a.num_errors == 0:
            return True
        return bool(np.all(np.abs(a.probs - b.probs) <= rel_tol * np.abs(b.probs)))

    def max_rel_error(self, other: "DetectorErrorModel") -> float:
        a, b = self.canonical(), other.canonical()
        if a.num_errors != b.num_errors or not np.array_equal(a._signed_targets(),
                                                                b._signed_targets()):
            return float("inf")
        if a.num_errors == 0:
            return 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.abs(a.probs - b.probs) / np.abs(b.probs)
        r[(a.probs == b.probs)] = 0.0
        return float(np.max(r))

    # ------------------------------------------------------------------
    @staticmethod
    def from_text(text: str, num_detectors: int | None = None,
                  num_observables: int | None = None) -> "DetectorErrorModel":
        """Parses stim DEM text made of ``error(p) D.. L..`` lines."""
        probs, lens, dets, obs_flags = [], [], [], []
        for line in text.splitlines():
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            m = _ERROR_RE.match(line)
            if not m:
                if line.startswith(("detector", "logical_observable")):
                    continue
                raise ValueError(f"unsupported DEM line: {line!r}")
            probs.append(float(m.group(1)))
            n = 0
            for tok in m.group(2).split():
                if tok[0] == "D":
                    dets.append(int(tok[1:]))
                    obs_flags.append(False)
                elif tok[0] == "L":
                    dets.append(int(tok[1:]))
                    obs_flags.append(True)
                elif tok == "^":
                    raise ValueError("decomposed errors ('^') are not supported")
                else:
                    raise ValueError(f"bad DEM target {tok!r}")
                n += 1
            lens.append(n)
        d_arr = np.array(dets, np.int64)
        o_arr = np.array(obs_flags, bool)
        if num_detectors is None:
            num_detectors = int(d_arr[~o_arr].max()) + 1 if (~o_arr).any() else 0
        if num_observables is None:
            num_observables = int(d_arr[o_arr].max()) + 1 if o_arr.any() else 0
        targets = np.where(o_arr, d_arr + num_detectors, d_arr).astype(np.uint32)
        offsets = np.zeros(len(probs) + 1, np.uint32)
        offsets[1:] = np.cumsum(lens)
        return DetectorErrorModel(np.array(probs, np.float64), offsets, targets, num_detectors,
                                  num_observables)

    @staticmethod
    def from_file(path, num_detectors: int | None = None,
                  num_observables: int | None = None) -> "DetectorErrorModel":
        with open(path) as f:
            return DetectorErrorModel.from_text(f.read(), num_detectors, num_observables)

    def to_stim(self):
        """The same model as a ``stim.DetectorErrorModel`` (needs stim)."""
        import stim  # noqa: PLC0415 -- optional dependency

        return stim.DetectorErrorModel(str(self))
