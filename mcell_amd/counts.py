"""Count output beside the diffuse-react path: MolOrRxnCountEvent's values
from the engine's counters, written as MCell4's CountBuffer writes them.

* ``CountBuffer`` (src4/count_buffer.cpp:20-170): one file per buffer; the
  .dat form holds one column as ``time value`` lines (CountItem::
  write_as_dat, :30-32, iostream's default formatting of a double: the
  shortest of %g at precision 6); the .gdat form holds several columns
  under a header ``#`` + "time" right-aligned to 14 characters + each
  column name right-aligned to 16 (or a space and the name when it has 14
  characters or more), then rows of `` time`` and ``  value`` in
  scientific form with 8 digits after the point and an exponent of at
  least two digits (write_gdat_value, :35-61).  Rows are kept until
  ``flush``; a buffer opened for appending writes no header.
* ``MolOrRxnCountEvent`` (src4/mol_or_rxn_count_event.cpp:607-716): every
  count item's value at the event's time (iteration x time unit) is the sum
  of its signed terms times the item's multiplier; a term counts the
  molecules of a species (in the whole replica or in one counted volume)
  or the reactions of one pathway slot carried out so far.  The counts
  come from the engine: mcg_count_species, mcg_count_species_by_volume and
  mcg_get_rxn_counts.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

GDAT_COLUMN_WIDTH = 14


def format_dat_value(d: float) -> str:
    """``std::ostream << double`` with the stream's defaults (%g, precision 6)."""
    return "%g" % d


def format_gdat_value(d: float) -> str:
    """write_gdat_value (count_buffer.cpp:35-61): scientific, 8 digits after
    the point, the exponent's sign and at least two digits."""
    return "%.*e" % (GDAT_COLUMN_WIDTH - 6, d)


class CountBuffer:
    """One output file of count columns (CountBuffer, count_buffer.h)."""

    def __init__(self, filename: str, column_names: List[str], fmt: str = "dat", append: bool = False):
        if fmt not in ("dat", "gdat"):
            raise ValueError("count output is 'dat' or 'gdat'")
        if fmt == "dat" and len(column_names) != 1:
            raise ValueError("a .dat buffer holds one column")
        self.filename, self.fmt, self.append = filename, fmt, append
        self.column_names = list(column_names)
        self.data = [[] for _ in column_names]  # per column: (time, value)
        self.fout = None

    def add(self, column: int, time: float, value: float):
        self.data[column].append((time, value))

    def _open(self):
        import os
        d = os.path.dirname(self.filename)
        if d:
            os.makedirs(d, exist_ok=True)
        self.fout = open(self.filename, "a" if self.append else "w")
        if self.fmt == "gdat" and not self.append:
            self.fout.write(self._gdat_header())

    def _gdat_header(self) -> str:
        h = "#" + "time".rjust(GDAT_COLUMN_WIDTH)
        for name in self.column_names:
            h += name.rjust(GDAT_COLUMN_WIDTH + 2) if GDAT_COLUMN_WIDTH > len(name) else " " + name
        return h + "\n"

    def flush(self):
        if self.fout is None:
            self._open()
        if self.fmt == "dat":
            for t, v in self.data[0]:
                self.fout.write("%s %s\n" % (format_dat_value(t), format_dat_value(v)))
        else:
            rows = len(self.data[0])
            if any(len(c) != rows for c in self.data):
                raise RuntimeError("count columns of one buffer differ in length")
            for r in range(rows):
                t = self.data[0][r][0]
                line = " " + format_gdat_value(t)
                for c in self.data:
                    if abs(c[r][0] - t) >= 1.0e-6 * max(1.0, abs(t)):  # cmp_eq(..., SQRT_EPS)
                        raise RuntimeError("mismatch in gdat column times")
                    line += "  " + format_gdat_value(c[r][1])
                self.fout.write(line + "\n")
        self.fout.flush()
        for c in self.data:
            c.clear()

    def flush_and_close(self):
        self.flush()
        self.fout.close()
        self.fout = None


@dataclass
class CountTerm:
    """A signed term: molecules of ``index`` (a species) or reactions of
    ``index`` (a pathway slot of mcg_get_rxn_counts); ``volume``: a counted
    volume index, or None for the whole replica."""
    kind: str
    index: int
    sign: int = 1
    volume: Optional[int] = None


@dataclass
class CountItem:
    buffer: CountBuffer
    column: int
    terms: List[CountTerm] = field(default_factory=list)
    multiplier: float = 1.0
    replica: int = 0


class MolOrRxnCountEvent:
    """The count event over the engine's counters (one call per event)."""

    def __init__(self, engine, items: List[CountItem], num_replicas: int, num_species: int, time_unit: float,
                 num_volumes: int = 0):
        for it in items:
            for t in it.terms:
                if t.kind not in ("species", "rxn"):
                    raise ValueError("a count term counts 'species' or 'rxn'")
                if t.volume is not None and not (0 <= t.volume < num_volumes):
                    raise ValueError("counted volume out of range")
        self.engine, self.items = engine, items
        self.num_replicas, self.num_species, self.num_volumes = num_replicas, num_species, num_volumes
        self.time_unit = time_unit

    def compute_counts(self) -> List[float]:
        need_vol = any(t.kind == "species" and t.volume is not None for it in self.items for t in it.terms)
        need_rxn = any(t.kind == "rxn" for it in self.items for t in it.terms)
        need_all = any(t.kind == "species" and t.volume is None for it in self.items for t in it.terms)
        sp = self.engine.count_species(self.num_replicas, self.num_species) if need_all else None
        spv = (self.engine.count_species_by_volume(self.num_replicas, self.num_volumes, self.num_species)
               if need_vol else None)
        rx = self.engine.rxn_counts() if need_rxn else None
        out = []
        for it in self.items:
            v = 0.0
            for t in it.terms:
                if t.kind == "species":
                    n = sp[it.replica, t.index] if t.volume is None else spv[it.replica, t.volume, t.index]
                else:
                    per_vol = rx[it.replica, :, t.index]
                    n = per_vol.sum() if t.volume is None else per_vol[t.volume]
                v += t.sign * float(n)
            out.append(v * it.multiplier)
        return out

    def step(self, iteration: int):
        time = iteration * self.time_unit
        for it, v in zip(self.items, self.compute_counts()):
            it.buffer.add(it.column, time, v)
