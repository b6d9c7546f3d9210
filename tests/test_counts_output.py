"""Count output (MolOrRxnCountEvent + CountBuffer, src4/mol_or_rxn_count_
event.cpp:607-716, src4/count_buffer.cpp:20-170) on the CPU.

* Number text against C++ iostreams themselves (compiled into the oracle's
  library): a default stream for .dat, std::scientific at precision 8 for
  .gdat, over counts, times and random magnitudes.
* File layout: .dat lines "time value"; the .gdat header widths and row
  layout; appending writes no header; a .gdat buffer refuses columns whose
  times differ.
* The event: signed terms times the multiplier over a stand-in engine's
  counters, at iteration x time unit.
"""
import numpy as np
import pytest

from mcell_amd import counts as C


def test_number_text_matches_iostreams(oracle_lib):
    from oracle import pyoracle as po
    rng = np.random.default_rng(3)
    vals = [0.0, -0.0, 1.0, 12.0, 999999.0, 1e6, 1234567.0, 1e-6, 2.5e-6, 1e-5 * 37, 0.1, 1.0 / 3.0,
            float("inf"), 123456789012.0, 1e100, 1e-100, 5e-324]
    vals += list(10.0 ** rng.uniform(-12, 12, size=300) * rng.choice([-1, 1], size=300))
    vals += list(rng.integers(0, 10 ** 9, size=100).astype(float))
    vals += list(np.arange(0, 2000) * 1e-6)
    for d in vals:
        assert C.format_dat_value(d) == po.format_count(d, 0), d
        assert C.format_gdat_value(d) == po.format_count(d, 1), d


def test_dat_and_gdat_files(tmp_path):
    b = C.CountBuffer(str(tmp_path / "out" / "A.dat"), ["A"])
    b.add(0, 0.0, 100)
    b.add(0, 1e-5, 97)
    b.flush()
    b.add(0, 2e-5, 1234567)
    b.flush_and_close()
    assert (tmp_path / "out" / "A.dat").read_text() == "0 100\n1e-05 97\n2e-05 1.23457e+06\n"
    g = C.CountBuffer(str(tmp_path / "counts.gdat"), ["A", "a_very_long_name_x"], fmt="gdat")
    g.add(0, 1e-6, 5)
    g.add(1, 1e-6, 0.5)
    g.flush_and_close()
    lines = (tmp_path / "counts.gdat").read_text().split("\n")
    assert lines[0] == "#" + " " * 10 + "time" + " " * 15 + "A" + " a_very_long_name_x"
    assert lines[1] == " 1.00000000e-06  5.00000000e+00  5.00000000e-01"
    a = C.CountBuffer(str(tmp_path / "counts.gdat"), ["A", "B"], fmt="gdat", append=True)
    a.add(0, 2e-6, 1)
    a.add(1, 2e-6, 2)
    a.flush_and_close()
    assert (tmp_path / "counts.gdat").read_text().split("\n")[2] == " 2.00000000e-06  1.00000000e+00  2.00000000e+00"
    bad = C.CountBuffer(str(tmp_path / "bad.gdat"), ["A", "B"], fmt="gdat")
    bad.add(0, 1e-6, 1)
    bad.add(1, 2e-6, 1)
    with pytest.raises(RuntimeError):
        bad.flush()
    with pytest.raises(ValueError):
        C.CountBuffer(str(tmp_path / "x.dat"), ["A", "B"])


class _Counters:
    """Stand-in for the engine's counters: 2 replicas, 3 species, 2 counted
    volumes, 5 reaction slots."""

    def count_species(self, r, s):
        return np.arange(r * s, dtype=np.uint64).reshape(r, s) + 10

    def count_species_by_volume(self, r, v, s):
        return np.arange(r * v * s, dtype=np.uint64).reshape(r, v, s)

    def rxn_counts(self):
        return np.arange(2 * 2 * 5, dtype=np.uint64).reshape(2, 2, 5) * 3


def test_count_event_terms(tmp_path):
    b = C.CountBuffer(str(tmp_path / "c.gdat"), ["x", "y", "z"], fmt="gdat")
    items = [C.CountItem(b, 0, [C.CountTerm("species", 1), C.CountTerm("species", 2, -1)], multiplier=2.0,
                         replica=1),
             C.CountItem(b, 1, [C.CountTerm("species", 0, volume=1), C.CountTerm("rxn", 4)]),
             C.CountItem(b, 2, [C.CountTerm("rxn", 3, volume=0)], replica=1)]
    ev = C.MolOrRxnCountEvent(_Counters(), items, num_replicas=2, num_species=3, time_unit=1e-6, num_volumes=2)
    sp, spv, rx = _Counters().count_species(2, 3), _Counters().count_species_by_volume(2, 2, 3), _Counters().rxn_counts()
    assert ev.compute_counts() == [2.0 * (float(sp[1, 1]) - float(sp[1, 2])),
                                   float(spv[0, 1, 0]) + float(rx[0, :, 4].sum()), float(rx[1, 0, 3])]
    ev.step(7)
    b.flush_and_close()
    row = (tmp_path / "c.gdat").read_text().split("\n")[1].split()
    assert float(row[0]) == 7e-6 and [float(x) for x in row[1:]] == ev.compute_counts()
    with pytest.raises(ValueError):
        C.MolOrRxnCountEvent(_Counters(), [C.CountItem(b, 0, [C.CountTerm("species", 0, volume=2)])], 2, 3, 1e-6, 2)
