"""The species table (the reference's PDG-table format, particle_decay.cpp:
23-129): the engine's reader (hbt_species_read, host code) against the
oracle's restatement, on the committed table with LF and CRLF line ends;
the anti-baryon entries and the zero entry of an undefined decay product
that the reference's reader creates; malformed files refused by both."""
import os

import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import hbt

TABLE = os.path.join(cases.HERE, "species_table.dat")


def as_map(t):
    return {int(m): (int(c), int(b), int(s))
            for m, c, b, s in zip(t["monval"], t["charge"], t["baryon"], t["strange"])}


@pytest.mark.parametrize("crlf", [False, True])
def test_engine_reader_equals_oracle(crlf, tmp_path):
    path = TABLE
    if crlf:
        path = str(tmp_path / "pdg.dat")
        with open(TABLE) as f, open(path, "w", newline="") as g:
            g.write(f.read().replace("\n", "\r\n"))
    e, o = as_map(hbt.read_species_table(path)), as_map(oracle.species_read(path))
    assert e == o
    # stable species as written, anti-baryons added, mesons not mirrored
    assert e[211] == (1, 0, 0) and e[-321] == (-1, 0, -1) and e[311] == (0, 0, 1)
    assert e[2212] == (1, 1, 0) and e[-2212] == (-1, -1, 0)
    assert e[3122] == (0, 1, -1) and e[-3122] == (0, -1, 1)
    assert e[-2224] == (-2, -1, 0) and -111 not in e and -211 in e
    # the reference's lookup of the Delta's decay products creates this one
    assert e[7777777] == (0, 0, 0)
    assert len(e) == 11 + 4 + 1  # 11 written, 4 anti-baryons, 1 decay product


def test_engine_species_order_keeps_first_insertion():
    t = hbt.read_species_table(TABLE)
    assert list(t["monval"][:3]) == [211, 111, -211]
    assert np.all(np.sort(t["monval"]) == np.sort(oracle.species_read(TABLE)["monval"]))


@pytest.mark.parametrize("text", ["211 Pion(+) 0.1 0.0 1 0 0 0 0 3 1 1\n211 1 1.0 211 0 0\n",
                                  "211 Pion(+) 0.1 0.0 1 0 0 zero 0 3 1 1\n",
                                  "211 Pion(+) 0.1 0.0 1 0 0 0 0 3 1 -1\n"])
def test_malformed_tables_are_refused(text, tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text(text)
    with pytest.raises(hbt.HbtError):
        hbt.read_species_table(str(p))
    with pytest.raises(ValueError):
        oracle.species_read(str(p))
