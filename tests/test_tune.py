"""The tuning table (no GPU): entries set, read back, removed and cleared
through the C ABI; candidate shapes; JSON tables."""
import pytest

from tbfsolver_amd import tune as Tu
from tbfsolver_amd._lib import TbfgError


def test_table_round_trip():
    Tu.clear()
    assert Tu.get_entry("sweep", (256, 192, 256), "mixed") == (0, 0, 0)
    Tu.set_entry("sweep", (256, 192, 256), "mixed", (32, 16, 256))
    Tu.set_entry("op", (256, 192, 256), "fp64", (-1, 0, 0))
    assert Tu.get_entry("sweep", (256, 192, 256), "mixed") == (32, 16, 256)
    assert Tu.get_entry("sweep", (256, 192, 256), "fp64") == (0, 0, 0)        # precision is part of the key
    assert Tu.get_entry("sweep_edge", (256, 192, 256), "mixed") == (0, 0, 0)  # and the family
    assert Tu.get_entry("op", (256, 192, 256), "fp64") == (-1, 0, 0)
    Tu.set_entry("sweep", (256, 192, 256), "mixed", (0, 0, 0))                # removes
    assert Tu.get_entry("sweep", (256, 192, 256), "mixed") == (0, 0, 0)
    Tu.clear()
    assert Tu.get_entry("op", (256, 192, 256), "fp64") == (0, 0, 0)
    with pytest.raises(TbfgError):
        Tu.check(Tu.lib().tbfg_tune_set(7, Tu._i3((1, 1, 1)), 4, Tu._i3((1, 1, 1))))  # no such family


def test_candidates_tile_the_level():
    assert Tu.candidates("sweep", (256, 192, 256), "mixed") == Tu.SWEEP_F
    assert Tu.candidates("sweep", (96, 32, 96), "mixed") == []              # below the plane-march size
    assert Tu.candidates("sweep", (192, 96, 128), "mixed") == [(64, 8, 256), (32, 16, 256), (64, 16, 512),
                                                                (32, 8, 128)]
    assert (64, 4, 128) in Tu.candidates("sweep", (256, 192, 256), "fp64")
    ops = Tu.candidates("op", (96, 40, 64), "mixed")
    assert (-1, 0, 0) in ops and (128, 512, 0) not in ops and (32, 256, 0) in ops


def test_json_tables(tmp_path):
    Tu.clear()
    table = {"precision": "mixed", "entries": {"128,96,128": {"op": {"shape": [64, 512, 0], "ms": 1.0, "tried": {}},
                                                             "sweep": {"shape": [32, 16, 256], "ms": 1.0,
                                                                       "tried": {}}}}}
    p = str(tmp_path / "t.json")
    Tu.save(table, p)
    assert Tu.load(p) == table
    assert Tu.get_entry("op", (128, 96, 128), "mixed") == (64, 512, 0)
    assert Tu.get_entry("sweep", (128, 96, 128), "mixed") == (32, 16, 256)
    Tu.clear()
