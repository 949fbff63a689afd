"""RuNNer ``input.data`` reader (datasets/runner.py:12-140, datasets/dataset.py:24-70).

``read_runner_blocks`` parses the text on the host into plain dicts; the
Dataset returns GPU ``Structure`` objects built from them.
"""

from __future__ import annotations

from collections import defaultdict
from pathlib import Path
from typing import Dict, Iterator, List, Optional


# The RuNNer input.data format, one record per line between "begin" and "end":
#   atom x y z element charge energy fx fy fz | lattice ax ay az | energy E | charge Q | comment text
# Per record keyword: (output field, column slice, converter); a slice of width one gives a scalar.
_RECORDS = {
    "atom": (("positions", slice(0, 3), float), ("elements", slice(3, 4), str), ("charges", slice(4, 5), float),
             ("energies", slice(5, 6), float), ("forces", slice(6, 9), float)),
    "lattice": (("lattice", slice(0, 3), float),),
    "energy": (("total_energy", slice(0, 1), float),),
    "charge": (("total_charge", slice(0, 1), float),),
}


def _fields(columns, rest: List[str]):
    for name, sl, conv in columns:
        vals = [conv(t) for t in rest[sl]]
        yield name, vals[0] if sl.stop - sl.start == 1 else vals


def read_runner_blocks(filename) -> Iterator[Dict[str, List]]:
    """Yield one dict per ``begin`` ... ``end`` block (the records of runner.py:89-123)."""
    block: Optional[Dict[str, List]] = None
    for raw in Path(filename).read_text().splitlines():
        words = raw.split()
        if not words:
            continue
        key, rest = words[0].lower(), words[1:]
        if key == "begin":
            block = defaultdict(list)
        elif block is None:
            continue
        elif key == "end":
            yield dict(block)
            block = None
        elif key == "comment":
            block["comment"].append(" ".join(rest))
        elif key in _RECORDS:
            for name, value in _fields(_RECORDS[key], rest):
                block[name].append(value)


class RunnerDataSource:
    def __init__(self, filename, dtype=None):
        self.filename = Path(filename)
        self.dtype = dtype

    def __len__(self) -> int:
        return sum(1 for _ in read_runner_blocks(self.filename))

    def __getitem__(self, index: int):
        for i, block in enumerate(read_runner_blocks(self.filename)):
            if i == index:
                return self._to_structure(block)
        raise IndexError(f"The given index {index} is out of bound (len={len(self)})")

    def read_structures(self):
        for block in read_runner_blocks(self.filename):
            yield self._to_structure(block)

    def _to_structure(self, block):
        from .atoms import Structure

        return Structure.from_dict(block, dtype=self.dtype)


class Dataset:
    def __init__(self, datasource: RunnerDataSource, persist: bool = False):
        self.datasource = datasource
        self.persist = persist
        self.cache: Dict[int, object] = {}

    @classmethod
    def from_runner(cls, filename, persist: bool = False, dtype=None) -> "Dataset":
        return cls(RunnerDataSource(filename, dtype), persist)

    def __len__(self) -> int:
        return len(self.datasource)

    def __getitem__(self, index: int):
        if self.persist and index in self.cache:
            return self.cache[index]
        s = self.datasource[index]
        if self.persist:
            self.cache[index] = s
        return s

    def preload(self) -> None:
        self.persist = True
        for i, s in enumerate(self.datasource.read_structures()):
            self.cache[i] = s
