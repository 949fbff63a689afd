"""RuNNer ``input.data`` reader (datasets/runner.py:12-140, datasets/dataset.py:24-70).

``read_runner_blocks`` parses the text on the host into plain dicts; the
Dataset returns GPU ``Structure`` objects built from them.
"""

from __future__ import annotations

from collections import defaultdict
from pathlib import Path
from typing import Dict, Iterator, List, Optional


def _tokenize(line: str):
    tok = line.split()
    if not tok:
        return None, []
    return tok[0].lower(), tok[1:]


def read_runner_blocks(filename) -> Iterator[Dict[str, List]]:
    """Yield one dict per ``begin`` ... ``end`` block (runner.py:89-123)."""
    data: Optional[Dict[str, List]] = None
    with open(str(filename)) as f:
        for line in f:
            key, tok = _tokenize(line)
            if key == "begin":
                data = defaultdict(list)
            elif data is None:
                continue
            elif key == "atom":
                data["positions"].append([float(t) for t in tok[:3]])
                data["elements"].append(tok[3])
                data["charges"].append(float(tok[4]))
                data["energies"].append(float(tok[5]))
                data["forces"].append([float(t) for t in tok[6:9]])
            elif key == "lattice":
                data["lattice"].append([float(t) for t in tok[:3]])
            elif key == "energy":
                data["total_energy"].append(float(tok[0]))
            elif key == "charge":
                data["total_charge"].append(float(tok[0]))
            elif key == "comment":
                data["comment"].append(" ".join(line.split()[1:]))
            elif key == "end":
                yield dict(data)
                data = None


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
