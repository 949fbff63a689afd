"""Hartree atomic units.

The constants are those of the reference (units.py:64-74) so that time
steps, masses, temperatures and energies mean the same thing in both.
"""

from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class PhysicalUnits:
    BOLTZMANN_CONSTANT: float
    TO_ANGSTROM: float
    TO_PICO_SECOND: float
    TO_BAR: float
    TO_ELECTRON_VOLT: float
    TO_ATOMIC_MASS: float
    TO_FEMTO_SECOND: float = field(init=False)
    TO_KILO_BAR: float = field(init=False)
    FROM_ANGSTROM: float = field(init=False)
    FROM_FEMTO_SECOND: float = field(init=False)
    FROM_ELECTRON_VOLT: float = field(init=False)
    FROM_ATOMIC_MASS: float = field(init=False)
    TO_MILLI_ELECTRON_VOLT: float = field(init=False)

    def __post_init__(self) -> None:
        self.TO_FEMTO_SECOND = self.TO_PICO_SECOND * 1000
        self.TO_KILO_BAR = self.TO_BAR * 0.001
        self.FROM_ANGSTROM = 1 / self.TO_ANGSTROM
        self.FROM_FEMTO_SECOND = 1 / self.TO_FEMTO_SECOND
        self.FROM_ELECTRON_VOLT = 1 / self.TO_ELECTRON_VOLT
        self.FROM_ATOMIC_MASS = 1 / self.TO_ATOMIC_MASS
        self.TO_MILLI_ELECTRON_VOLT = self.TO_ELECTRON_VOLT * 1000


units = PhysicalUnits(
    BOLTZMANN_CONSTANT=3.166811563e-6,
    TO_ANGSTROM=5.29177249e-01,
    TO_PICO_SECOND=2.418884326e-05,
    TO_BAR=2.942102648e08,
    TO_ELECTRON_VOLT=2.7211407953e01,
    TO_ATOMIC_MASS=5.48579957163e-4,
)
