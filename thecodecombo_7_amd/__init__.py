"""MI355X-native nonbonded path of TheCodeCombo_7 (Verlet list, LJ + real-space
Ewald pair kernel, smooth PME, velocity Verlet) as hand-written HIP behind a
C ABI (include/ttc.h, libttc.so).

Import is cheap: the HIP library is loaded on first use by `nonbond.Box`;
a missing library raises ImportError there (no CPU fallback exists).
"""
from . import system  # noqa: F401

__all__ = ["system", "nonbond", "build"]
