"""lcodepy_amd: the quasistatic 3D xi step of lcodePy on AMD MI355X (gfx950).

The plasma push, the deposit of rho/j, the DST/DCT field solves and the beam
push/deposit run as hand-written HIP kernels (liblcodepy_amd.so, C ABI in
include/lcodepy_amd.h); this package mirrors the reference's Python
interface around them: Config, Simulation, PusherAndSolver3D,
Plane2d3vPlasmaSolver and the plasma3d / beam3d factories."""
__version__ = '0.1.0'


def __getattr__(name):
    if name == 'Simulation':
        from .simulation import Simulation
        return Simulation
    if name == 'Config':
        from .config import Config
        return Config
    raise AttributeError(name)
