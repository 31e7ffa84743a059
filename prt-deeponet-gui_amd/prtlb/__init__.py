"""prtlb -- MI355X-native PRT-LB: D2Q9/D3Q19 TRT Stokes and the coupled
advection-diffusion-reaction stage, behind the reference's Python entry
points (shahram444/PRT-DeepONet-GUI, 3D/tools/prtlb_2d.py and prtlb_3d.py).

    from prtlb import stokes_d3q19, solve_adr, check_physics

Every solve runs on the GPU through libprtlb_hip.so; there is no CPU path.
"""
import numpy as np

from ._lib import PrtError
from .adr import TransportSolver, solve_adr
from .chemistry import Chemistry, default_chemistry
from .geometry import (PORE, SOLID, WALL, assert_finite_distance,
                       coat_with_biofilm, disk_pack, edt_device, geodesic_device,
                       keep_inlet_connected, sphere_pack)
from .physics import (check_physics, check_physics_device, outlet_concentration,
                      permeability)
from .stokes import (FlowConvergence, StokesSolver, plane_flux, stokes_d2q9,
                     stokes_d3q19, trt_minus)

# module constants the dataset builders import (3D/tools/prtlb_3d.py:491-515)
INLET = np.array([1.0, 1.0, 0.0, 0.0], np.float32)
KS_A = np.float32(0.15)
ALPHA = np.float32(0.4)
Y = np.float32(0.04)
KD = np.float32(0.006)
B0 = np.float32(0.1)
ATTACHED = 3

__all__ = [
    "SOLID", "WALL", "PORE", "INLET", "KS_A", "ALPHA", "Y", "KD", "B0", "ATTACHED",
    "stokes_d2q9", "stokes_d3q19", "solve_adr", "check_physics",
    "keep_inlet_connected", "coat_with_biofilm", "assert_finite_distance",
    "Chemistry", "default_chemistry", "permeability", "outlet_concentration",
    "StokesSolver", "TransportSolver", "FlowConvergence", "trt_minus", "plane_flux",
    "disk_pack", "sphere_pack", "PrtError", "check_physics_device", "geodesic_device",
    "edt_device",
]
