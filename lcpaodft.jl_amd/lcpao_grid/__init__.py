"""lcpao_grid -- MI355X-native real-space grid path of LCPAODFT.jl.

Python host-side mirror of the reference's grid routines over the HIP
library liblcpao_grid.so (C ABI: include/lcpao_grid.h).
"""
from ._lib import BF16X3, FP32, FP64, GridError, load  # noqa: F401
from .grid import (GridContext, Calc_MatrixElements_dVH_Vxc_nc, Calc_MatrixElements_dVH_Vxc_off,  # noqa: F401
                   Calc_MatrixElements_dVH_Vxc_on, Force3_nc, Force3_nospin, Force3_spin, Set_Density_Grid,
                   Set_Density_Grid_nc, Set_Density_Grid_nonpol, Set_Density_Grid_pol, Set_dOrbitals_Grid,
                   Set_Orbitals_Grid)
from .pao import PAO, Read_PAO  # noqa: F401
from .ucell import UCell  # noqa: F401
