"""adh_kraken_amd -- MI355X-native SW2 Newton step of AdH-Kraken behind the reference's C call surface.

The compute lives in libadhk.so (hand-written gfx950 HIP kernels + C host driver, include/adhk.h);
this package is the Python host-side mirror of the reference interface (``sw2``) and the build step.
"""
from . import _lib  # noqa: F401
from .sw2 import (FLUME_C1, FLUME_C2, FlumeSpec, Mesh, SModelSuper, apply_Dirichlet_BC,  # noqa: F401
                  assemble_jacobian, assemble_residual, check_diag, elem_jacobians, elem_jacobians_mixed,
                  elem_residuals, fe_newton,
                  fe_sw2_init, flume_ensemble, flume_mesh, get_system, series_average, sparsity,
                  time_loop)
from .channel import (C5_HYDROGRAPH, CHANNEL_C3, CHANNEL_C5, ChannelSpec, bind_hydrograph,  # noqa: F401,E402
                      channel_mesh, inflow_segments)
