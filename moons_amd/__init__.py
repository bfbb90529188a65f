"""moons_amd -- MI355X-native MOONS per-timestep hot path (HIP, gfx950).

The compute lives in libmoons_amd.so (C ABI: include/moons_amd.h); this
package is the Python host mirror of the reference's SF/mesh/PCG surface.
"""
from ._lib import LIB_PATH, MoonsError, load  # noqa: F401
from .core import (  # noqa: F401
    BC,
    DL,
    FFT_solver_SF,
    PCG_solver_SF,
    PCG_solver_VF,
    MHD_step,
    Energy_step,
    Stream,
    UNet,
    mfma_check,
    SF,
    Lap_uniform_SF,
    apply_BCs,
    apply_BCs_implicit,
    clean_div,
    clean_div_FFT,
    curl,
    del_op,
    div,
    grad,
    init,
    iter_solver_params,
    mesh,
    nodes_roberts_both,
    nodes_uniform,
    synchronize,
    volume,
    interp,
    edgeCrossFace_E,
    advect_B,
    export_raw,
    import_raw,
    gf_write,
    gf_read,
    raw_write,
    raw_read,
)
