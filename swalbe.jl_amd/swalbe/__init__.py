"""swalbe -- Swalbe.jl's thin-film lattice Boltzmann step (2D D2Q9, 1D D1Q3) on MI355X.

The Python face of libswalbe_hip.so: the reference's parameter structs, a
GPU-resident state and the ops / time loop / run_* drivers by the
reference's names (without the trailing `!`).

    import swalbe as Swalbe
    sys = Swalbe.SysConst(Lx=256, Ly=256, param=Swalbe.Taumucs(Tmax=1000))
    state = Swalbe.Sys(sys, "GPU")
    Swalbe.equilibrium(state, sys)
    Swalbe.time_loop(sys, state)
"""
from ._lib import DomainError, SwalbeError, device_count, lib
from .core import (BGKandStream, DeviceArray, HipState, Sys, SysConst, Taumucs, cospi, equilibrium,
                   fast_32, fast_93, filmpressure, fluid_dry, force_sum, grad_f, grad_f_simple,
                   h_grad_p, h_grad_p_raw, inclination, laplacian_f, measure, moments, power_broad,
                   slippage, store_mode, surfacearea, t0, thermal, time_loop, wetted)
from .initialvalues import (Ellipsoid, VPolygon, boxpattern, ellipsepattern, normals, randinterface,
                            random_film, rivulet, sinewave, sinewave1D, singledroplet, torus,
                            trianglepattern, two_droplets_2d, uniforms)
from .batch import (INIT_CAP, INIT_FLAT, INIT_NORMAL, INIT_TWO_CAPS, Batch, FilmInit, device_memory,
                    film_bytes, film_init, tune_shape)
from .io import load_state, restart_from_height, save_snapshots, save_state, snapshot
from .simulate import (run_dropletforced, run_dropletpatterned, run_dropletrelax, run_flat,
                       run_gamma, run_random, run_rayleightaylor)
from . import one_d
from .one_d import (HipState1D, SysConst_1D, SysConstWithBound_1D, ensemble_time_loop, grad_gamma,
                    obslist, obslist1D, resident_max_length, two_droplets)

lib()  # the HIP library is required: fail at import if it is missing

__version__ = "1.0.0"
