"""Loader and ctypes signatures of liblcpao_grid.so (the C ABI in include/lcpao_grid.h).

The shared library is built in-tree by ``csrc/Makefile`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for the contractions.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblcpao_grid.so")

LCPAO_OK, LCPAO_EINVAL, LCPAO_ERANGE, LCPAO_EDEVICE, LCPAO_ESTATE, LCPAO_EMESH = range(6)
FP64, FP32, BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
lp = C.POINTER(C.c_int64)
dpp = C.POINTER(dp)
ipp = C.POINTER(ip)


class Geom(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("atom_base", C.c_int32), ("ngrid", C.c_int32 * 3), ("cpycell", C.c_int32),
        ("gridvol", C.c_double), ("no", ip), ("gridn", ip), ("grid_list", ipp), ("cell_list", ipp),
        ("npairs", C.c_int64), ("pair_atom", ip), ("pair_natn", ip), ("pair_ncn", ip), ("numolg", ip),
        ("olg1", ipp), ("olg2", ipp), ("latvecs", dp), ("origin", dp), ("gxyz", dp), ("atom2spe", ip),
    ]


class Pao(C.Structure):
    _fields_ = [("maxl", C.c_int32), ("num_basis", C.c_int32 * 4), ("nmesh", C.c_int32), ("rv", dp), ("rwf", dpp)]


class AtomicRadial(C.Structure):
    _fields_ = [("nmesh_pao", C.c_int32), ("pao_xv", dp), ("pao_rv", dp), ("atomic_den", dp),
                ("nmesh_vps", C.c_int32), ("vps_xv", dp), ("vps_rv", dp), ("atomic_pcc", dp)]


# every extern "C" symbol declared in include/lcpao_grid.h
EXPORTS = [
    "lcpao_grid_create", "lcpao_grid_set_tables", "lcpao_grid_set_dtables", "lcpao_grid_build_tables",
    "lcpao_grid_set_positions", "lcpao_grid_get_tables", "lcpao_grid_hamiltonian", "lcpao_grid_density", "lcpao_grid_force3",
    "lcpao_grid_hamiltonian_dev", "lcpao_grid_density_dev", "lcpao_grid_force3_dev", "lcpao_grid_force3_dev_noalloc",
    "lcpao_grid_set_stream",
    "lcpao_grid_profile", "lcpao_grid_kernel_stats", "lcpao_grid_tune", "lcpao_grid_tuned", "lcpao_grid_sizes",
    "lcpao_grid_last_error", "lcpao_grid_destroy", "lcpao_ucell_build", "lcpao_ucell_scalars",
    "lcpao_ucell_copy", "lcpao_ucell_last_error", "lcpao_ucell_destroy",
    "lcpao_batch_create", "lcpao_batch_assemble", "lcpao_batch_set_precision", "lcpao_batch_tune", "lcpao_batch_tuned", "lcpao_batch_sizes",
    "lcpao_batch_last_error", "lcpao_batch_destroy",
    "lcpao_grid_aden_pcc", "lcpao_grid_aden_pcc_dev", "lcpao_grid_poisson", "lcpao_grid_poisson_dev",
    "lcpao_grid_bloch_sum", "lcpao_grid_bloch_sum_dev", "lcpao_grid_bloch_sizes", "lcpao_grid_dm_states",
    "lcpao_grid_dm_states_dev", "lcpao_grid_bloch_sum_nc", "lcpao_grid_bloch_sum_nc_dev", "lcpao_grid_dm_states_nc",
    "lcpao_grid_dm_states_nc_dev", "lcpao_grid_tune_load", "lcpao_grid_tune_save", "lcpao_grid_block_forms",
    "lcpao_grid_block_layout",
]

_lib = None


class GridError(RuntimeError):
    """Raised for a non-zero status of the C ABI (mirrors the reference's error(...))."""

    def __init__(self, code, msg):
        super().__init__("lcpao_grid error %d: %s" % (code, msg))
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("liblcpao_grid.so is not built (run __graft_entry__.build() or make -C lcpaodft.jl_amd/csrc)")
    # torch ships its own HIP and HSA runtimes under the same sonames.  Loaded
    # first, they are the ones this library binds to; loaded after it, torch
    # brings a second pair into the process and its runtime then finds no
    # GPU.  One runtime per process: torch (when installed) goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.lcpao_grid_create.argtypes = [C.POINTER(vp), C.POINTER(Geom), C.c_int]
    L.lcpao_grid_set_tables.argtypes = [vp, dpp]
    L.lcpao_grid_set_dtables.argtypes = [vp, dpp, dpp, dpp]
    L.lcpao_grid_build_tables.argtypes = [vp, C.c_int32, C.POINTER(Pao), C.c_int32]
    L.lcpao_grid_set_positions.argtypes = [vp, dp]
    L.lcpao_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    L.lcpao_batch_assemble.argtypes = [vp, dpp, dpp, dpp, dpp, C.c_int32]
    L.lcpao_batch_tune.argtypes = [vp, C.c_int32]
    L.lcpao_batch_set_precision.argtypes = [vp, C.c_int32]
    L.lcpao_batch_tuned.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
    L.lcpao_batch_sizes.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), lp]
    L.lcpao_batch_last_error.argtypes = [vp]
    L.lcpao_batch_last_error.restype = C.c_char_p
    L.lcpao_batch_destroy.argtypes = [vp]
    L.lcpao_batch_destroy.restype = None
    L.lcpao_grid_get_tables.argtypes = [vp, C.c_int32, dpp]
    L.lcpao_grid_hamiltonian.argtypes = [vp, C.c_int32, dpp, dpp, C.c_int32]
    L.lcpao_grid_density.argtypes = [vp, C.c_int32, dpp, C.c_int64, dpp, C.c_int32]
    L.lcpao_grid_force3.argtypes = [vp, C.c_int32, dpp, dpp, C.c_int64, dp, dp, C.c_int32]
    L.lcpao_grid_hamiltonian_dev.argtypes = [vp, C.c_int32, dpp, dpp, C.c_int32]
    L.lcpao_grid_density_dev.argtypes = [vp, C.c_int32, dpp, C.c_int64, dpp, C.c_int32]
    L.lcpao_grid_force3_dev.argtypes = [vp, C.c_int32, dpp, dpp, C.c_int64, dp, dp, C.c_int32]
    L.lcpao_grid_force3_dev_noalloc.argtypes = [vp, C.c_int32, dpp, dpp, C.c_int64, dp, dp]
    L.lcpao_grid_sizes.argtypes = [vp, lp, lp, lp, lp]
    L.lcpao_grid_block_forms.argtypes = [vp, ip]
    L.lcpao_grid_block_layout.argtypes = [C.POINTER(Geom), ip, ip, lp]
    ar = C.POINTER(AtomicRadial)
    L.lcpao_grid_aden_pcc.argtypes = [vp, C.c_int32, C.c_int32, ar, dp, dp, dp, dp, dpp]
    L.lcpao_grid_aden_pcc_dev.argtypes = [vp, C.c_int32, C.c_int32, ar, dp, dp, dp, dp, dpp]
    L.lcpao_grid_poisson.argtypes = [vp, C.c_int32, dpp, dp, dp, dp]
    L.lcpao_grid_poisson_dev.argtypes = [vp, C.c_int32, dpp, dp, dp, dp]
    L.lcpao_grid_bloch_sum.argtypes = [vp, dp, C.c_int32, dpp, dpp]
    L.lcpao_grid_bloch_sum_dev.argtypes = [vp, dp, C.c_int32, dpp, dpp]
    L.lcpao_grid_bloch_sizes.argtypes = [vp, lp, lp]
    L.lcpao_grid_dm_states.argtypes = [vp, C.c_int32, dp, C.c_int64, dpp, C.c_double, dp]
    L.lcpao_grid_dm_states_dev.argtypes = [vp, C.c_int32, dp, C.c_int64, dpp, C.c_double, dp]
    L.lcpao_grid_tune_load.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int32)]
    L.lcpao_grid_tune_save.argtypes = [vp, C.c_char_p]
    L.lcpao_grid_bloch_sum_nc.argtypes = [vp, dp, dpp, dpp, dp]
    L.lcpao_grid_bloch_sum_nc_dev.argtypes = [vp, dp, dpp, dpp, dp]
    L.lcpao_grid_dm_states_nc.argtypes = [vp, C.c_int32, dp, C.c_int64, dpp, C.c_double, dpp, dpp]
    L.lcpao_grid_dm_states_nc_dev.argtypes = [vp, C.c_int32, dp, C.c_int64, dpp, C.c_double, dpp, dpp]
    L.lcpao_grid_set_stream.argtypes = [vp, vp]
    L.lcpao_grid_profile.argtypes = [vp, C.c_int32]
    L.lcpao_grid_kernel_stats.argtypes = [vp, C.c_int32, lp, dp]
    L.lcpao_grid_tune.argtypes = [vp, C.c_int32, C.c_int32]
    L.lcpao_grid_tuned.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int32]
    L.lcpao_grid_last_error.argtypes = [vp]
    L.lcpao_grid_last_error.restype = C.c_char_p
    L.lcpao_grid_destroy.argtypes = [vp]
    L.lcpao_grid_destroy.restype = None
    L.lcpao_ucell_build.argtypes = [C.POINTER(vp), dp, C.c_int32, dp, dp, ip, ip, dp, C.c_int32, C.c_int32, C.c_int32]
    L.lcpao_ucell_scalars.argtypes = [vp, ip, lp, lp, lp, lp, lp, dp, lp, lp]
    L.lcpao_ucell_copy.argtypes = [vp] + [ip, lp, ip, ip, ip, lp, ip, ip, ip, ip, ip, lp, ip, lp, ip, ip]
    L.lcpao_ucell_last_error.restype = C.c_char_p
    L.lcpao_ucell_destroy.argtypes = [vp]
    L.lcpao_ucell_destroy.restype = None
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != LCPAO_OK:
        msg = load().lcpao_grid_last_error(ctx).decode()
        raise GridError(rc, msg)
