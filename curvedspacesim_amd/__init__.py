"""curvedSpaceSim geodesic force path on MI355X (gfx950): exact polyhedral geodesics, pair forces and the
displacement walk as hand-written HIP kernels behind a C-ABI (include/csg.h), with a C++ host layer that keeps the
reference's simulation/model/force/updater classes (include/csg_host.h, curvedspacesim_amd/csrc/host/)."""
from ._native import (CsgError, Context, System, bumpy_sphere, icosphere, jittered_torus, read_off, HARMONIC, GAUSSIAN,
                      TRUNC_GAUSSIAN, NO_POTENTIAL, PAIR_UNREACHED, PAIR_BEYOND, PAIR_SENTINEL, SCREEN_FEATURES,
                      SCREEN_PARAMS, LDS_VARIANTS, LDS_VARIANT_NAMES, save_checkpoint, load_checkpoint, ERR_ARGUMENT, ERR_DEVICE,
                      ERR_MESH, ERR_STATE, ERR_WALK, ERR_CAPACITY, lookup_tuning, record_tuning)

__all__ = ["CsgError", "Context", "System", "bumpy_sphere", "icosphere", "jittered_torus", "read_off", "HARMONIC", "GAUSSIAN",
           "TRUNC_GAUSSIAN", "NO_POTENTIAL", "PAIR_UNREACHED", "PAIR_BEYOND", "PAIR_SENTINEL", "SCREEN_FEATURES",
           "SCREEN_PARAMS", "LDS_VARIANTS", "LDS_VARIANT_NAMES", "save_checkpoint", "load_checkpoint", "ERR_ARGUMENT",
           "ERR_DEVICE", "ERR_MESH", "ERR_STATE", "ERR_WALK", "ERR_CAPACITY", "lookup_tuning", "record_tuning"]
