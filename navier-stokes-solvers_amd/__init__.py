"""ctypes binding of libnsspectral.so, the MI355X pseudospectral vorticity solver.

The library's C-ABI is the reference's spectral API (include/ns_spectral_solver.h,
mirroring src/spectral/ns_spectral_solver.h:70-113 of the reference) plus the
forcing, configuration and time-step helpers.  This module mirrors the
NSSpectralData struct so Python callers (tests, bench.py) can read and write
the host arrays exactly like C callers do (data->u, data->omega_hat, ...).

Loading fails loudly when the library is missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsspectral.so")
DRIVER_PATH = os.path.join(_HERE, "ns_spectral_run")

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)

SOLVER_SUCCESS, SOLVER_MAX_ITER, SOLVER_DIVERGED, SOLVER_MEM_ERROR, SOLVER_INVALID_PARAM = range(5)
NS_PRECISION_FP64, NS_PRECISION_FP32 = 0, 1
NS_SCHEME_RK4, NS_SCHEME_IF_RK4 = 0, 1
NS_FORCING_WHITE_NOISE, NS_FORCING_COLORED_NOISE = 0, 1
NSD_FP64, NSD_FP32 = 0, 1

# every entry point the headers in include/ declare and the library exports
EXPORTED = [
    # ns_spectral_solver.h (reference API)
    "ns_spectral_allocate", "ns_spectral_free", "ns_spectral_initialize_grids",
    "ns_spectral_initialize_fft_plans", "ns_spectral_initialize_taylor_green",
    "ns_spectral_physical_to_spectral", "ns_spectral_spectral_to_physical",
    "ns_spectral_apply_dealiasing", "ns_spectral_compute_derivatives",
    "ns_spectral_compute_velocity_from_vorticity", "ns_spectral_compute_nonlinear_terms",
    "ns_spectral_advance_rk4", "ns_spectral_compute_cfl_timestep", "ns_spectral_compute_energy",
    "ns_spectral_compute_enstrophy", "ns_spectral_analyze_convergence",
    "ns_spectral_analyze_spectrum", "ns_spectral_check_resolution",
    "ns_spectral_hyperviscosity_default", "ns_spectral_hyperviscosity_set",
    "ns_spectral_hyperviscosity_update_weights", "ns_spectral_hyperviscosity_apply",
    "ns_spectral_hyperviscosity_post_step", "ns_spectral_output_solution",
    # ns_spectral_solver.h (additions)
    "ns_spectral_last_status", "ns_spectral_last_error", "ns_spectral_set_precision",
    "ns_spectral_set_scheme", "ns_spectral_set_drag", "ns_spectral_set_forcing",
    "ns_spectral_initialize_random_phase", "ns_spectral_sync_device", "ns_spectral_advance_n",
    "ns_spectral_sync_host", "ns_spectral_device_diagnostics", "ns_spectral_configure",
    "ns_spectral_create_from_config", "ns_spectral_convergence_metrics",
    "ns_spectral_set_error_control", "ns_spectral_error_control_stats",
    # forcing.h
    "ns_forcing_config_default", "ns_forcing_state_init", "ns_forcing_state_free",
    "ns_forcing_generate",
    # ns_config.h
    "ns_config_init", "ns_config_free", "ns_config_set", "ns_config_get_string",
    "ns_config_get_int", "ns_config_get_double", "ns_config_get_bool", "ns_config_has_key",
    "ns_config_load_files", "ns_config_validate",
    # time_integration.h
    "compute_cfl_condition", "compute_viscous_timestep", "adaptive_timestep",
    # ns_device.h
    "nsd_create", "nsd_destroy", "nsd_error", "nsd_set_params", "nsd_precision",
    "nsd_state_upload", "nsd_state_download", "nsd_sfd_upload", "nsd_sfd_download",
    "nsd_state_random_phase", "nsd_set_time", "nsd_get_time", "nsd_set_forcing_counter",
    "nsd_get_forcing_counter", "nsd_advance", "nsd_advance_graph", "nsd_state_derived",
    "nsd_state_energy", "nsd_synchronize", "nsd_profile", "nsd_profile_read", "nsd_grids",
    "nsd_hv_weights", "nsd_physical_to_spectral", "nsd_spectral_to_physical", "nsd_dealias",
    "nsd_velocity", "nsd_derivatives", "nsd_nonlinear", "nsd_hv_apply", "nsd_hv_post_step",
    "nsd_taylor_green", "nsd_phys_reduce", "nsd_spec_reduce", "nsd_spectrum",
    "nsd_forcing_generate", "nsd_bandwidth_probe",
    # ns_device.h (batches)
    "nsd_create_batch", "nsd_batch_members", "nsd_set_member_params", "nsd_batch_random_phase",
    "nsd_batch_energy", "nsd_batch_scalars", "nsd_member_upload", "nsd_member_download",
    "nsd_multi_create", "nsd_multi_destroy", "nsd_multi_error", "nsd_multi_advance",
    "nsd_multi_synchronize", "nsd_autotune", "nsd_set_variants", "nsd_forcing_ou",
    "nsd_forcing_modes", "nsd_physical_to_spectral_noalloc", "nsd_errctl_stats",
    "nsd_row_stagger",
    # checkpoint.h
    "ns_checkpoint_write_spectral", "ns_checkpoint_read_spectral", "ns_checkpoint_error",
    # ns_closure.h
    "nsd_closure_create", "nsd_closure_destroy", "nsd_closure_error", "nsd_closure_param_count",
    "nsd_closure_params", "nsd_closure_forward", "nsd_closure_gradient", "nsd_closure_set_scales",
    "nsd_closure_get_scales", "nsd_closure_size", "nsd_closure_dataset_reserve",
    "nsd_closure_dataset_slot", "nsd_closure_dataset_fill", "nsd_closure_dataset_read",
    "nsd_closure_train",
    "nsd_closure_eval_device", "nsd_closure_eval_batch", "nsd_closure_max_batch",
    "nsd_attach_closure", "nsd_subgrid_sample", "nsd_subgrid_dataset",
    # vtk_writer.h
    "ns_vtk_write_spectral_legacy", "ns_vtk_write_spectral_xml", "ns_vtk_update_pvd",
    # ns_ensemble.h
    "ns_ensemble_config_default", "ns_ensemble_create", "ns_ensemble_create_error",
    "ns_ensemble_free", "ns_ensemble_members", "ns_ensemble_advance", "ns_ensemble_synchronize",
    "ns_ensemble_diagnostics", "ns_ensemble_member_state", "ns_ensemble_last_error",
]


class NSSpectralHyperviscosityConfig(C.Structure):
    _fields_ = [("enabled", C.c_int), ("coefficient", C.c_double), ("order", C.c_int),
                ("cutoff_ratio", C.c_double), ("sfd_strength", C.c_double),
                ("sfd_timescale", C.c_double), ("sfd_enabled", C.c_int), ("kmax", C.c_double),
                ("cutoff_wavenumber", C.c_double), ("inv_bandwidth", C.c_double)]


class NSSpectralData(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nkx", C.c_int), ("nky", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Re", C.c_double), ("dt", C.c_double),
                ("t", C.c_double), ("cfl_factor", C.c_double)] + \
               [(n, c_double_p) for n in ("u", "v", "p", "omega", "psi", "x", "y")] + \
               [(n, c_double_p) for n in ("u_hat", "v_hat", "omega_hat", "psi_hat",
                                          "nonlinear_hat", "rk1_hat", "rk2_hat", "rk3_hat",
                                          "rk4_hat", "omega_tmp_hat")] + \
               [("kx", c_double_p), ("ky", c_double_p), ("k2", c_double_p),
                ("forward_plan", C.c_void_p), ("backward_plan", C.c_void_p),
                ("dealias_mask", c_int_p),
                ("hyperviscosity", NSSpectralHyperviscosityConfig),
                ("hyperviscosity_weights", c_double_p), ("sfd_hat", c_double_p),
                ("device", C.c_void_p)]


class NSForcingConfig(C.Structure):
    _fields_ = [("kmin", C.c_double), ("kmax", C.c_double), ("energy_target", C.c_double),
                ("correlation_time", C.c_double), ("seed", C.c_ulong), ("type", C.c_int)]


class NSForcingState(C.Structure):
    _fields_ = [("config", NSForcingConfig), ("nkx", C.c_int), ("nky", C.c_int),
                ("spectral_size", C.c_int), ("active_modes", C.c_int),
                ("weights", c_double_p), ("colored_state", c_double_p),
                ("have_spare", C.c_uint), ("spare_gaussian", C.c_double),
                ("rng_state", C.c_ulonglong), ("active_index", c_int_p),
                ("calls", C.c_ulonglong)]


class NSConfig(C.Structure):
    _fields_ = [("entries", C.c_void_p), ("count", C.c_size_t), ("capacity", C.c_size_t)]


class NSConfigSchemaEntry(C.Structure):
    _fields_ = [("key", C.c_char_p), ("type", C.c_int), ("required", C.c_int),
                ("min_value", C.c_double), ("max_value", C.c_double), ("has_min", C.c_int),
                ("has_max", C.c_int), ("allowed_values", C.POINTER(C.c_char_p)),
                ("allowed_count", C.c_size_t)]


class AdaptiveTimeParams(C.Structure):
    _fields_ = [("cfl_factor", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("safety_factor", C.c_double), ("adapt_enabled", C.c_int)]


class NsdParams(C.Structure):  # ns_device.h
    _fields_ = [("Re", C.c_double), ("drag", C.c_double), ("cfl_factor", C.c_double),
                ("scheme", C.c_int), ("hv_on", C.c_int), ("hv_order", C.c_int),
                ("hv_coeff", C.c_double), ("hv_kc", C.c_double), ("hv_invbw", C.c_double),
                ("sfd_on", C.c_int), ("sfd_strength", C.c_double), ("sfd_tau", C.c_double),
                ("force_on", C.c_int), ("force_colored", C.c_int), ("force_kmin", C.c_double),
                ("force_kmax", C.c_double), ("force_energy_target", C.c_double),
                ("force_tau", C.c_double), ("force_rate", C.c_double),
                ("force_seed", C.c_ulonglong), ("errctl", C.c_int), ("err_rtol", C.c_double),
                ("err_atol", C.c_double), ("err_safety", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double)]


class NSEnsembleConfig(C.Structure):  # ns_ensemble.h
    _fields_ = [("Lx", C.c_double), ("Ly", C.c_double), ("precision", C.c_int),
                ("scheme", C.c_int), ("cfl_factor", C.c_double), ("drag", C.c_double),
                ("nu4", C.c_double), ("forcing_enabled", C.c_int), ("forcing", NSForcingConfig),
                ("injection_rate", C.c_double), ("k_peak", C.c_double), ("energy", C.c_double),
                ("autotune", C.c_int)]


_lib = None


def load():
    """Load libnsspectral.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library missing: {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(NSSpectralData)
    L.ns_spectral_allocate.restype = P
    L.ns_spectral_allocate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
    for n in ("ns_spectral_free", "ns_spectral_initialize_grids",
              "ns_spectral_initialize_fft_plans", "ns_spectral_initialize_taylor_green",
              "ns_spectral_compute_derivatives", "ns_spectral_compute_velocity_from_vorticity",
              "ns_spectral_compute_nonlinear_terms", "ns_spectral_advance_rk4",
              "ns_spectral_analyze_spectrum", "ns_spectral_hyperviscosity_update_weights"):
        getattr(L, n).argtypes = [P]
        getattr(L, n).restype = None
    for n in ("ns_spectral_compute_cfl_timestep", "ns_spectral_compute_energy",
              "ns_spectral_compute_enstrophy"):
        getattr(L, n).argtypes = [P]
        getattr(L, n).restype = C.c_double
    L.ns_spectral_check_resolution.argtypes = [P]
    L.ns_spectral_check_resolution.restype = C.c_int
    L.ns_spectral_analyze_convergence.argtypes = [P, C.c_int]
    L.ns_spectral_physical_to_spectral.argtypes = [P, c_double_p, c_double_p]
    L.ns_spectral_spectral_to_physical.argtypes = [P, c_double_p, c_double_p]
    L.ns_spectral_apply_dealiasing.argtypes = [P, c_double_p]
    HVP = C.POINTER(NSSpectralHyperviscosityConfig)
    L.ns_spectral_hyperviscosity_default.argtypes = [HVP]
    L.ns_spectral_hyperviscosity_set.argtypes = [P, HVP]
    L.ns_spectral_hyperviscosity_apply.argtypes = [P, c_double_p, c_double_p]
    L.ns_spectral_hyperviscosity_post_step.argtypes = [P, c_double_p, C.c_double]
    L.ns_spectral_output_solution.argtypes = [P, C.c_char_p]
    L.ns_spectral_last_status.argtypes = [P]
    L.ns_spectral_last_status.restype = C.c_int
    L.ns_spectral_last_error.argtypes = [P]
    L.ns_spectral_last_error.restype = C.c_char_p
    for n, extra in (("ns_spectral_set_precision", [C.c_int]), ("ns_spectral_set_scheme", [C.c_int]),
                     ("ns_spectral_set_drag", [C.c_double]),
                     ("ns_spectral_set_forcing", [C.POINTER(NSForcingConfig), C.c_double]),
                     ("ns_spectral_initialize_random_phase",
                      [C.c_ulonglong, C.c_double, C.c_double]),
                     ("ns_spectral_sync_device", []), ("ns_spectral_sync_host", []),
                     ("ns_spectral_advance_n", [C.c_int, C.c_int]),
                     ("ns_spectral_device_diagnostics", [c_double_p, c_double_p]),
                     ("ns_spectral_convergence_metrics", [c_double_p]),
                     ("ns_spectral_set_error_control",
                      [C.POINTER(AdaptiveTimeParams), C.c_double, C.c_double]),
                     ("ns_spectral_error_control_stats",
                      [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong), c_double_p,
                       c_double_p]),
                     ("ns_spectral_configure", [C.POINTER(NSConfig)])):
        getattr(L, n).argtypes = [P] + extra
        getattr(L, n).restype = C.c_int
    L.ns_spectral_create_from_config.argtypes = [C.POINTER(NSConfig)]
    L.ns_spectral_create_from_config.restype = P
    FS = C.POINTER(NSForcingState)
    L.ns_forcing_config_default.argtypes = [C.POINTER(NSForcingConfig)]
    L.ns_forcing_state_init.argtypes = [FS, C.POINTER(NSForcingConfig), c_double_p, C.c_int,
                                        C.c_int]
    L.ns_forcing_state_init.restype = C.c_int
    L.ns_forcing_state_free.argtypes = [FS]
    L.ns_forcing_generate.argtypes = [FS, C.c_double, c_double_p, c_double_p]
    CP = C.POINTER(NSConfig)
    L.ns_config_init.argtypes = [CP]
    L.ns_config_free.argtypes = [CP]
    L.ns_config_set.argtypes = [CP, C.c_char_p, C.c_char_p]
    L.ns_config_set.restype = C.c_int
    L.ns_config_get_string.argtypes = [CP, C.c_char_p, C.c_char_p]
    L.ns_config_get_string.restype = C.c_char_p
    L.ns_config_get_int.argtypes = [CP, C.c_char_p, C.c_int]
    L.ns_config_get_int.restype = C.c_int
    L.ns_config_get_double.argtypes = [CP, C.c_char_p, C.c_double]
    L.ns_config_get_double.restype = C.c_double
    L.ns_config_get_bool.argtypes = [CP, C.c_char_p, C.c_int]
    L.ns_config_get_bool.restype = C.c_int
    L.ns_config_has_key.argtypes = [CP, C.c_char_p]
    L.ns_config_has_key.restype = C.c_int
    L.ns_config_load_files.argtypes = [CP, C.POINTER(C.c_char_p), C.c_size_t, C.c_char_p,
                                       C.c_size_t]
    L.ns_config_load_files.restype = C.c_int
    L.ns_config_validate.argtypes = [CP, C.POINTER(NSConfigSchemaEntry), C.c_size_t, C.c_char_p,
                                     C.c_size_t]
    L.ns_config_validate.restype = C.c_int
    L.compute_cfl_condition.argtypes = [C.c_double] * 5
    L.compute_cfl_condition.restype = C.c_double
    L.compute_viscous_timestep.argtypes = [C.c_double] * 3
    L.compute_viscous_timestep.restype = C.c_double
    L.adaptive_timestep.argtypes = [C.c_double] * 4 + [C.POINTER(AdaptiveTimeParams)]
    L.adaptive_timestep.restype = C.c_double
    # ns_device.h (used by bench.py for kernel timing)
    L.nsd_profile.argtypes = [C.c_void_p, C.c_int]
    L.nsd_profile.restype = C.c_int
    L.nsd_profile_read.argtypes = [C.c_void_p, c_double_p, c_double_p, c_double_p, c_int_p]
    L.nsd_profile_read.restype = C.c_int
    L.nsd_advance.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.nsd_advance.restype = C.c_int
    L.nsd_advance_graph.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.nsd_advance_graph.restype = C.c_int
    L.nsd_bandwidth_probe.argtypes = [C.c_void_p, C.c_size_t, C.c_int, c_double_p]
    L.nsd_bandwidth_probe.restype = C.c_int
    L.nsd_synchronize.argtypes = [C.c_void_p]
    L.nsd_synchronize.restype = C.c_int
    V, D, I = C.c_void_p, c_double_p, c_int_p
    ull = C.c_ulonglong
    dev = {
        "nsd_create": ([C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_char_p,
                        C.c_size_t], V),
        "nsd_destroy": ([V], None),
        "nsd_error": ([V], C.c_char_p),
        "nsd_set_params": ([V, C.POINTER(NsdParams)], C.c_int),
        "nsd_precision": ([V], C.c_int),
        "nsd_state_upload": ([V, D], C.c_int),
        "nsd_state_download": ([V, D], C.c_int),
        "nsd_sfd_upload": ([V, D], C.c_int),
        "nsd_sfd_download": ([V, D], C.c_int),
        "nsd_state_random_phase": ([V, ull, C.c_double, C.c_double], C.c_int),
        "nsd_set_time": ([V, C.c_double, C.c_double], C.c_int),
        "nsd_get_time": ([V, D, D, I], C.c_int),
        "nsd_set_forcing_counter": ([V, ull], C.c_int),
        "nsd_get_forcing_counter": ([V, C.POINTER(ull)], C.c_int),
        "nsd_state_derived": ([V, D, D, D, D, D, D], C.c_int),
        "nsd_state_energy": ([V, D], C.c_int),
        "nsd_grids": ([V, D, D, D, D, D, I], C.c_int),
        "nsd_hv_weights": ([V, C.c_int, C.c_double, C.c_int, C.c_double, D, D, D, D], C.c_int),
        "nsd_physical_to_spectral": ([V, D, D], C.c_int),
        "nsd_spectral_to_physical": ([V, D, D], C.c_int),
        "nsd_dealias": ([V, D], C.c_int),
        "nsd_velocity": ([V, D, D, D, D], C.c_int),
        "nsd_derivatives": ([V, D, D, D, D, D, D, D], C.c_int),
        "nsd_nonlinear": ([V, D, D, D], C.c_int),
        "nsd_hv_apply": ([V, D, D, D], C.c_int),
        "nsd_hv_post_step": ([V, D, D, C.c_double], C.c_int),
        "nsd_taylor_green": ([V, D, D, D, D], C.c_int),
        "nsd_phys_reduce": ([V, D, D, D, D], C.c_int),
        "nsd_spec_reduce": ([V, D, D], C.c_int),
        "nsd_spectrum": ([V, D, D, D, I, C.c_int], C.c_int),
        "nsd_forcing_generate": ([C.c_int, I, C.c_int, ull, ull, C.c_int, C.c_double,
                                  C.c_double, C.c_double, D, D, C.c_char_p, C.c_size_t],
                                 C.c_int),
    }
    ULLP = C.POINTER(ull)
    dev.update({
        "nsd_create_batch": ([C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                              C.c_char_p, C.c_size_t], V),
        "nsd_batch_members": ([V], C.c_int),
        "nsd_set_member_params": ([V, D, ULLP], C.c_int),
        "nsd_batch_random_phase": ([V, ULLP, C.c_double, C.c_double], C.c_int),
        "nsd_batch_energy": ([V, D, D], C.c_int),
        "nsd_batch_scalars": ([V, D, D, I], C.c_int),
        "nsd_member_upload": ([V, C.c_int, D], C.c_int),
        "nsd_member_download": ([V, C.c_int, D], C.c_int),
        "nsd_multi_create": ([C.POINTER(V), C.c_int], V),
        "nsd_multi_destroy": ([V], None),
        "nsd_multi_error": ([V], C.c_char_p),
        "nsd_multi_advance": ([V, C.c_int, C.c_int], C.c_int),
        "nsd_multi_synchronize": ([V], C.c_int),
        "nsd_autotune": ([V, C.c_int, C.c_int, I, I], C.c_int),
        "nsd_set_variants": ([V, C.c_int, C.c_int], C.c_int),
        "nsd_forcing_ou": ([V, D, C.c_int, C.c_int], C.c_int),
        "nsd_forcing_modes": ([V], C.c_int),
        "nsd_physical_to_spectral_noalloc": ([C.c_int, C.c_int, V, V, V, C.c_char_p,
                                              C.c_size_t], C.c_int),
        "nsd_errctl_stats": ([V, C.c_int, ULLP, ULLP, D, D], C.c_int),
        "nsd_row_stagger": ([V, C.c_int], C.c_int),
        "ns_checkpoint_write_spectral": ([C.c_char_p, C.POINTER(NSSpectralData)], C.c_int),
        "ns_checkpoint_read_spectral": ([C.c_char_p, C.POINTER(NSSpectralData)], C.c_int),
        "ns_checkpoint_error": ([], C.c_char_p),
        "ns_vtk_write_spectral_legacy": ([C.c_char_p, C.POINTER(NSSpectralData), C.c_double],
                                         C.c_int),
        "ns_vtk_write_spectral_xml": ([C.c_char_p, C.POINTER(NSSpectralData), C.c_double],
                                      C.c_int),
        "ns_vtk_update_pvd": ([C.c_char_p, C.c_char_p, C.c_double], C.c_int),
        "nsd_closure_create": ([C.c_int, C.c_int, ull, C.c_char_p, C.c_size_t], V),
        "nsd_closure_destroy": ([V], None),
        "nsd_closure_error": ([V], C.c_char_p),
        "nsd_closure_param_count": ([V], C.c_int),
        "nsd_closure_params": ([V, C.POINTER(C.c_float), C.c_int], C.c_int),
        "nsd_closure_forward": ([V, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)],
                                C.c_int),
        "nsd_closure_gradient": ([V, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                  D, C.POINTER(C.c_float)], C.c_int),
        "nsd_closure_set_scales": ([V, C.c_double, C.c_double], C.c_int),
        "nsd_closure_get_scales": ([V, D, D], C.c_int),
        "nsd_closure_size": ([V], C.c_int),
        "nsd_closure_dataset_reserve": ([V, C.c_int], C.c_int),
        "nsd_closure_dataset_slot": ([V, C.c_int, C.POINTER(V), C.POINTER(V)], C.c_int),
        "nsd_closure_dataset_fill": ([V, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)],
                                     C.c_int),
        "nsd_closure_dataset_read": ([V, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)],
                                     C.c_int),
        "nsd_closure_train": ([V, C.c_int, C.c_int, C.c_int, C.c_double, ull, D, D, D], C.c_int),
        "nsd_closure_eval_device": ([V, V, V, V], C.c_int),
        "nsd_closure_eval_batch": ([V, V, V, C.c_int, V], C.c_int),
        "nsd_closure_max_batch": ([V], C.c_int),
        "nsd_attach_closure": ([V, V], C.c_int),
        "nsd_subgrid_sample": ([V, C.c_int, V, V, V], C.c_int),
        "nsd_subgrid_dataset": ([V, V, V, C.c_int, C.c_int, C.c_int], C.c_int),
        "ns_ensemble_config_default": ([C.POINTER(NSEnsembleConfig)], None),
        "ns_ensemble_create": ([C.c_int, I, D, ULLP, C.POINTER(NSEnsembleConfig)], V),
        "ns_ensemble_create_error": ([], C.c_char_p),
        "ns_ensemble_free": ([V], None),
        "ns_ensemble_members": ([V], C.c_int),
        "ns_ensemble_advance": ([V, C.c_int], C.c_int),
        "ns_ensemble_synchronize": ([V], C.c_int),
        "ns_ensemble_diagnostics": ([V, D, D, D, I], C.c_int),
        "ns_ensemble_member_state": ([V, C.c_int, D], C.c_int),
        "ns_ensemble_last_error": ([V], C.c_char_p),
    })
    for name, (args, res) in dev.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = res
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(c_double_p)


class Spectral:
    """One solver instance; numpy views share memory with the C host arrays."""

    def __init__(self, nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, Re=100.0):
        self.L = load()
        self.p = self.L.ns_spectral_allocate(nx, ny, Lx, Ly, Re)
        if not self.p:
            raise RuntimeError("ns_spectral_allocate failed (no HIP device or bad size)")
        self.d = self.p.contents
        self.nx, self.ny, self.nkx = nx, ny, nx // 2 + 1

    def close(self):
        if getattr(self, "p", None):
            self.L.ns_spectral_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- views
    def phys(self, name):
        return np.ctypeslib.as_array(getattr(self.d, name), shape=(self.ny, self.nx))

    def spec(self, name):
        a = np.ctypeslib.as_array(getattr(self.d, name), shape=(self.ny, self.nkx * 2))
        return a.view(np.complex128)

    @property
    def kx(self):
        return np.ctypeslib.as_array(self.d.kx, shape=(self.nkx,))

    @property
    def ky(self):
        return np.ctypeslib.as_array(self.d.ky, shape=(self.ny,))

    @property
    def k2(self):
        return np.ctypeslib.as_array(self.d.k2, shape=(self.ny, self.nkx))

    @property
    def mask(self):
        return np.ctypeslib.as_array(self.d.dealias_mask, shape=(self.ny, self.nkx))

    @property
    def hv_weights(self):
        return np.ctypeslib.as_array(self.d.hyperviscosity_weights, shape=(self.ny, self.nkx))

    @property
    def xgrid(self):
        return np.ctypeslib.as_array(self.d.x, shape=(self.nx,))

    @property
    def ygrid(self):
        return np.ctypeslib.as_array(self.d.y, shape=(self.ny,))

    def status(self):
        return self.L.ns_spectral_last_status(self.p)

    def error(self):
        return self.L.ns_spectral_last_error(self.p).decode()

    def check(self, rc=None):
        st = self.status() if rc is None else rc
        if st != SOLVER_SUCCESS:
            raise RuntimeError(f"solver status {st}: {self.error()}")

    # ---- reference API
    def init(self, taylor_green=True, derivatives=True):
        self.L.ns_spectral_initialize_grids(self.p)
        self.L.ns_spectral_initialize_fft_plans(self.p)
        if taylor_green:
            self.L.ns_spectral_initialize_taylor_green(self.p)
            if derivatives:
                self.L.ns_spectral_compute_derivatives(self.p)
        self.check()

    def p2s(self, field, out=None):
        field = np.ascontiguousarray(field, dtype=np.float64)
        out = np.zeros((self.ny, self.nkx), np.complex128) if out is None else out
        self.L.ns_spectral_physical_to_spectral(self.p, _dp(field), _dp(out))
        self.check()
        return out

    def s2p(self, field_hat):
        fh = np.ascontiguousarray(field_hat, dtype=np.complex128)
        out = np.zeros((self.ny, self.nx))
        self.L.ns_spectral_spectral_to_physical(self.p, _dp(fh), _dp(out))
        self.check()
        return out

    def dealias(self, field_hat):
        fh = np.ascontiguousarray(field_hat, dtype=np.complex128).copy()
        self.L.ns_spectral_apply_dealiasing(self.p, _dp(fh))
        self.check()
        return fh

    def advance_rk4(self):
        self.L.ns_spectral_advance_rk4(self.p)
        return self.status()

    def cfl(self):
        return self.L.ns_spectral_compute_cfl_timestep(self.p)

    def energy(self):
        return self.L.ns_spectral_compute_energy(self.p)

    def enstrophy(self):
        return self.L.ns_spectral_compute_enstrophy(self.p)

    def hv_set(self, **kw):
        c = NSSpectralHyperviscosityConfig()
        self.L.ns_spectral_hyperviscosity_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        self.L.ns_spectral_hyperviscosity_set(self.p, C.byref(c))
        self.check()

    # ---- additions
    def set_precision(self, fp32):
        self.check(self.L.ns_spectral_set_precision(self.p, NS_PRECISION_FP32 if fp32
                                                    else NS_PRECISION_FP64))

    def set_scheme(self, integrating_factor):
        self.check(self.L.ns_spectral_set_scheme(self.p, NS_SCHEME_IF_RK4 if integrating_factor
                                                 else NS_SCHEME_RK4))

    def set_drag(self, mu):
        self.check(self.L.ns_spectral_set_drag(self.p, mu))

    def set_forcing(self, kmin, kmax, rate, seed=42, colored=False, energy_target=1.0,
                    correlation_time=0.5):
        f = NSForcingConfig(kmin, kmax, energy_target, correlation_time, seed,
                            NS_FORCING_COLORED_NOISE if colored else NS_FORCING_WHITE_NOISE)
        self.check(self.L.ns_spectral_set_forcing(self.p, C.byref(f), rate))

    def random_phase(self, seed, k_peak=10.0, energy=1.0):
        self.check(self.L.ns_spectral_initialize_random_phase(self.p, seed, k_peak, energy))

    def sync_device(self):
        self.check(self.L.ns_spectral_sync_device(self.p))

    def advance_n(self, n, adaptive):
        return self.L.ns_spectral_advance_n(self.p, n, 1 if adaptive else 0)

    def sync_host(self):
        return self.L.ns_spectral_sync_host(self.p)

    def device_diagnostics(self):
        e, z = C.c_double(), C.c_double()
        self.check(self.L.ns_spectral_device_diagnostics(self.p, C.byref(e), C.byref(z)))
        return e.value, z.value

    def set_error_control(self, rtol, atol=0.0, safety=0.9, dt_min=1e-8, dt_max=0.1,
                          cfl_factor=None, enabled=True):
        p = AdaptiveTimeParams(self.d.cfl_factor if cfl_factor is None else cfl_factor,
                               dt_min, dt_max, safety, 1 if enabled else 0)
        self.check(self.L.ns_spectral_set_error_control(self.p, C.byref(p), rtol, atol))

    def error_control_stats(self):
        """(accepted, rejected, last error norm, next step limit)"""
        a, r = C.c_ulonglong(), C.c_ulonglong()
        e, dtn = C.c_double(), C.c_double()
        self.check(self.L.ns_spectral_error_control_stats(self.p, C.byref(a), C.byref(r),
                                                          C.byref(e), C.byref(dtn)))
        return a.value, r.value, e.value, dtn.value

    def convergence_metrics(self):
        """{E, Z, max|omega_hat|, spectral error, t, dt} of analyze_convergence"""
        out = np.zeros(6)
        self.check(self.L.ns_spectral_convergence_metrics(self.p, _dp(out)))
        return out

    @property
    def ctx(self):
        """NsdContext* of this instance (first field of the device struct)."""
        return C.cast(self.d.device, C.POINTER(C.c_void_p))[0]


class Ensemble:
    """Independent solves of different sizes, Reynolds numbers and seeds
    (ns_ensemble.h).  Keyword arguments set NSEnsembleConfig fields; `forcing`
    takes (kmin, kmax, injection_rate)."""

    def __init__(self, sizes, Re, seeds, forcing=None, **cfg):
        self.L = load()
        c = NSEnsembleConfig()
        self.L.ns_ensemble_config_default(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        if forcing is not None:
            c.forcing_enabled = 1
            c.forcing.kmin, c.forcing.kmax = float(forcing[0]), float(forcing[1])
            c.forcing.energy_target = 1.0
            c.forcing.type = NS_FORCING_WHITE_NOISE
            c.injection_rate = float(forcing[2])
        self.cfg = c
        self.sizes = [int(n) for n in sizes]
        m = len(self.sizes)
        n_arr = (C.c_int * m)(*self.sizes)
        re_arr = (C.c_double * m)(*[float(r) for r in Re])
        sd_arr = (C.c_ulonglong * m)(*[int(s) for s in seeds])
        self.p = self.L.ns_ensemble_create(m, n_arr, re_arr, sd_arr, C.byref(c))
        if not self.p:
            raise RuntimeError("ns_ensemble_create: " + self.L.ns_ensemble_create_error().decode())
        self.members = m

    def close(self):
        if getattr(self, "p", None):
            self.L.ns_ensemble_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            raise RuntimeError(f"ensemble call failed ({rc}): "
                               + self.L.ns_ensemble_last_error(self.p).decode())
        return rc

    def advance(self, nsteps):
        return self.check(self.L.ns_ensemble_advance(self.p, nsteps))

    def synchronize(self):
        return self.check(self.L.ns_ensemble_synchronize(self.p))

    def diagnostics(self):
        m = self.members
        t, e, z = np.zeros(m), np.zeros(m), np.zeros(m)
        st = np.zeros(m, dtype=np.int32)
        self.check(self.L.ns_ensemble_diagnostics(self.p, _dp(t), _dp(e), _dp(z),
                                                  st.ctypes.data_as(c_int_p)))
        return t, e, z, st

    def state(self, member):
        n = self.sizes[member]
        out = np.zeros((n, n // 2 + 1), dtype=np.complex128)
        self.check(self.L.ns_ensemble_member_state(self.p, member, _dp(out)))
        return out
