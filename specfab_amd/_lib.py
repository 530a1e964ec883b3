"""ctypes binding of libspecfab_gpu.so (C-ABI: include/specfab_gpu.h).

The library is the only compute path: if it is missing or no HIP device is
present, every entry point raises instead of computing anything on the CPU.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("SFG_LIBRARY", os.path.join(HERE, "libspecfab_gpu.so"))

SFG_LROT, SFG_DDRX, SFG_CDRX, SFG_REG = 1, 2, 4, 8
SFG_EULER, SFG_RK4 = 0, 1
SFG_FP64, SFG_FP32C = 0, 1
PROC_BITS = {"LROT": SFG_LROT, "DDRX": SFG_DDRX, "CDRX": SFG_CDRX, "REG": SFG_REG}

# every exported symbol of include/specfab_gpu.h with (restype, argtypes)
_c = ctypes
_vp, _i, _i64, _d = _c.c_void_p, _c.c_int, _c.c_int64, _c.c_double
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)


class EvolveDesc(_c.Structure):
    _fields_ = [("L", _c.c_int), ("P", _c.c_int64), ("procs", _c.c_int), ("scheme", _c.c_int),
                ("precision", _c.c_int), ("nsteps", _c.c_int), ("dt", _c.c_double), ("iota", _c.c_double),
                ("zeta", _c.c_double), ("ugrad_dev", _c.c_void_p), ("tau_dev", _c.c_void_p),
                ("gam0_dev", _c.c_void_p), ("lam_dev", _c.c_void_p), ("steps_dev", _c.c_void_p),
                ("x_dev", _c.c_void_p), ("status_dev", _c.c_void_p), ("reg_nu", _c.c_double),
                ("reg_expo", _c.c_double)]


class HomogParams(_c.Structure):
    _fields_ = [("Eij_grain", _c.c_double * 2), ("alpha", _c.c_double), ("n_grain", _c.c_int),
                ("Lame", _c.c_double * 5), ("rho", _c.c_double), ("vi_alpha", _c.c_double),
                ("eij_tol", _c.c_double)]


class SurrogateStats(_c.Structure):
    _fields_ = [("n_fallback", _c.c_int64), ("n_audit", _c.c_int64), ("audit_max_err_Eij", _c.c_double),
                ("audit_mean_err_Eij", _c.c_double), ("audit_max_err_Vi", _c.c_double),
                ("audit_mean_err_Vi", _c.c_double), ("n_bound", _c.c_int64)]


_fp = _c.POINTER(_c.c_float)

SIGNATURES = {
    "sfg_ctx_create": (_i, [_i, _c.POINTER(_vp)]),
    "sfg_ctx_destroy": (_i, [_vp]),
    "sfg_last_error": (_c.c_char_p, []),
    "sfg_ctx_set_timing": (_i, [_vp, _i]),
    "sfg_autotune": (_i, [_vp, _i, _i, _i64, _i, _vp, _ip, _dp, _i]),
    "sfg_ctx_read_timing": (_i, [_vp, _dp, _ip]),
    "sfg_nlm_len": (_i, [_i]),
    "sfg_lm": (_i, [_i, _ip]),
    "sfg_M_LROT": (_i, [_vp, _i, _dp, _dp, _d, _d, _dp]),
    "sfg_M_DDRX": (_i, [_vp, _i, _dp, _dp, _dp]),
    "sfg_M_DDRX_src": (_i, [_vp, _i, _dp, _dp]),
    "sfg_M_CDRX": (_i, [_vp, _i, _dp]),
    "sfg_M_REG": (_i, [_vp, _i, _dp, _dp]),
    "sfg_M_LROT_batch": (_i, [_vp, _i, _i64, _vp, _vp, _d, _d, _vp, _vp]),
    "sfg_M_REG_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_M_DDRX_src_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_rnlm_len": (_i, [_i]),
    "sfg_reduce_M": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "sfg_nlm_to_rnlm": (_i, [_vp, _i, _vp, _vp]),
    "sfg_rnlm_to_nlm": (_i, [_vp, _i, _vp, _vp]),
    "sfg_reduce_M_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sfg_nlm_to_rnlm_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_rnlm_to_nlm_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_nlm_to_state": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_state_to_nlm": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_evolve": (_i, [_vp, _c.POINTER(EvolveDesc), _vp]),
    "sfg_a2_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_a4mandel_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_a2": (_i, [_vp, _i, _dp, _dp]),
    "sfg_Eij_tranisotropic_batch": (_i, [_vp, _i, _i64, _vp, _i, _vp, _dp, _d, _i, _vp, _vp, _vp, _vp, _vp]),
    "sfg_eig": (_i, [_vp, _i, _dp, _dp, _dp]),
    "sfg_Eij_tranisotropic": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "sfg_ev_D2": (_i, [_vp, _i, _dp, _dp, _dp]),
    "sfg_a4": (_i, [_vp, _i, _dp, _dp]),
    "sfg_a4_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_reg_calibration": (_i, [_i, _dp, _dp]),
    "sfg_a2_to_nlm": (_i, [_vp, _dp, _dp]),
    "sfg_a4_to_nlm": (_i, [_vp, _dp, _dp]),
    "sfg_a2_to_state_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_a4_to_state_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_Gamma0": (_i, [_vp, _dp, _c.c_double, _c.c_double, _c.c_double, _dp]),
    "sfg_Gamma0_batch": (_i, [_vp, _i64, _vp, _vp, _c.c_double, _c.c_double, _c.c_double, _vp, _vp]),
    "sfg_apply_bounds": (_i, [_vp, _i, _dp, _dp]),
    "sfg_apply_bounds_batch": (_i, [_vp, _i, _i64, _vp, _vp]),
    "sfg_Cij_to_Lame_tranisotropic": (_i, [_dp, _dp]),
    "sfg_Vi_elastic_tranisotropic": (_i, [_vp, _i, _dp, _d, _dp, _d, _i, _dp, _dp, _dp]),
    "sfg_Vi_elastic_tranisotropic_batch": (_i, [_vp, _i, _i64, _vp, _d, _dp, _d, _i, _dp, _dp, _vp, _vp, _vp]),
    "sfg_Vi_electromagnetic_tranisotropic": (_i, [_vp, _i, _dp, _d, _d, _d, _i, _dp, _dp, _dp]),
    "sfg_Vi_electromagnetic_tranisotropic_batch": (_i, [_vp, _i, _i64, _vp, _d, _d, _d, _i, _dp, _dp, _vp, _vp]),
    "sfg_plan_create": (_i, [_vp, _i, _c.POINTER(EvolveDesc), _i, _i, _vp, _c.POINTER(_vp)]),
    "sfg_plan_run": (_i, [_vp, _vp]),
    "sfg_plan_destroy": (_i, [_vp]),
    "sfg_rotate_nlm": (_i, [_vp, _i, _dp, _d, _d, _dp]),
    "sfg_rotate_nlm_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp]),
    "sfg_a4_eigentensors": (_i, [_vp, _i, _dp, _dp, _dp]),
    "sfg_a4_eigentensors_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp]),
    "sfg_nlm_isvalid": (_i, [_vp, _i64, _dp, _dp, _c.POINTER(_c.c_int32)]),
    "sfg_state_isvalid_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp]),
    "sfg_Cij_to_Lame_orthotropic": (_i, [_dp, _dp]),
    "sfg_Eij_orthotropic": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "sfg_Eij_orthotropic_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _i, _vp, _dp, _d, _i, _vp, _vp]),
    "sfg_Vi_elastic_orthotropic": (_i, [_vp, _i, _dp, _dp, _dp, _d, _dp, _d, _i, _dp, _dp, _dp]),
    "sfg_Vi_elastic_orthotropic_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _d, _dp, _d, _i, _dp, _dp, _vp, _vp]),
    "sfg_Eij_orthotropic_discrete": (_i, [_vp, _i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "sfg_Eij_orthotropic_discrete_batch": (_i, [_vp, _i64, _i64, _vp, _vp, _vp, _i, _vp, _dp, _d, _i, _vp, _vp]),
    "sfg_Vi_elastic_orthotropic_discrete": (_i, [_vp, _i64, _dp, _dp, _dp, _d, _dp, _d, _i, _dp, _dp, _dp]),
    "sfg_Vi_elastic_orthotropic_discrete_batch": (_i, [_vp, _i64, _i64, _vp, _vp, _vp, _d, _dp, _d, _i, _dp, _dp,
                                                       _vp, _vp]),
    "sfg_Evw_tranisotropic": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "sfg_Evw_tranisotropic_batch": (_i, [_vp, _i, _i64, _vp, _i, _vp, _dp, _d, _i, _vp, _vp, _vp]),
    "sfg_Evw_orthotropic": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _i, _dp]),
    "sfg_Evw_orthotropic_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _i, _vp, _dp, _d, _i, _vp, _vp]),
    "sfg_moments68": (_i, [_vp, _i, _dp, _dp, _dp]),
    "sfg_moments68_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp]),
    "sfg_M_LROT_reduced_batch": (_i, [_vp, _i, _i64, _vp, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    "sfg_M_DDRX_src_reduced_batch": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sfg_eig3": (_i, [_vp, _dp, _dp, _dp]),
    "sfg_eig3_batch": (_i, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "sfg_Qnorm_tranisotropic": (_i, [_vp, _i, _dp, _d, _dp, _dp]),
    "sfg_Qnorm_tranisotropic_batch": (_i, [_vp, _i, _i64, _vp, _d, _dp, _vp, _vp, _vp]),
    "sfg_ev_D": (_i, [_vp, _i, _dp, _dp, _i, _dp]),
    "sfg_ev_D_batch": (_i, [_vp, _i, _i64, _vp, _vp, _i, _vp, _vp]),
    "sfg_E_CAFFE": (_i, [_vp, _i, _dp, _dp, _d, _d, _i, _dp]),
    "sfg_E_CAFFE_batch": (_i, [_vp, _i, _i64, _vp, _vp, _d, _d, _i, _vp, _vp]),
    "sfg_ri_LROT": (_i, [_vp, _i64, _i, _dp, _d, _dp, _dp, _d, _dp]),
    "sfg_ri_LROT_batch": (_i, [_vp, _i64, _i, _vp, _d, _vp, _vp, _d, _vp, _vp]),
    "sfg_ri_to_nlm": (_i, [_vp, _i, _i64, _dp, _dp, _dp]),
    "sfg_ri_to_state_batch": (_i, [_vp, _i, _i64, _i64, _vp, _vp, _vp, _vp]),
    "sfg_evolve_noalloc": (_i, [_vp, _c.POINTER(EvolveDesc), _vp]),
    "sfg_mlp_nparams": (_i64, []),
    "sfg_mlp_create": (_i, [_vp, _c.c_uint64, _c.POINTER(_vp)]),
    "sfg_mlp_destroy": (_i, [_vp]),
    "sfg_mlp_get_params": (_i, [_vp, _fp]),
    "sfg_mlp_set_params": (_i, [_vp, _fp]),
    "sfg_mlp_set_normalization": (_i, [_vp, _fp, _fp, _fp, _fp]),
    "sfg_mlp_get_normalization": (_i, [_vp, _fp]),
    "sfg_mlp_set_envelope": (_i, [_vp, _dp, _dp]),
    "sfg_mlp_set_error_model": (_i, [_vp, _dp]),
    "sfg_mlp_train_step": (_i, [_vp, _i64, _vp, _vp, _c.c_float, _fp, _vp]),
    "sfg_mlp_gradients": (_i, [_vp, _i64, _vp, _vp, _fp, _fp]),
    "sfg_mlp_forward": (_i, [_vp, _i64, _vp, _vp, _vp]),
    "sfg_surrogate_features": (_i, [_vp, _i, _i64, _vp, _vp, _vp, _vp]),
    "sfg_surrogate_targets": (_i, [_i64, _vp, _vp, _vp, _vp]),
    "sfg_surrogate_homogenize": (_i, [_vp, _vp, _i, _i64, _vp, _vp, _c.POINTER(HomogParams), _d, _c.c_uint64,
                                      _vp, _vp, _vp, _c.POINTER(SurrogateStats), _vp]),
}


class SFGError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libspecfab_gpu.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise SFGError("libspecfab_gpu.so not found at %s: build it with `python -m specfab_amd.build`"
                           % LIBPATH)
        L = _c.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().sfg_last_error()
        raise SFGError("libspecfab_gpu error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc


class Context:
    """Owns an sfg_ctx on one HIP device (coupling tables, stream, workspace)."""

    def __init__(self, device=0):
        self.handle = _vp()
        check(lib().sfg_ctx_create(int(device), _c.byref(self.handle)))
        self.device = device

    def close(self):
        if self.handle:
            lib().sfg_ctx_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)
