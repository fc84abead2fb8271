"""ctypes binding of libnqs_hip.so (the C ABI in include/nqs_hip.h).

The library is loaded from this directory; there is no fallback path: when the library or a
GPU is missing, the compute entry points raise.  Device memory is torch (ROCm) tensors, the
stream is torch's current stream.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnqs_hip.so")

NQS_OK, NQS_ERR_ARG, NQS_ERR_HIP, NQS_ERR_WORKSPACE, NQS_ERR_NUMERIC, NQS_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5
NQS_FP64, NQS_MIXED = 0, 1
NQS_GEMM_BF16X3, NQS_GEMM_FP32, NQS_GEMM_BF16 = 0, 1, 2
OP_LOGPSI, OP_ELOC, OP_MCMC, OP_GRAD, OP_SPIN, OP_VJP = 0, 1, 2, 3, 4, 5

c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class ModelDesc(ctypes.Structure):
    _fields_ = [("n_elec", c_i32), ("n_g", c_i32), ("embed_dim", c_i32), ("depth", c_i32),
                ("n_heads", c_i32), ("ff_mult", c_i32), ("n_freq", c_i32), ("n_dets", c_i32),
                ("learnable_env", c_i32), ("use_jastrow", c_i32), ("jastrow_kind", c_i32),
                ("n_knots", c_i32), ("precision", c_i32), ("gemm", c_i32),
                ("cell", c_dbl * 4), ("cell_inv", c_dbl * 4), ("cusp", c_dbl)]


class HamDesc(ctypes.Structure):
    _fields_ = [("m_star", c_dbl), ("lambda_r", c_dbl), ("h_z", c_dbl), ("n_field", c_i32),
                ("n_pin", c_i32), ("field_dev", c_vp), ("pin_dev", c_vp), ("cell", c_dbl * 4),
                ("field_scale", c_dbl)]


class EwaldDesc(ctypes.Structure):
    _fields_ = [("alpha", c_dbl), ("eta", c_dbl), ("v_g0", c_dbl), ("zeta_m", c_dbl),
                ("n_img", c_i32), ("n_g", c_i32), ("img_dev", c_vp), ("g_dev", c_vp), ("rg_dev", c_vp)]


class MinsrDesc(ctypes.Structure):
    _fields_ = [("n_walkers", c_i64), ("n_params", c_i64), ("ldo", c_i64), ("o_dtype", c_i32), ("gram", c_i32),
                ("refine", c_i32), ("reserved", c_i32), ("lam", c_dbl)]


# (name, restype, argtypes) of every exported symbol; the CPU tests check each one resolves.
SIGNATURES = [
    ("nqs_version", c_i32, []),
    ("nqs_desc_size", c_i64, [c_i32]),
    ("nqs_last_error", ctypes.c_char_p, []),
    ("nqs_param_count", c_i64, [ctypes.POINTER(ModelDesc)]),
    ("nqs_param_leaf", c_i32, [ctypes.POINTER(ModelDesc), c_i32, ctypes.c_char_p, c_i32,
                               ctypes.POINTER(c_i64), ctypes.POINTER(c_i32), ctypes.POINTER(c_i64),
                               ctypes.POINTER(c_i32)]),
    ("nqs_packed_bytes", c_i64, [ctypes.POINTER(ModelDesc)]),
    ("nqs_pack_params", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("nqs_workspace_bytes", c_i64, [ctypes.POINTER(ModelDesc), c_i32, c_i64]),
    ("nqs_psiformer_fwd", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_psiformer_fwd_noalloc", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    ("nqs_eloc", c_i32, [ctypes.POINTER(ModelDesc), ctypes.POINTER(HamDesc), c_vp, c_vp, c_vp, c_i64, c_vp,
                         c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_psiformer_grad", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp,
                                   c_i64, c_vp]),
    ("nqs_psiformer_vjp", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                                  c_vp]),
    ("nqs_minsr_workspace_bytes", c_i64, [ctypes.POINTER(MinsrDesc)]),
    ("nqs_minsr_direction", c_i32, [ctypes.POINTER(MinsrDesc), c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_minsr_gram", c_i32, [ctypes.POINTER(MinsrDesc), c_vp, c_vp, c_i64, c_vp]),
    ("nqs_minsr_factor", c_i32, [ctypes.POINTER(MinsrDesc), c_vp, c_vp, c_i64, c_vp]),
    ("nqs_minsr_solve", c_i32, [ctypes.POINTER(MinsrDesc), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_clip_energies", c_i32, [c_vp, c_i64, c_dbl, c_vp, c_vp, c_vp]),
    ("nqs_minsr_batched_workspace_bytes", c_i64, [ctypes.POINTER(MinsrDesc), c_i32]),
    ("nqs_minsr_direction_batched", c_i32, [ctypes.POINTER(MinsrDesc), c_i32, ctypes.POINTER(c_vp),
                                            ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), c_vp, c_vp, c_i64, c_vp]),
    ("nqs_vmc_gradient_workspace_bytes", c_i64, [ctypes.POINTER(MinsrDesc)]),
    ("nqs_vmc_gradient", c_i32, [ctypes.POINTER(MinsrDesc), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_adam_update", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_i32,
                                c_vp]),
    ("nqs_spin_estimators", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp]),
    ("nqs_spin_fourier", c_i32, [c_vp, c_vp, c_i64, c_i32, c_vp, c_i32, c_vp, c_vp]),
    ("nqs_spin_texture", c_i32, [c_vp, c_vp, c_i64, c_i32, ctypes.POINTER(c_dbl), c_i32, c_vp, c_vp, c_vp]),
    ("nqs_ratio_stats", c_i32, [c_vp, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp]),
    ("nqs_minsr_profile", c_i32, [c_i32, ctypes.POINTER(ctypes.c_float)]),
    ("nqs_autotune_enable", c_i32, [c_i32]),
    ("nqs_autotune_reset", c_i32, []),
    ("nqs_autotune_report", c_i64, [ctypes.c_char_p, c_i64]),
    ("nqs_ewald", c_i32, [ctypes.POINTER(EwaldDesc), c_vp, c_i64, c_i32, c_vp, c_vp]),
    ("nqs_mcmc_sweeps", c_i32, [ctypes.POINTER(ModelDesc), c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_dbl,
                                c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
]

_lib = None


def load():
    """Load the library once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def autotune_report():
    """The autotuner's choices so far, one line per tuned shape class."""
    lib = load()
    n = lib.nqs_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    lib.nqs_autotune_report(buf, n + 1)
    return buf.value.decode()


def check(rc, what):
    if rc != NQS_OK:
        msg = load().nqs_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with code {rc}: {msg}")


def ptr(t):
    """Device pointer of a contiguous torch tensor (or None -> NULL)."""
    if t is None:
        return None
    assert t.is_contiguous(), "tensor must be contiguous"
    return ctypes.c_void_p(t.data_ptr())


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device():
    """The GPU every product tensor lives on; raises when no GPU is present."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the nqs HIP path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


_WS = {}


def workspace(nbytes, tag="default"):
    """A cached device byte buffer of at least nbytes (grown on demand, reused across calls)."""
    import torch
    buf = _WS.get(tag)
    if buf is None or buf.numel() < nbytes:
        _WS.pop(tag, None)
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())
        _WS[tag] = buf
    return buf


def release_workspaces():
    """Drop every cached workspace buffer (their memory returns to torch's caching allocator)."""
    _WS.clear()
