"""Python view of librsf.so, the MI355X-native RSF level-set engine.

Thin ctypes binding over the C ABI declared in include/rsf.h.  Device entry
points take torch tensors on a HIP device (PyTorch-ROCm provides device memory
and streams); host entry points take numpy arrays.  Nothing here computes: if
librsf.so is missing, importing the binding raises.

Volumes are numpy/torch arrays of shape (Z, Y, X) = (l, h, w), C order, the
layout of the reference's X-fastest tira::volume (CUDA_RSF/kernel.cu:1010).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "librsf.so")
CLI_PATH = os.path.join(PKG_DIR, "rsf")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_UNSUPPORTED = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4
ERR_IO = 5
ERR_NONFINITE = 6

DTYPE_U8 = 0
DTYPE_F32 = 1
MODE_FAST = 0
MODE_LITERAL = 1


class RSFError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rsf status {status}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("sigma", C.c_float),
        ("lambda1", C.c_float),
        ("lambda2", C.c_float),
        ("nu", C.c_float),
        ("mu", C.c_float),
        ("dt", C.c_float),
        ("epsilon", C.c_float),
        ("sigma2", C.c_float),
        ("iterations", C.c_int32),
        ("mode", C.c_int32),
    ]


class Block(C.Structure):
    _fields_ = [("offset", C.c_int64), ("w", C.c_int64), ("h", C.c_int64), ("l", C.c_int64)]


def make_params(sigma=3.0, lambda1=1.0, lambda2=None, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                sigma2=0.0, iterations=1, mode=MODE_LITERAL) -> Params:
    if lambda2 is None:
        lambda2 = lambda1
    return Params(sigma, lambda1, lambda2, nu, mu, dt, epsilon, sigma2, int(iterations), int(mode))


def reference_defaults() -> Params:
    p = Params()
    lib().rsf_params_default(C.byref(p))
    return p


_lib = None

# name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_F = C.c_float
_PP = C.POINTER(Params)
_PB = C.POINTER(Block)
_SIGS = {
    "rsf_abi_version": (_I, []),
    "rsf_last_error": (C.c_char_p, []),
    "rsf_status_string": (C.c_char_p, [_I]),
    "rsf_params_default": (None, [_PP]),
    "rsf_taps": (_I, [_F, _P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "rsf_context_create": (_I, [_I, _P, C.POINTER(_P)]),
    "rsf_context_destroy": (_I, [_P]),
    "rsf_context_set_stream": (_I, [_P, _P]),
    "rsf_context_hold_image": (_I, [_P, _P]),
    "rsf_context_release_image": (_I, [_P]),
    "rsf_workspace_size": (_I, [_PP, _I64, _I64, _I64, C.POINTER(_I64)]),
    "rsf_step_device": (_I, [_P, _P, _I, _P, _P, _I64, _I64, _I64, _PP]),
    "rsf_step_device_noalloc": (_I, [_P, _P, _I, _P, _P, _I64, _I64, _I64, _PP]),
    "rsf_run_device": (_I, [_P, _P, _I, _P, _I64, _I64, _I64, _PP, _P]),
    "rsf_run_host": (_I, [_P, _P, _I, _P, _I64, _I64, _I64, _PP, _P]),
    "rsf_last_loop_ms": (_I, [_P, C.POINTER(_F)]),
    "rsf_run_batched_device": (_I, [_P, _PB, _I, _P, _I, _P, _PP]),
    "rsf_run_blocks_sequential_device": (_I, [_P, _PB, _I, _P, _I, _P, _PP]),
    "rsf_autotune": (_I, [_P, _PB, _I, _I, _PP, C.c_char_p, _I]),
    "rsf_profile_sweeps": (_I, [_P, _PB, _I, _P, _I, _P, _PP, C.POINTER(_F), C.POINTER(_F)]),
    "rsf_variant_count": (_I, [_I, C.POINTER(_I)]),
    "rsf_variant_info": (_I, [_I, _I, C.c_char_p, _I, C.POINTER(_I), C.POINTER(_I)]),
    "rsf_step_adjoint_device": (_I, [_P, _PB, _I, _P, _I, _P, _P, _P, _P, _PP]),
    "rsf_step_blocks_device": (_I, [_P, _PB, _I, _P, _I, _P, _P, _PP]),
    "rsf_block_grid": (_I, [_I64] * 8 + [_PB, _P, _I, C.POINTER(_I), C.POINTER(_I64)]),
    "rsf_disassemble_device": (_I, [_P, _P, _I, _I64, _I64, _I64, _PB, _P, _I, _P]),
    "rsf_reassemble_device": (_I, [_P, _P] + [_I64] * 8 + [_P]),
    "rsf_run_volume_blocks_device": (_I, [_P, _P, _I, _P] + [_I64] * 7 + [_PP]),
    "rsf_run_batched_host": (_I, [_P, _PB, _I, _P, _I, _P, _PP]),
    "rsf_run_volume_blocks_host": (_I, [_P, _P, _I, _P] + [_I64] * 7 + [_PP]),
    "rsf_npy_info": (_I, [C.c_char_p, C.POINTER(_I64), C.POINTER(_I), C.POINTER(_I)]),
    "rsf_npy_load_device": (_I, [_P, C.c_char_p, _P, _I, _I64]),
    "rsf_npy_save_device": (_I, [_P, C.c_char_p, _P, _I64, _I64, _I64]),
    "rsf_run_npy": (_I, [_P, C.c_char_p, C.c_char_p, _I, C.c_char_p, _PP, _P]),
    "rsf_tile_vote_device": (_I, [_P, _P, _I64, _I64, _I64, _P, _I, _P, _I, _P, _I, _I64, _I64, _I64, _P, _P]),
    "rsf_synth_tubes": (_I, [_P, _P, _I64, _I64, _I64, _I, _F, _F, _F, _F, C.c_uint64]),
    "rsf_initial_phi": (_I, [_P, _P, _I64, _I64, _I64, _F, _F]),
    "rsf_sdf_from_seed": (_I, [_P, _P, _I64, _I64, _I64, _F]),
    "rsf_sdf_from_seed_host": (_I, [_P, _P, _P, _I64, _I64, _I64, _F]),
    "rsf_sdf_from_seed_device": (_I, [_P, _P, _I, _P, _I64, _I64, _I64, _F]),
    "rsf_heaviside_cal": (_I, [_P, _P, _I, _I, _I]),
    "rsf_deri_heaviside_cal": (_I, [_P, _P, _I, _I, _I]),
    "rsf_inverse_cal": (_I, [_P, _P, _I, _I, _I]),
    "rsf_multiplication_cal": (_I, [_P, _P, _P, _I, _I, _I]),
    "rsf_division_cal": (_I, [_P, _P, _P, _I, _I, _I]),
    "rsf_gradient_cal": (_I, [_P, _P, _P, _P, _I, _I, _I]),
    "rsf_gradientmagnitude_cal": (_I, [_P, _P, _P, _P, _I, _I, _I]),
    "rsf_division_nDIV": (_I, [_P] * 7 + [_I, _I, _I]),
    "rsf_gradient_cal_3_input": (_I, [_P] * 6 + [_I, _I, _I]),
    "rsf_normalized_div_plus": (_I, [_P] * 4 + [_I, _I, _I]),
    "rsf_DIV_cal": (_I, [_P] * 4 + [_I, _I, _I]),
    "rsf_adddevice_convolution_seperable": (_I, [_P, _P, _I, _I, _I, _F, _P, C.c_uint]),
    "rsf_pad_border_cal": (_I, [_P, _P, _I, _I, _I, _I]),
    "rsf_adddevice": (_I, [_P] * 5 + [_I, _I, _I, _F]),
    "rsf_Final_PHI_cal": (_I, [_P] * 7 + [_F, _F, _F, _F, _I, _I, _I]),
    "rsf_step_unfused_device": (_I, [_P, _P, _P, _P, _I64, _I64, _I64, _PP]),
}
EXPORTED = tuple(_SIGS)


def lib():
    """Load librsf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `make -C {PKG_DIR}`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != OK:
        raise RSFError(status, lib().rsf_last_error().decode(errors="replace"))


def _ptr(a) -> int:
    """Address of a numpy array or a torch tensor."""
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    if a is None:
        return 0
    return int(a)


def _sync(*arrays):
    """Orders torch's pending work on these tensors before a library call:
    the context runs on its own non-blocking stream."""
    for a in arrays:
        if hasattr(a, "is_cuda") and a.is_cuda:
            import torch

            torch.cuda.current_stream(a.device).synchronize()
            return


def _device_tensor(t, name, numel=None, dtypes=("torch.float32",), shape=None):
    """The device-pointer entry points hand their tensors' addresses to HIP
    kernels: refuse anything that is not a contiguous device tensor of the
    right dtype holding at least `numel` elements (or exactly `shape`)."""
    if not (hasattr(t, "is_cuda") and hasattr(t, "data_ptr")):
        raise TypeError(f"{name}: expected a torch tensor on the HIP device, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is on {t.device}, not on the HIP device")
    if str(t.dtype) not in dtypes:
        raise TypeError(f"{name}: dtype {t.dtype}, expected {' or '.join(dtypes)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if numel is not None and t.numel() < numel:
        raise ValueError(f"{name}: {t.numel()} elements, the blocks span {numel}")
    return t


_IMAGE_DTYPES = ("torch.uint8", "torch.float32")


def _extent(blocks) -> int:
    """Voxels spanned by packed blocks [(offset, w, h, l), ...]."""
    if not blocks:
        raise ValueError("no blocks")
    return max(int(off) + int(w) * int(h) * int(l) for (off, w, h, l) in blocks)


def _dtype_code(a) -> int:
    s = str(getattr(a, "dtype", ""))
    if s in ("uint8", "torch.uint8"):
        return DTYPE_U8
    if s in ("float32", "torch.float32"):
        return DTYPE_F32
    raise TypeError(f"image dtype must be uint8 or float32, got {s}")


def variants(sweep: int):
    """Kernel variants of sweep 1 or 2 compiled into the library: a list of
    (name, taps, literal) -- taps 0 for sweep 2 (any tap count)."""
    n = C.c_int()
    check(lib().rsf_variant_count(int(sweep), C.byref(n)))
    out = []
    for k in range(n.value):
        buf = C.create_string_buffer(64)
        P, lit = C.c_int(), C.c_int()
        check(lib().rsf_variant_info(int(sweep), k, buf, 64, C.byref(P), C.byref(lit)))
        out.append((buf.value.decode(), P.value, bool(lit.value)))
    return out


def taps(sigma: float):
    buf = np.zeros(63, np.float32)
    P, m = C.c_int(), C.c_int()
    check(lib().rsf_taps(sigma, _ptr(buf), 63, C.byref(P), C.byref(m)))
    return buf[: P.value].copy(), P.value, m.value


# --------------------------------------------------------------- synthetic data
def synth_tubes(shape, tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=1):
    """Tube phantom (Z,Y,X) uint8 and its 0/1 mask."""
    l, h, w = shape
    img = np.empty(shape, np.uint8)
    mask = np.empty(shape, np.uint8)
    check(lib().rsf_synth_tubes(_ptr(img), _ptr(mask), w, h, l, int(tubes), rmin, rmax, bias, noise, int(seed)))
    return img, mask


def initial_phi(image: np.ndarray, blur_sigma=1.0, level=110.0):
    image = np.ascontiguousarray(image, np.uint8)
    l, h, w = image.shape
    phi = np.empty(image.shape, np.float32)
    check(lib().rsf_initial_phi(_ptr(image), _ptr(phi), w, h, l, blur_sigma, level))
    return phi


def sdf_from_seed(seed: np.ndarray, shift=13.0):
    seed = np.ascontiguousarray(seed, np.float32)
    l, h, w = seed.shape
    phi = np.empty(seed.shape, np.float32)
    check(lib().rsf_sdf_from_seed(_ptr(seed), _ptr(phi), w, h, l, shift))
    return phi


# -------------------------------------------------------------------- context
class Context:
    """Owns an rsf_context (one device, one stream)."""

    def __init__(self, device: int = 0, stream=None):
        h = C.c_void_p()
        check(lib().rsf_context_create(int(device), stream, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().rsf_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        check(lib().rsf_context_set_stream(self._h, stream))

    # ---- host-pointer forms
    def run_host(self, image: np.ndarray, phi: np.ndarray, params: Params, per_iteration=False):
        image = np.ascontiguousarray(image)
        phi = np.ascontiguousarray(phi, np.float32).copy()
        l, h, w = phi.shape
        if image.shape != phi.shape:
            raise ValueError("image and phi shapes differ")
        ms = np.zeros(max(params.iterations, 1), np.float32) if per_iteration else None
        check(lib().rsf_run_host(self._h, _ptr(image), _dtype_code(image), _ptr(phi), w, h, l,
                                 C.byref(params), _ptr(ms) if ms is not None else None))
        return (phi, ms) if per_iteration else phi

    # ---- device-pointer forms (torch tensors on the HIP device)
    def step(self, image, phi_in, phi_out, params: Params):
        _device_tensor(phi_in, "phi_in")
        if phi_in.dim() != 3:
            raise ValueError("phi_in must be a (l, h, w) volume")
        l, h, w = phi_in.shape
        _device_tensor(phi_out, "phi_out", shape=phi_in.shape)
        _device_tensor(image, "image", dtypes=_IMAGE_DTYPES, shape=phi_in.shape)
        _sync(image, phi_in, phi_out)
        check(lib().rsf_step_device(self._h, _ptr(image), _dtype_code(image), _ptr(phi_in), _ptr(phi_out),
                                    w, h, l, C.byref(params)))

    def run(self, image, phi, params: Params):
        _device_tensor(phi, "phi")
        if phi.dim() != 3:
            raise ValueError("phi must be a (l, h, w) volume")
        l, h, w = phi.shape
        _device_tensor(image, "image", dtypes=_IMAGE_DTYPES, shape=phi.shape)
        _sync(image, phi)
        check(lib().rsf_run_device(self._h, _ptr(image), _dtype_code(image), _ptr(phi), w, h, l,
                                   C.byref(params), None))
        return self.last_loop_ms()

    def last_loop_ms(self) -> float:
        ms = C.c_float()
        check(lib().rsf_last_loop_ms(self._h, C.byref(ms)))
        return ms.value

    def run_batched(self, blocks, image, phi, params: Params):
        arr = blocks_array(blocks)
        n = _extent(blocks)
        _device_tensor(phi, "phi", numel=n)
        _device_tensor(image, "image", numel=n, dtypes=_IMAGE_DTYPES)
        _sync(image, phi)
        check(lib().rsf_run_batched_device(self._h, arr, len(blocks), _ptr(image), _dtype_code(image),
                                           _ptr(phi), C.byref(params)))
        return self.last_loop_ms()

    def run_sequential(self, blocks, image, phi, params: Params):
        arr = blocks_array(blocks)
        n = _extent(blocks)
        _device_tensor(phi, "phi", numel=n)
        _device_tensor(image, "image", numel=n, dtypes=_IMAGE_DTYPES)
        _sync(image, phi)
        check(lib().rsf_run_blocks_sequential_device(self._h, arr, len(blocks), _ptr(image),
                                                     _dtype_code(image), _ptr(phi), C.byref(params)))
        return self.last_loop_ms()

    def step_blocks(self, blocks, image, phi_in, phi_out, params: Params):
        arr = blocks_array(blocks)
        n = _extent(blocks)
        _device_tensor(phi_in, "phi_in", numel=n)
        _device_tensor(phi_out, "phi_out", numel=n)
        _device_tensor(image, "image", numel=n, dtypes=_IMAGE_DTYPES)
        _sync(image, phi_in, phi_out)
        check(lib().rsf_step_blocks_device(self._h, arr, len(blocks), _ptr(image), _dtype_code(image),
                                           _ptr(phi_in), _ptr(phi_out), C.byref(params)))

    def adjoint(self, blocks, image, phi, grad_out, grad_in, grad_params, params: Params):
        arr = blocks_array(blocks)
        n = _extent(blocks)
        for t, name in ((phi, "phi"), (grad_out, "grad_out"), (grad_in, "grad_in")):
            _device_tensor(t, name, numel=n)
        _device_tensor(grad_params, "grad_params", numel=4)
        _device_tensor(image, "image", numel=n, dtypes=_IMAGE_DTYPES)
        _sync(image, phi, grad_out, grad_in, grad_params)
        check(lib().rsf_step_adjoint_device(self._h, arr, len(blocks), _ptr(image), _dtype_code(image),
                                            _ptr(phi), _ptr(grad_out), _ptr(grad_in), _ptr(grad_params),
                                            C.byref(params)))

    def profile_sweeps(self, blocks, image, phi, params: Params):
        """Mean device ms of one sweep-1 and one sweep-2 launch over params.iterations
        eager iterations (HIP events on the context stream)."""
        arr = blocks_array(blocks)
        a, b = C.c_float(), C.c_float()
        n = _extent(blocks)
        _device_tensor(phi, "phi", numel=n)
        _device_tensor(image, "image", numel=n, dtypes=_IMAGE_DTYPES)
        _sync(image, phi)
        check(lib().rsf_profile_sweeps(self._h, arr, len(blocks), _ptr(image), _dtype_code(image), _ptr(phi),
                                       C.byref(params), C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, blocks, image_dtype: int, params: Params) -> str:
        arr = blocks_array(blocks)
        buf = C.create_string_buffer(512)
        check(lib().rsf_autotune(self._h, arr, len(blocks), int(image_dtype), C.byref(params), buf, 512))
        return buf.value.decode()

    def disassemble(self, volume, grid, packed):
        """Copy the blocks of `grid` (a BlockGrid) out of volume (l, h, w) into packed."""
        _device_tensor(volume, "volume", dtypes=_IMAGE_DTYPES)
        if volume.dim() != 3:
            raise ValueError("volume must be (l, h, w)")
        l, h, w = volume.shape
        _device_tensor(packed, "packed", numel=_extent(grid.blocks), dtypes=(str(volume.dtype),))
        _sync(volume, packed)
        check(lib().rsf_disassemble_device(self._h, _ptr(volume), _dtype_code(volume), w, h, l, grid.c_blocks,
                                           grid.c_origins.ctypes.data, len(grid.blocks), _ptr(packed)))

    def reassemble(self, packed, grid, volume):
        """Blend the packed blocks of `grid` into volume (l, h, w), float32."""
        _device_tensor(volume, "volume", shape=grid.shape)
        l, h, w = volume.shape
        _device_tensor(packed, "packed", numel=_extent(grid.blocks))
        _sync(packed, volume)
        check(lib().rsf_reassemble_device(self._h, _ptr(packed), w, h, l, *grid.block_size, grid.overlap,
                                          grid.min_size, _ptr(volume)))

    def run_volume_blocks(self, image, phi, block_size, overlap, params: Params):
        """Whole block pipeline in place on phi (l, h, w): block_size = (bw, bh, bl)."""
        _device_tensor(phi, "phi")
        if phi.dim() != 3:
            raise ValueError("phi must be a (l, h, w) volume")
        _device_tensor(image, "image", dtypes=_IMAGE_DTYPES, shape=phi.shape)
        l, h, w = phi.shape
        bw, bh, bl = block_size
        _sync(image, phi)
        check(lib().rsf_run_volume_blocks_device(self._h, _ptr(image), _dtype_code(image), _ptr(phi), w, h, l,
                                                 bw, bh, bl, overlap, C.byref(params)))

    def sdf_from_seed(self, seed, phi, shift=13.0):
        """phi0 = sdf(seed - shift) on the device (seed uint8 or float32, phi float32)."""
        _device_tensor(seed, "seed", dtypes=_IMAGE_DTYPES)
        if seed.dim() != 3:
            raise ValueError("seed must be a (l, h, w) volume")
        _device_tensor(phi, "phi", shape=seed.shape)
        l, h, w = seed.shape
        _sync(seed, phi)
        check(lib().rsf_sdf_from_seed_device(self._h, _ptr(seed), _dtype_code(seed), _ptr(phi), w, h, l, shift))

    def hold_image(self, image):
        """Promise that `image` stays unchanged until release_image(): its float copy
        and blur are then computed once instead of on every step."""
        _sync(image)
        check(lib().rsf_context_hold_image(self._h, _ptr(image)))

    def release_image(self):
        check(lib().rsf_context_release_image(self._h))

    def tile_vote(self, pred, tile, starts, shape, prob, mask):
        """rsf_tile_vote_device: pred (T, tz, ty, tx) float32, tile = (tz, ty, tx),
        starts = (z0s, y0s, x0s) per axis, shape = (l, h, w); prob float64 and mask
        uint8 device tensors of that shape receive the vote."""
        z0, y0, x0 = (np.ascontiguousarray(a, np.int64) for a in starts)
        l, h, w = shape
        _device_tensor(pred, "pred", numel=len(z0) * len(y0) * len(x0) * int(np.prod(tile)))
        _device_tensor(prob, "prob", dtypes=("torch.float64",), shape=(l, h, w))
        _device_tensor(mask, "mask", dtypes=("torch.uint8",), shape=(l, h, w))
        _sync(pred, prob, mask)
        check(lib().rsf_tile_vote_device(self._h, _ptr(pred), *tile, z0.ctypes.data, len(z0), y0.ctypes.data,
                                         len(y0), x0.ctypes.data, len(x0), w, h, l, _ptr(prob), _ptr(mask)))

    def npy_load(self, path, dst):
        """Stream a 3-D .npy file into the device tensor dst (uint8 as is, else float32)."""
        _device_tensor(dst, "dst", dtypes=_IMAGE_DTYPES)
        _sync(dst)
        check(lib().rsf_npy_load_device(self._h, os.fsencode(path), _ptr(dst), _dtype_code(dst),
                                        dst.numel() * dst.element_size()))

    def npy_save(self, path, src):
        """Stream the float32 device volume src (l, h, w) into a .npy file."""
        _device_tensor(src, "src")
        if src.dim() != 3:
            raise ValueError("src must be a (l, h, w) volume")
        l, h, w = src.shape
        _sync(src)
        check(lib().rsf_npy_save_device(self._h, os.fsencode(path), _ptr(src), w, h, l))

    def run_npy(self, image_path, seed_path, out_path, params: Params, seed_is_phi0=False, per_iteration=False):
        """The rsf3d run file to file on the device; returns per-iteration ms if asked."""
        ms = np.zeros(max(params.iterations, 1), np.float32) if per_iteration else None
        check(lib().rsf_run_npy(self._h, os.fsencode(image_path), os.fsencode(seed_path), int(seed_is_phi0),
                                os.fsencode(out_path), C.byref(params), _ptr(ms) if ms is not None else None))
        return ms

    def step_unfused(self, image_f32, phi_in, phi_out, params: Params):
        _device_tensor(phi_in, "phi_in")
        if phi_in.dim() != 3:
            raise ValueError("phi_in must be a (l, h, w) volume")
        l, h, w = phi_in.shape
        _device_tensor(phi_out, "phi_out", shape=phi_in.shape)
        _device_tensor(image_f32, "image", shape=phi_in.shape)
        _sync(image_f32, phi_in, phi_out)
        check(lib().rsf_step_unfused_device(self._h, _ptr(image_f32), _ptr(phi_in), _ptr(phi_out), w, h, l,
                                            C.byref(params)))


def blocks_array(blocks):
    """[(offset, w, h, l), ...] -> ctypes array of Block."""
    arr = (Block * len(blocks))()
    for k, (off, w, h, l) in enumerate(blocks):
        arr[k] = Block(int(off), int(w), int(h), int(l))
    return arr


class BlockGrid:
    """Overlapping block grid of a (l, h, w) volume, as rsf_block_grid lays it out.

    blocks: [(offset, w, h, l)] in packed order; origins: (n, 3) int64 (x0, y0, z0)."""

    def __init__(self, shape, block_size, overlap, min_size=1):
        l, h, w = shape
        self.shape = (int(l), int(h), int(w))
        self.block_size = tuple(int(b) for b in block_size)  # (bw, bh, bl)
        self.overlap, self.min_size = int(overlap), int(min_size)
        n, tot = C.c_int(), C.c_int64()
        args = (w, h, l, *self.block_size, self.overlap, self.min_size)
        check(lib().rsf_block_grid(*args, None, None, 0, C.byref(n), C.byref(tot)))
        self.c_blocks = (Block * n.value)()
        self.c_origins = np.zeros((n.value, 3), np.int64)
        check(lib().rsf_block_grid(*args, self.c_blocks, self.c_origins.ctypes.data, n.value, C.byref(n),
                                   C.byref(tot)))
        self.blocks = [(b.offset, b.w, b.h, b.l) for b in self.c_blocks]
        self.origins = self.c_origins
        self.packed_voxels = tot.value


def npy_info(path):
    """(shape, is_uint8) of a .npy file, from its header."""
    shape = (C.c_int64 * 8)()
    nd, u8 = C.c_int(), C.c_int()
    check(lib().rsf_npy_info(os.fsencode(path), shape, C.byref(nd), C.byref(u8)))
    if nd.value > 8:
        raise ValueError(f"{path}: {nd.value} dimensions (at most 8 supported)")
    return tuple(shape[k] for k in range(nd.value)), bool(u8.value)


def pack_blocks(shapes):
    """Consecutive packing of (l, h, w) shapes -> (blocks, total voxels)."""
    blocks, off = [], 0
    for (l, h, w) in shapes:
        blocks.append((off, w, h, l))
        off += w * h * l
    return blocks, off
