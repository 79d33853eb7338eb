"""Python host binding of libgecco_hip.so (the MI355X block-contraction engine).

Mirrors GeCCo's kernel interface for the blocked contraction path:
  contr_blk1blk2_blocked_mm(...)  -> contract/contr_blk1blk2_blocked_mm.f:2-20
  contr_blk1blk2_wmaps_c(...)     -> contract/contr_blk1blk2_wmaps_c.f:2-19
and exposes the grouped-launch plan (Plan) used by the block enumerator.

A kernel call is described by a dict whose keys are the reference's argument
names (xfac, ncblk_op1, ..., lstrop1, ..., map_info_12c, ..., map_ex2cnta).
Errors raise GeccoHipError with the library's message (the reference stops
through quit(), baserout/quit.F:36-50).  There is no CPU fallback: if the
library is missing, importing this module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'libgecco_hip.so')

I64 = ctypes.c_int64
PI64 = ctypes.POINTER(ctypes.c_int64)
PD = ctypes.POINTER(ctypes.c_double)


class GeccoHipError(RuntimeError):
    pass


if not os.path.exists(LIBPATH):
    raise ImportError('libgecco_hip.so is not built (run __graft_entry__.build())')
_lib = ctypes.CDLL(LIBPATH)


class BlockArgs(ctypes.Structure):
    _fields_ = [('xfac', ctypes.c_double),
                ('tra_op1', I64), ('tra_op2', I64), ('tra_op1op2', I64)] + \
        [(n, I64) for n in ('ncblk_op1', 'nablk_op1', 'ncblk_ex1', 'nablk_ex1',
                            'ncblk_op2', 'nablk_op2', 'ncblk_ex2', 'nablk_ex2',
                            'ncblk_cnt', 'nablk_cnt', 'ncblk_op1op2', 'nablk_op1op2')] + \
        [(n, PI64) for n in ('hpvxop1c', 'hpvxop1a', 'hpvxop2c', 'hpvxop2a',
                             'hpvxop1op2c', 'hpvxop1op2a', 'lstrop1', 'lstrop2',
                             'lstrop1op2', 'lstr_ex1', 'lstr_ex2', 'lstr_cnt',
                             'map_info_12c', 'map_info_12a', 'map_info_1c',
                             'map_info_1a', 'map_info_2c', 'map_info_2a',
                             'map_ex1ex2c', 'map_ex1ex2a', 'map_ex1cntc',
                             'map_ex1cnta', 'map_ex2cntc', 'map_ex2cnta')] + \
        [('buf_op1', ctypes.c_int32), ('buf_op2', ctypes.c_int32),
         ('buf_op1op2', ctypes.c_int32),
         ('off_op1', I64), ('off_op2', I64), ('off_op1op2', I64)]


DIM_KEYS = ('ncblk_op1', 'nablk_op1', 'ncblk_ex1', 'nablk_ex1', 'ncblk_op2',
            'nablk_op2', 'ncblk_ex2', 'nablk_ex2', 'ncblk_cnt', 'nablk_cnt',
            'ncblk_op1op2', 'nablk_op1op2')
ARR_KEYS = ('hpvxop1c', 'hpvxop1a', 'hpvxop2c', 'hpvxop2a', 'hpvxop1op2c',
            'hpvxop1op2a', 'lstrop1', 'lstrop2', 'lstrop1op2', 'lstr_ex1',
            'lstr_ex2', 'lstr_cnt', 'map_info_12c', 'map_info_12a', 'map_info_1c',
            'map_info_1a', 'map_info_2c', 'map_info_2a', 'map_ex1ex2c',
            'map_ex1ex2a', 'map_ex1cntc', 'map_ex1cnta', 'map_ex2cntc', 'map_ex2cnta')

FORTRAN_SYMBOLS = ('contr_blk1blk2_blocked_mm_', 'contr_blk1blk2_wmaps_c_')
C_SYMBOLS = ('gecco_hip_last_error', 'gecco_hip_plan_create', 'gecco_hip_plan_destroy',
             'gecco_hip_plan_add_block', 'gecco_hip_plan_finalize',
             'gecco_hip_plan_execute', 'gecco_hip_plan_stats', 'gecco_hip_malloc',
             'gecco_hip_free', 'gecco_hip_memcpy_htod', 'gecco_hip_memcpy_dtoh',
             'gecco_hip_synchronize')

_lib.gecco_hip_last_error.restype = ctypes.c_int
_lib.gecco_hip_last_error.argtypes = [ctypes.c_char_p, I64]
for _n in C_SYMBOLS[1:]:
    getattr(_lib, _n).restype = ctypes.c_int
_lib.gecco_hip_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
_lib.gecco_hip_plan_destroy.argtypes = [ctypes.c_void_p]
_lib.gecco_hip_plan_add_block.argtypes = [ctypes.c_void_p, ctypes.POINTER(BlockArgs)]
_lib.gecco_hip_plan_finalize.argtypes = [ctypes.c_void_p]
_lib.gecco_hip_plan_execute.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), I64,
                                        ctypes.c_void_p]
_lib.gecco_hip_plan_stats.argtypes = [ctypes.c_void_p, PD, PI64, PI64, PD, PD]


def lib():
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    code = _lib.gecco_hip_last_error(buf, 512)
    return code, buf.value.decode(errors='replace')


def _check(rc):
    if rc:
        code, msg = last_error()
        raise GeccoHipError(f'gecco_hip error {code}: {msg}')


def _arrays(blk):
    """int64 numpy copies of the integer array arguments (kept alive)."""
    out = {}
    for k in ARR_KEYS:
        v = list(blk[k])
        out[k] = np.asarray(v if v else [0], dtype=np.int64)
    return out


def fill_args(blk, arrs, buf=(0, 1, 2), off=(0, 0, 0)):
    a = BlockArgs()
    a.xfac = float(blk['xfac'])
    a.tra_op1 = int(blk.get('tra_op1', 0))
    a.tra_op2 = int(blk.get('tra_op2', 0))
    a.tra_op1op2 = int(blk.get('tra_op1op2', 0))
    for k in DIM_KEYS:
        setattr(a, k, int(blk[k]))
    for k in ARR_KEYS:
        setattr(a, k, arrs[k].ctypes.data_as(PI64))
    a.buf_op1, a.buf_op2, a.buf_op1op2 = buf
    a.off_op1, a.off_op2, a.off_op1op2 = off
    return a


def _blocked_mm_call(sym, blk, x1, x2, xr):
    arrs = _arrays(blk)
    f = getattr(_lib, sym)
    f.restype = None
    ref = lambda v: ctypes.byref(I64(int(v)))
    args = [ctypes.byref(ctypes.c_double(float(blk['xfac']))), xr, x1, x2,
            ref(blk.get('tra_op1', 0)), ref(blk.get('tra_op2', 0)), ref(blk.get('tra_op1op2', 0))]
    if sym == 'contr_blk1blk2_blocked_mm_':
        scr = ctypes.c_double(0.0)
        args += [ctypes.byref(scr), ref(1), ref(200), ref(-1)]
    args += [ref(blk[k]) for k in DIM_KEYS]
    args += [arrs[k].ctypes.data_as(PI64) for k in ARR_KEYS]
    f(*args)
    code, msg = last_error()
    if code:
        raise GeccoHipError(f'{sym}: error {code}: {msg}')


def _ptr(x):
    """Device or host address of a torch tensor / numpy array / int."""
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    if hasattr(x, 'data_ptr'):
        return ctypes.c_void_p(x.data_ptr())
    return ctypes.c_void_p(x.ctypes.data)


def contr_blk1blk2_blocked_mm(blk, xop1, xop2, xop1op2):
    """Route-3 kernel call with the reference's arguments.  xop* are the block
    starts (numpy host arrays are staged; torch CUDA tensors are used in place);
    xop1op2 is updated in place."""
    _blocked_mm_call('contr_blk1blk2_blocked_mm_', blk, _ptr(xop1), _ptr(xop2), _ptr(xop1op2))


def contr_blk1blk2_wmaps_c(blk, xop1, xop2, xop1op2):
    """Route-2 kernel call (same arithmetic result as route 3)."""
    _blocked_mm_call('contr_blk1blk2_wmaps_c_', blk, _ptr(xop1), _ptr(xop2), _ptr(xop1op2))


class Plan:
    """Grouped launch of many kernel calls (gecco_hip_plan_*)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(_lib.gecco_hip_plan_create(ctypes.byref(self._h)))
        self.nblocks = 0

    def add(self, blk, buf=(0, 1, 2), off=None):
        """Record one kernel call; buf = slots of (op1, op2, result), off =
        element offsets of the block starts (default: blk['off1'],...)."""
        if off is None:
            off = (blk['off1'], blk['off2'], blk['off12'])
        arrs = _arrays(blk)
        a = fill_args(blk, arrs, buf, off)
        _check(_lib.gecco_hip_plan_add_block(self._h, ctypes.byref(a)))
        self.nblocks += 1

    def finalize(self):
        _check(_lib.gecco_hip_plan_finalize(self._h))

    def execute(self, bufs, stream=None):
        arr = (ctypes.c_void_p * len(bufs))(*[_ptr(b).value for b in bufs])
        s = ctypes.c_void_p(0 if stream is None else int(stream))
        _check(_lib.gecco_hip_plan_execute(self._h, arr, len(bufs), s))

    def stats(self):
        f = ctypes.c_double()
        nb = I64()
        nt = I64()
        pb = ctypes.c_double()
        ib = ctypes.c_double()
        _check(_lib.gecco_hip_plan_stats(self._h, ctypes.byref(f), ctypes.byref(nb),
                                         ctypes.byref(nt), ctypes.byref(pb), ctypes.byref(ib)))
        return dict(flops=f.value, nblocks=nb.value, ntasks=nt.value,
                    payload_bytes=pb.value, index_bytes=ib.value)

    def __del__(self):
        try:
            if self._h:
                _lib.gecco_hip_plan_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass
