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
import struct

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
             'gecco_hip_synchronize', 'gecco_hip_plan_set_timing', 'gecco_hip_plan_timing',
             'gecco_hip_plan_block_info', 'gecco_hip_plan_set_materialize',
             'gecco_hip_plan_workspace', 'gecco_hip_plan_block_index',
             'gecco_hip_plan_set_constant', 'gecco_hip_plan_set_chain',
             'gecco_hip_plan_set_group', 'gecco_hip_plan_set_skip', 'gecco_hip_plan_set_tuning',
             'gecco_hip_plan_keep_host', 'gecco_hip_plan_retune', 'gecco_hip_plan_set_precision',
             'gecco_hip_task_order', 'gecco_hip_plan_compact', 'gecco_hip_plan_live_tasks')

_lib.gecco_hip_last_error.restype = ctypes.c_int
_lib.gecco_hip_last_error.argtypes = [ctypes.c_char_p, I64]
for _n in C_SYMBOLS[1:]:
    getattr(_lib, _n).restype = ctypes.c_int
_lib.gecco_hip_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
_lib.gecco_hip_task_order.argtypes = [I64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
_lib.gecco_hip_plan_destroy.argtypes = [ctypes.c_void_p]
_lib.gecco_hip_plan_add_block.argtypes = [ctypes.c_void_p, ctypes.POINTER(BlockArgs)]
_lib.gecco_hip_plan_finalize.argtypes = [ctypes.c_void_p]
_lib.gecco_hip_plan_set_group.argtypes = [ctypes.c_void_p, ctypes.c_int32]
_lib.gecco_hip_plan_set_tuning.argtypes = [ctypes.c_void_p, I64, I64, I64, I64]
_lib.gecco_hip_plan_keep_host.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_set_precision.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_retune.argtypes = [ctypes.c_void_p, I64, I64, I64, I64]
_lib.gecco_hip_plan_set_skip.argtypes = [ctypes.c_void_p, ctypes.c_void_p, I64]
_lib.gecco_hip_plan_compact.argtypes = [ctypes.c_void_p, I64, ctypes.c_void_p]
_lib.gecco_hip_plan_live_tasks.argtypes = [ctypes.c_void_p, PI64, I64]
_lib.gecco_hip_plan_execute.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), I64,
                                        ctypes.c_void_p]
_lib.gecco_hip_plan_stats.argtypes = [ctypes.c_void_p, PD, PI64, PI64, PD, PD]
_lib.gecco_hip_plan_set_timing.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_timing.argtypes = [ctypes.c_void_p, PD, PI64, PD]
_lib.gecco_hip_plan_block_info.argtypes = [ctypes.c_void_p, I64, PI64, I64]
_lib.gecco_hip_plan_set_materialize.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_set_constant.argtypes = [ctypes.c_void_p, ctypes.c_int32]
_lib.gecco_hip_plan_set_chain.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_set_screen.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_set_deterministic.argtypes = [ctypes.c_void_p, I64]
_lib.gecco_hip_plan_workspace.argtypes = [ctypes.c_void_p, PI64, PI64, PD, PI64]
_lib.gecco_hip_plan_block_index.argtypes = [ctypes.c_void_p, I64, I64, ctypes.POINTER(ctypes.c_int32),
                                            I64, ctypes.POINTER(ctypes.c_int32), I64]


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
    if sym in ('contr_blk1blk2_blocked_mm_', 'contr_blk1blk2_blocked_mm_noalloc_'):
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


def contr_blk1blk2_blocked_mm_noalloc(blk, xop1, xop2, xop1op2):
    """The allocation-free variant (contr_blk1blk2_blocked_mm_noalloc_):
    device tensors only; no device memory is allocated."""
    _blocked_mm_call('contr_blk1blk2_blocked_mm_noalloc_', blk, _ptr(xop1), _ptr(xop2), _ptr(xop1op2))


def _iocc(c, a):
    """GeCCo's iocc(ngastp, 2) of one vertex (column 1 C, column 2 A, rows
    H, P, V, X) from occupation dicts, as int64 in Fortran order."""
    return np.concatenate([_occ4(c), _occ4(a)]).astype(np.int64)


def contr_op1op2(xfac, op1, op2, op1op2, cnt, xop1, xop2, xop1op2, irreps, update=True,
                 tra=(False, False, False), use_tr=True):
    """One contraction with contr_op1op2's semantics through
    gecco_hip_contr_op1op2_: op1/op2/op1op2 = (C occupation, A occupation)
    of the operators as the contraction uses them, cnt = (contracted part of
    op1's C side, of op1's A side), xop* the whole lists (numpy host arrays
    or torch CUDA tensors; xop1op2 updated in place), irreps = (nsym,
    IRREPs of the hole orbitals, IRREPs of the particle orbitals), tra the
    dagger flags (the list stored as the adjoint)."""
    nsym, ih, ip = irreps
    ih = np.ascontiguousarray(ih, dtype=np.int64)
    ip = np.ascontiguousarray(ip, dtype=np.int64)
    occ = [_iocc(*o) for o in (op1, op2, op1op2, cnt)]
    f = _lib.gecco_hip_contr_op1op2_
    f.restype = None
    ref = lambda v: ctypes.byref(I64(int(v)))
    err = I64()
    f(ctypes.byref(ctypes.c_double(float(xfac))), ref(bool(update)),
      *[o.ctypes.data_as(PI64) for o in occ], ref(tra[0]), ref(tra[1]), ref(tra[2]), ref(bool(use_tr)),
      ref(nsym), ref(len(ih)), ih.ctypes.data_as(PI64), ref(len(ip)), ip.ctypes.data_as(PI64),
      _ptr(xop1), _ptr(xop2), _ptr(xop1op2), ctypes.byref(err))
    if err.value:
        raise GeccoHipError(f'gecco_hip_contr_op1op2_: error {err.value}: {last_error()[1]}')


def contr_op1op2_n(xfac, op1, op2, op1op2, ex1, ex2, cnt, merge, xop1, xop2, xop1op2, irreps,
                   update=True, use_tr=True, op1op2tmp=None, reo=None):
    """A contraction of operators with several vertices through
    gecco_hip_contr_op1op2_n_: op1/op2/op1op2 per-vertex (C, A) occupation
    lists (lists of dicts), ex1/ex2/cnt per-vertex (C, A) dict pairs, merge
    the four short-form merge maps; with reo = (from_vtx, to_vtx, (shift C,
    shift A), sign_reo) (vertices from 0 here) the kernel calls write
    op1op2tmp, which is reordered into op1op2."""
    nsym, ih, ip = irreps
    ih = np.ascontiguousarray(ih, dtype=np.int64)
    ip = np.ascontiguousarray(ip, dtype=np.int64)
    vo = lambda o: np.concatenate([_iocc(c, a) for c, a in zip(*o)])
    tmp = op1op2 if op1op2tmp is None else op1op2tmp
    arrs = [vo(op1), vo(op2), vo(op1op2), vo(tmp), _iocc_n(ex1), _iocc_n(ex2), _iocc_n(cnt)]
    mm = [np.asarray(list(m) or [0], dtype=np.int64) for m in merge]
    frm, to, shift, sgn = reo if reo is not None else (0, 0, ({}, {}), 1)
    sh = _iocc(*shift)
    f = _lib.gecco_hip_contr_op1op2_n_
    f.restype = None
    ref = lambda v: ctypes.byref(I64(int(v)))
    err = I64()
    P = lambda a: a.ctypes.data_as(PI64)
    f(ctypes.byref(ctypes.c_double(float(xfac))), ref(bool(update)), ref(len(op1[0])), P(arrs[0]),
      ref(len(op2[0])), P(arrs[1]), ref(len(op1op2[0])), P(arrs[2]), P(arrs[3]), ref(len(cnt)),
      P(arrs[4]), P(arrs[5]), P(arrs[6]), *[P(m) for m in mm], ref(reo is not None), ref(frm + 1),
      ref(to + 1), P(sh), ref(sgn), ref(bool(use_tr)), ref(nsym), ref(len(ih)), P(ih), ref(len(ip)),
      P(ip), _ptr(xop1), _ptr(xop2), _ptr(xop1op2), ctypes.byref(err))
    if err.value:
        raise GeccoHipError(f'gecco_hip_contr_op1op2_n_: error {err.value}: {last_error()[1]}')


def reo_op(fac, ori, reo, frm, to, shift, xori, xreo, irreps, update=True, sign_reo=1):
    """Reorder of a whole list with reo_op_wmaps_c's semantics through
    gecco_hip_reo_op_: ori/reo = (C occupations, A occupations) per vertex
    (lists of dicts), the shift string (C dict, A dict) moved from vertex frm
    to vertex to (0-based here), xori/xreo whole lists (xreo updated in
    place), irreps as for contr_op1op2."""
    nsym, ih, ip = irreps
    ih = np.ascontiguousarray(ih, dtype=np.int64)
    ip = np.ascontiguousarray(ip, dtype=np.int64)
    occ = [np.concatenate([_iocc(c, a) for c, a in zip(*o)]) for o in (ori, reo)]
    sh = _iocc(*shift)
    f = _lib.gecco_hip_reo_op_
    f.restype = None
    ref = lambda v: ctypes.byref(I64(int(v)))
    err = I64()
    f(ctypes.byref(ctypes.c_double(float(fac))), ref(bool(update)), ref(sign_reo),
      ref(len(ori[0])), occ[0].ctypes.data_as(PI64), ref(len(reo[0])), occ[1].ctypes.data_as(PI64),
      ref(frm + 1), ref(to + 1), sh.ctypes.data_as(PI64), ref(nsym), ref(len(ih)),
      ih.ctypes.data_as(PI64), ref(len(ip)), ip.ctypes.data_as(PI64), _ptr(xori), _ptr(xreo),
      ctypes.byref(err))
    if err.value:
        raise GeccoHipError(f'gecco_hip_reo_op_: error {err.value}: {last_error()[1]}')


def call_cache(clear=False):
    """Number of one-call plans cached by the Fortran-signature entries
    (after freeing them all when clear)."""
    n = I64()
    _check(_lib.gecco_hip_call_cache(ctypes.byref(n), int(bool(clear))))
    return n.value


def noalloc_scratch_bytes():
    """Size of the fixed device array of the allocation-free entry."""
    f = _lib.gecco_hip_noalloc_scratch_bytes
    f.restype = I64
    return int(f())


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
        elem = getattr(self, '_elem', 8)
        for b in bufs:
            if hasattr(b, 'element_size') and b.element_size() != elem:
                raise GeccoHipError(f'plan of {8 * elem}-bit lists given a {b.dtype} buffer')
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

    def set_timing(self, on=True):
        _check(_lib.gecco_hip_plan_set_timing(self._h, int(bool(on))))

    def timing(self):
        """{'ms': [gemm, gather, sym, materialise, direct], 'launches': [...],
        'flops': [gemm, gather, direct]} accumulated since the previous call."""
        ms = (ctypes.c_double * 5)()
        n = (I64 * 5)()
        fl = (ctypes.c_double * 3)()
        _check(_lib.gecco_hip_plan_timing(self._h, ms, n, fl))
        return dict(ms=list(ms), launches=list(n), flops=list(fl))

    def block_info(self, i):
        info = (I64 * 30)()
        _check(_lib.gecco_hip_plan_block_info(self._h, i, info, 30))
        return list(info)

    def block_operands(self, i):
        """(buffer slot, element offset) of A and B in kernel orientation
        and the factor of kernel call i."""
        info = self.block_info(i)
        xfac = struct.unpack('<d', struct.pack('<q', info[29]))[0]
        return (info[25], info[26]), (info[27], info[28]), xfac

    def block_index(self, i, which):
        """Resolved signed index data (first, second) of operand which (0 A,
        1 B, 2 C) of kernel call i, kernel orientation; before finalize."""
        info = self.block_info(i)
        M, N, K = info[0], info[1], info[2]
        F, S = (K, M) if which == 0 else (K, N) if which == 1 else (M, N)
        table = info[5 + which]
        first = np.zeros(F * S if table else F, dtype=np.int32)
        second = np.zeros(1 if table else S, dtype=np.int32)
        _check(_lib.gecco_hip_plan_block_index(
            self._h, i, which, first.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), first.size,
            second.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), second.size))
        if table:
            return first.reshape(S, F).T, None
        return first, second

    def set_constant(self, slot):
        """Slot `slot` keeps its data between executes (materialised gathers
        of it are reused); before finalize."""
        _check(_lib.gecco_hip_plan_set_constant(self._h, int(slot)))

    def set_screen(self, on=True):
        """Norm screening: skip kernel calls whose operand blocks have a
        squared norm below 1e-28, checked on the device at every execute
        (the reference's zero-block test); before finalize."""
        _check(_lib.gecco_hip_plan_set_screen(self._h, 1 if on else 0))

    def set_deterministic(self, on=True):
        """Fixed-order summation: bit-reproducible results
        (gecco_hip_plan_set_deterministic); before finalize."""
        _check(_lib.gecco_hip_plan_set_deterministic(self._h, int(bool(on))))

    def set_chain(self, on):
        """Run strided-GEMM kernel calls that add into the same result block
        through the same result map as one GEMM over their concatenated
        contraction ranges (default on); before finalize."""
        _check(_lib.gecco_hip_plan_set_chain(self._h, 1 if on else 0))

    def set_tuning(self, variant=None, target_tasks=None, min_kchunk=None, max_kchunk=None):
        """Strided-GEMM tile variant, task-split target and contraction
        chunk bounds (None keeps the current value); before finalize."""
        v = lambda x: -1 if x is None else int(x)
        _check(_lib.gecco_hip_plan_set_tuning(self._h, v(variant), v(target_tasks), v(min_kchunk),
                                              v(max_kchunk)))

    def set_precision(self, precision):
        """'f64' (default), 'f32' or 'bf16x3' (float32 lists whose GEMM
        tiles run as split-bfloat16 three-term products: tile variant 8);
        before finalize."""
        elem = {'f64': 8, 'f32': 4, 'bf16x3': 4}[precision]
        _check(_lib.gecco_hip_plan_set_precision(self._h, elem))
        if precision == 'bf16x3':
            self.set_tuning(variant=8)
        self._elem = elem

    def keep_host(self, on=True):
        """Before finalize: keep the host index data (for retune); after
        finalize, keep_host(False) releases it."""
        _check(_lib.gecco_hip_plan_keep_host(self._h, 1 if on else 0))

    def retune(self, variant=None, target_tasks=None, min_kchunk=None, max_kchunk=None):
        """Finalize again with other launch settings (plan kept its host
        data)."""
        v = lambda x: -1 if x is None else int(x)
        _check(_lib.gecco_hip_plan_retune(self._h, v(variant), v(target_tasks), v(min_kchunk),
                                          v(max_kchunk)))

    def set_group(self, group):
        """Record the following kernel calls, list initialisations and
        symmetrisations as instance group `group` (default 0); before
        finalize."""
        _check(_lib.gecco_hip_plan_set_group(self._h, int(group)))

    def set_skip(self, flags):
        """Device int32 tensor of one flag per instance group (None: off);
        every execute then skips the work of group g while flags[g] != 0."""
        if flags is None:
            _check(_lib.gecco_hip_plan_set_skip(self._h, None, 0))
            self._skip = None
            return
        if str(flags.dtype) != 'torch.int32' or not flags.is_contiguous():
            raise GeccoHipError('set_skip: flags must be a contiguous int32 tensor')
        _check(_lib.gecco_hip_plan_set_skip(self._h, _ptr(flags), flags.numel()))
        self._skip = flags            # keeps the memory alive

    def compact(self, on=True, stream=None):
        """on: launch only the tasks of groups whose skip flag is 0 now (the
        live-task lists are made on `stream`, which is then synchronized);
        off: the full task lists again (call it before clearing a flag).
        Each compaction rewrites the live lists in place: capture a graph
        of the compacted plan again after the next call."""
        s = ctypes.c_void_p(0 if stream is None else int(stream))
        _check(_lib.gecco_hip_plan_compact(self._h, 1 if on else 0, s))

    def live_tasks(self):
        """Task counts the launches use now: strided-GEMM, gather kernel,
        initialisation, symmetrisation, constant-slot gathers, other gathers."""
        out = (I64 * 6)()
        _check(_lib.gecco_hip_plan_live_tasks(self._h, out, 6))
        return [int(x) for x in out]

    def set_materialize(self, min_dim):
        """Materialise gathered operands of kernel calls with min(M, N) >=
        min_dim into scratch (0 disables); before finalize."""
        _check(_lib.gecco_hip_plan_set_materialize(self._h, int(min_dim)))

    def workspace(self):
        sb = I64()
        nj = I64()
        gb = ctypes.c_double()
        st = I64()
        _check(_lib.gecco_hip_plan_workspace(self._h, ctypes.byref(sb), ctypes.byref(nj),
                                             ctypes.byref(gb), ctypes.byref(st)))
        return dict(scratch_bytes=sb.value, gather_jobs=nj.value, gather_bytes=gb.value,
                    staged_tasks=st.value)

    def __del__(self):
        try:
            if self._h:
                _lib.gecco_hip_plan_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass


# ----------------------------------------------------------- plan builder
MODEL_SYMBOLS = ('gecco_hip_orbs_create', 'gecco_hip_orbs_destroy', 'gecco_hip_orbs_lenstr',
                 'gecco_hip_orbs_strmap', 'gecco_hip_orbs_flipmap', 'gecco_hip_list_create',
                 'gecco_hip_list_destroy', 'gecco_hip_list_length', 'gecco_hip_list_ndist',
                 'gecco_hip_list_dist', 'gecco_hip_plan_add_contraction',
                 'gecco_hip_plan_add_sym_ab', 'gecco_hip_list_spinorbs',
                 'gecco_hip_list_diagonal', 'gecco_hip_plan_add_trace', 'gecco_hip_list_create_n',
                 'gecco_hip_plan_add_reorder')
SOLVER_SYMBOLS = ('gecco_hip_diis_create', 'gecco_hip_diis_destroy', 'gecco_hip_diis_reset',
                  'gecco_hip_diis_step', 'gecco_hip_diis_weights', 'gecco_hip_diis_matrix',
                  'gecco_hip_diis_step_async', 'gecco_hip_diis_status', 'gecco_hip_diis_set_flag',
                  'gecco_hip_diis_set_scale', 'gecco_hip_diis_group_create',
                  'gecco_hip_diis_group_step_async', 'gecco_hip_diis_group_destroy')
MLP_SYMBOLS = ('gecco_hip_mlp_infer', 'gecco_hip_mlp_scale', 'gecco_hip_mlp_grad', 'gecco_hip_mlp_adam')
MLP_P = 17281
MLP_SCALE_WORK = 2 + 2 * 512
MLP_GRAD_GROUPS = 256
MLP_GRAD_WORK = MLP_GRAD_GROUPS * MLP_P + 2 * MLP_GRAD_GROUPS
VP = ctypes.c_void_p
_lib.gecco_hip_orbs_create.argtypes = [I64, I64, PI64, I64, PI64, ctypes.POINTER(VP)]
_lib.gecco_hip_orbs_create.restype = ctypes.c_int
_lib.gecco_hip_orbs_destroy.argtypes = [VP]
_lib.gecco_hip_orbs_lenstr.argtypes = [VP, I64, I64, I64, I64]
_lib.gecco_hip_orbs_lenstr.restype = I64
for _n in ('gecco_hip_orbs_strmap',):
    getattr(_lib, _n).argtypes = [VP, I64, I64, I64, I64, I64, I64, I64, PI64, I64, PI64]
    getattr(_lib, _n).restype = ctypes.c_int
_lib.gecco_hip_orbs_flipmap.argtypes = [VP, I64, I64, I64, I64, PI64, I64, PI64]
_lib.gecco_hip_orbs_flipmap.restype = ctypes.c_int
_lib.gecco_hip_list_create.argtypes = [VP, PI64, PI64, I64, ctypes.POINTER(VP)]
_lib.gecco_hip_list_create.restype = ctypes.c_int
_lib.gecco_hip_list_destroy.argtypes = [VP]
_lib.gecco_hip_list_length.argtypes = [VP]
_lib.gecco_hip_list_length.restype = I64
_lib.gecco_hip_list_ndist.argtypes = [VP]
_lib.gecco_hip_list_ndist.restype = I64
_lib.gecco_hip_list_dist.argtypes = [VP, I64, PI64, I64]
_lib.gecco_hip_list_dist.restype = ctypes.c_int
_lib.gecco_hip_plan_add_contraction.argtypes = [
    VP, VP, VP, VP, PI64, PI64, ctypes.c_double, I64, ctypes.c_int32, I64, ctypes.c_int32, I64,
    ctypes.c_int32, I64, PI64, PD]
_lib.gecco_hip_plan_add_contraction.restype = ctypes.c_int
_lib.gecco_hip_plan_add_contraction_n.argtypes = [
    VP, VP, VP, VP, I64, PI64, PI64, PI64, PI64, PI64, PI64, PI64, ctypes.c_double, I64,
    ctypes.c_int32, I64, ctypes.c_int32, I64, ctypes.c_int32, I64, ctypes.POINTER(I64),
    ctypes.POINTER(ctypes.c_double)]
_lib.gecco_hip_plan_add_contraction_n.restype = ctypes.c_int
_lib.gecco_hip_plan_add_sym_ab.argtypes = [VP, VP, ctypes.c_int32, I64, ctypes.c_double,
                                           ctypes.c_double]
_lib.gecco_hip_plan_add_sym_ab.restype = ctypes.c_int
_lib.gecco_hip_list_spinorbs.argtypes = [VP, ctypes.POINTER(ctypes.c_int16), I64]
_lib.gecco_hip_list_spinorbs.restype = ctypes.c_int
_lib.gecco_hip_list_diagonal.argtypes = [VP, PD, PD, PD, I64]
_lib.gecco_hip_plan_add_trace.argtypes = [VP, VP, VP, PI64, ctypes.c_double, ctypes.c_int32, I64,
                                          ctypes.c_int32, I64, PI64]
_lib.gecco_hip_plan_add_trace.restype = ctypes.c_int
_lib.gecco_hip_list_create_n.argtypes = [VP, I64, PI64, I64, ctypes.POINTER(VP)]
_lib.gecco_hip_list_create_n.restype = ctypes.c_int
_lib.gecco_hip_plan_add_reorder.argtypes = [VP, VP, VP, I64, I64, PI64, PI64, ctypes.c_double, I64,
                                            ctypes.c_int32, I64, ctypes.c_int32, I64, PI64]
_lib.gecco_hip_plan_add_reorder.restype = ctypes.c_int
_lib.gecco_hip_list_diagonal.restype = ctypes.c_int
_lib.gecco_hip_plan_add_init.argtypes = [VP, ctypes.c_int32, I64, I64, ctypes.c_int32, I64,
                                         ctypes.c_double]
_lib.gecco_hip_plan_add_init.restype = ctypes.c_int
_lib.gecco_hip_plan_add_sum.argtypes = _lib.gecco_hip_plan_add_init.argtypes
_lib.gecco_hip_plan_add_sum.restype = ctypes.c_int
_lib.gecco_hip_plan_add_spin_project.argtypes = [VP, VP, ctypes.c_int32, I64]
_lib.gecco_hip_plan_add_spin_project.restype = ctypes.c_int
_lib.gecco_hip_diis_create.argtypes = [I64, I64, ctypes.POINTER(VP)]
_lib.gecco_hip_diis_destroy.argtypes = [VP]
_lib.gecco_hip_diis_reset.argtypes = [VP]
_lib.gecco_hip_diis_step.argtypes = [VP, VP, VP, VP, VP, ctypes.c_double, PD, PD, PI64, VP]
_lib.gecco_hip_diis_step_async.argtypes = [VP, VP, VP, VP, VP, ctypes.c_double, VP]
_lib.gecco_hip_diis_status.argtypes = [VP, I64, PD, PD, PI64]
_lib.gecco_hip_diis_set_flag.argtypes = [VP, VP]
_lib.gecco_hip_diis_set_scale.argtypes = [VP, VP]
_lib.gecco_hip_diis_group_create.argtypes = [ctypes.POINTER(VP), I64, ctypes.POINTER(VP),
                                             ctypes.POINTER(VP), ctypes.POINTER(VP),
                                             ctypes.POINTER(VP), ctypes.POINTER(VP)]
_lib.gecco_hip_diis_group_step_async.argtypes = [VP, ctypes.c_double, VP]
_lib.gecco_hip_diis_group_destroy.argtypes = [VP]
_lib.gecco_hip_mlp_infer.argtypes = [VP, VP, I64, VP, VP]
_lib.gecco_hip_mlp_scale.argtypes = [VP, VP, VP, VP, I64, ctypes.c_double, VP, VP, VP]
_lib.gecco_hip_mlp_grad.argtypes = [VP, VP, VP, VP, I64, VP, VP, VP, VP]
_lib.gecco_hip_mlp_adam.argtypes = [VP, VP, VP, VP, I64, ctypes.c_double, VP]
for _n in MLP_SYMBOLS:
    getattr(_lib, _n).restype = ctypes.c_int
_lib.gecco_hip_diis_weights.argtypes = [VP, PD, I64, PI64]
_lib.gecco_hip_diis_matrix.argtypes = [VP, PD, I64, PI64]
for _n in SOLVER_SYMBOLS:
    getattr(_lib, _n).restype = ctypes.c_int

HOLE, PART = 1, 2


def _occ4(d):
    a = np.zeros(4, dtype=np.int64)
    for h, n in d.items():
        a[h - 1] = n
    return a


class Orbs:
    """Orbital spaces (hole / particle IRREPs) of the model."""

    def __init__(self, h_irreps, p_irreps, nsym):
        h = np.asarray(list(h_irreps) or [1], dtype=np.int64)
        p = np.asarray(list(p_irreps) or [1], dtype=np.int64)
        self._h = VP()
        self.nsym = nsym
        _check(_lib.gecco_hip_orbs_create(nsym, len(h_irreps), h.ctypes.data_as(PI64),
                                          len(p_irreps), p.ctypes.data_as(PI64),
                                          ctypes.byref(self._h)))

    def lenstr(self, hpvx, nel, idxms, igam):
        return _lib.gecco_hip_orbs_lenstr(self._h, hpvx, nel, idxms, igam)

    def _getmap(self, fn, *args):
        n = 1 << 16
        while True:
            out = np.zeros(n, dtype=np.int64)
            ln = I64()
            rc = fn(self._h, *args, out.ctypes.data_as(PI64), n, ctypes.byref(ln))
            if rc == 0:
                return out[:ln.value].tolist()
            if ln.value > n:
                n = ln.value
                continue
            _check(rc)

    def strmap(self, hpvx, n1, idxms1, g1, n2, idxms2, g2):
        return self._getmap(_lib.gecco_hip_orbs_strmap, hpvx, n1, idxms1, g1, n2, idxms2, g2)

    def flipmap(self, hpvx, nel, idxms, g):
        return self._getmap(_lib.gecco_hip_orbs_flipmap, hpvx, nel, idxms, g)

    def __del__(self):
        try:
            _lib.gecco_hip_orbs_destroy(self._h)
        except Exception:
            pass


class MEList:
    """ME-list layout of an operator (operators/set_op_dim2.f).  occ_c/occ_a:
    dict hpvx->count for one vertex, or lists of such dicts (one per
    vertex)."""

    def __init__(self, orbs, occ_c, occ_a, absym=1):
        self.orbs = orbs
        if isinstance(occ_c, dict):
            occ_c, occ_a = [occ_c], [occ_a]
        self.vocc_c = [dict(x) for x in occ_c]
        self.vocc_a = [dict(x) for x in occ_a]
        self.njoined = len(self.vocc_c)
        self.occ_c, self.occ_a = {}, {}
        for x in self.vocc_c:
            for h, n in x.items():
                self.occ_c[h] = self.occ_c.get(h, 0) + n
        for x in self.vocc_a:
            for h, n in x.items():
                self.occ_a[h] = self.occ_a.get(h, 0) + n
        self.absym = absym
        occ = np.concatenate([np.concatenate([_occ4(c), _occ4(a)])
                              for c, a in zip(self.vocc_c, self.vocc_a)]).astype(np.int64)
        self._h = VP()
        _check(_lib.gecco_hip_list_create_n(orbs._h, self.njoined, occ.ctypes.data_as(PI64), absym,
                                            ctypes.byref(self._h)))
        self.length = _lib.gecco_hip_list_length(self._h)

    def dists(self):
        out = []
        buf = np.zeros(64, dtype=np.int64)
        for i in range(_lib.gecco_hip_list_ndist(self._h)):
            _check(_lib.gecco_hip_list_dist(self._h, i, buf.ctypes.data_as(PI64), 64))
            off, ln, msa, ga, ncc, nca = buf[:6].tolist()
            comps = buf[6:6 + 5 * (ncc + nca)].reshape(-1, 5).tolist()
            out.append(dict(off=off, len=ln, msa_tot=msa, igama=ga, comps_c=comps[:ncc],
                            comps_a=comps[ncc:]))
        return out

    @property
    def nel(self):
        return sum(self.occ_c.values()) + sum(self.occ_a.values())

    def spinorbs(self):
        """(length, nel) int16 array of the spin-orbitals of every element."""
        out = np.zeros((self.length, max(1, self.nel)), dtype=np.int16)
        _check(_lib.gecco_hip_list_spinorbs(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                            out.size))
        return out[:, :self.nel]

    def diagonal(self, eps_h, eps_p):
        """Diagonal preconditioner (dia4op_ev): per element, the orbital
        energies of the C strings minus those of the A strings."""
        eh = np.ascontiguousarray(eps_h, dtype=np.float64)
        ep = np.ascontiguousarray(eps_p, dtype=np.float64)
        out = np.zeros(max(1, self.length))
        _check(_lib.gecco_hip_list_diagonal(self._h, eh.ctypes.data_as(PD), ep.ctypes.data_as(PD),
                                            out.ctypes.data_as(PD), out.size))
        return out[:self.length]

    def adjoint(self):
        """Adjoint view X^+ of this one-vertex list (gecco_hip_list_adjoint):
        C and A swapped, the same storage; a contraction that takes the view
        runs its kernel calls with the transpose flag (the vertex's dagger)."""
        v = MEList.__new__(MEList)
        v.orbs = self.orbs
        v.vocc_c, v.vocc_a = [dict(x) for x in self.vocc_a], [dict(x) for x in self.vocc_c]
        v.njoined = self.njoined
        v.occ_c, v.occ_a = dict(self.occ_a), dict(self.occ_c)
        v.absym = self.absym
        v._h = VP()
        _check(_lib.gecco_hip_list_adjoint(self._h, ctypes.byref(v._h)))
        v.length = _lib.gecco_hip_list_length(v._h)
        v.base = self        # keeps the stored list alive
        return v

    def comp_spaces(self):
        """Space (1=H, 2=P) of each electron column of spinorbs()."""
        from itertools import chain
        cols = []
        for side in (self.occ_c, self.occ_a):
            for h in (4, 2, 3, 1):
                cols += [h] * side.get(h, 0)
        return cols

    def __del__(self):
        try:
            _lib.gecco_hip_list_destroy(self._h)
        except Exception:
            pass


def _plan_add_contraction(self, op1, op2, res, cnt_c, cnt_a, fact=1.0, use_tr=True,
                          buf=(0, 1, 2), off=(0, 0, 0)):
    """Record all kernel calls of Res += fact*Op1.Op2 (contr_op1op2)."""
    cc, ca = _occ4(cnt_c), _occ4(cnt_a)
    nb = I64()
    fl = ctypes.c_double()
    _check(_lib.gecco_hip_plan_add_contraction(
        self._h, op1._h, op2._h, res._h, cc.ctypes.data_as(PI64), ca.ctypes.data_as(PI64),
        float(fact), int(bool(use_tr)), buf[0], off[0], buf[1], off[1], buf[2], off[2],
        ctypes.byref(nb), ctypes.byref(fl)))
    self.nblocks += nb.value
    return nb.value, fl.value


def _iocc_n(vocc):
    """Per-vertex (occ_c, occ_a) dicts -> iocc(4, 2, nvtx), C column first."""
    if not vocc:
        return np.zeros(8, dtype=np.int64)
    return np.concatenate([np.concatenate([_occ4(c), _occ4(a)]) for c, a in vocc]).astype(np.int64)


def _plan_add_contraction_n(self, op1, op2, res, ex1, ex2, cnt, merge1, merge2, merge12,
                            merge21, fact=1.0, use_tr=True, buf=(0, 1, 2), off=(0, 0, 0)):
    """Record all kernel calls of Res += fact*Op1.Op2 for operators with
    several vertices (gecco_hip_plan_add_contraction_n): ex1 / ex2 / cnt are
    per-vertex (occ_c, occ_a) dicts, merge* the short-form merge maps."""
    e1, e2, cn = _iocc_n(ex1), _iocc_n(ex2), _iocc_n(cnt)
    mm = [np.asarray(list(m) or [0], dtype=np.int64) for m in (merge1, merge2, merge12, merge21)]
    nb = I64()
    fl = ctypes.c_double()
    _check(_lib.gecco_hip_plan_add_contraction_n(
        self._h, op1._h, op2._h, res._h, len(cnt), e1.ctypes.data_as(PI64),
        e2.ctypes.data_as(PI64), cn.ctypes.data_as(PI64), *[m.ctypes.data_as(PI64) for m in mm],
        float(fact), int(bool(use_tr)), buf[0], off[0], buf[1], off[1], buf[2], off[2],
        ctypes.byref(nb), ctypes.byref(fl)))
    self.nblocks += nb.value
    return nb.value, fl.value


def _plan_add_sym_ab(self, lst, buf, off=0, prefac=0.5, relfac=None):
    """Spin-flip symmetrisation of a result list (sym_ab_blk)."""
    if relfac is None:
        relfac = float(lst.absym)
    _check(_lib.gecco_hip_plan_add_sym_ab(self._h, lst._h, buf, off, prefac, relfac))


def _plan_add_init(self, buf_dst, length, buf_src=-1, factor=1.0, off_dst=0, off_src=0):
    """Initialise slot buf_dst[off_dst:off_dst+length] to factor * slot
    buf_src (or to zero when buf_src < 0) at the start of every execute."""
    _check(_lib.gecco_hip_plan_add_init(self._h, int(buf_dst), int(off_dst), int(length),
                                        int(buf_src), int(off_src), float(factor)))


def _plan_add_trace(self, op, trop, cnt, fact=1.0, buf=(0, 1), off=(0, 0)):
    """Record OPtr += fact * sum_J OP(J.IC, J.IA) (trace_op); buf/off =
    slots and offsets of (OP, OPtr).  Returns the number of element terms."""
    c = _occ4(cnt)
    n = I64()
    _check(_lib.gecco_hip_plan_add_trace(self._h, op._h, trop._h, c.ctypes.data_as(PI64),
                                         float(fact), buf[0], off[0], buf[1], off[1],
                                         ctypes.byref(n)))
    return n.value


def _plan_add_reorder(self, ori, reo, frm, to, shift_c, shift_a, fact=1.0, sign_reo=1,
                      buf=(0, 1), off=(0, 0)):
    """Record REO += fact*sign_reo * ORI with the shift string K (shift_c,
    shift_a) moved from vertex frm of ORI to vertex to of REO (reo_op_wmaps_c).
    Returns the number of element terms."""
    kc, ka = _occ4(shift_c), _occ4(shift_a)
    n = I64()
    _check(_lib.gecco_hip_plan_add_reorder(self._h, ori._h, reo._h, int(frm), int(to),
                                           kc.ctypes.data_as(PI64), ka.ctypes.data_as(PI64),
                                           float(fact), int(sign_reo), buf[0], off[0], buf[1],
                                           off[1], ctypes.byref(n)))
    return n.value


Plan.add_contraction = _plan_add_contraction
Plan.add_contraction_n = _plan_add_contraction_n
Plan.add_reorder = _plan_add_reorder
Plan.add_trace = _plan_add_trace
Plan.add_sym_ab = _plan_add_sym_ab
Plan.add_init = _plan_add_init


def _plan_add_sum(self, buf_dst, length, buf_src, factor=1.0, off_dst=0, off_src=0):
    """dst += factor * src over `length` elements (an ADD item), in the
    initialisation launch; atomic, so several sums into one list may share
    the plan, but an initialisation of the same list must be in an earlier
    plan."""
    _check(_lib.gecco_hip_plan_add_sum(self._h, int(buf_dst), int(off_dst), int(length), int(buf_src),
                                       int(off_src), float(factor)))


Plan.add_sum = _plan_add_sum


def _plan_add_spin_project(self, lst, buf, off=0):
    """Singlet spin projection of list `lst` in slot `buf` after the plan's
    contractions (gecco_hip_plan_add_spin_project)."""
    _check(_lib.gecco_hip_plan_add_spin_project(self._h, lst._h, int(buf), int(off)))


Plan.add_spin_project = _plan_add_spin_project


class DIIS:
    """Diagonal-preconditioned Pulay DIIS on a device-resident amplitude
    vector of n doubles (gecco_hip_diis_*)."""

    def __init__(self, n, maxsub=8):
        self._h = VP()
        _check(_lib.gecco_hip_diis_create(int(n), int(maxsub), ctypes.byref(self._h)))
        self.n = int(n)
        self.maxsub = int(maxsub)

    def reset(self):
        _check(_lib.gecco_hip_diis_reset(self._h))

    def step(self, t, rest, diag, energy=None, tol=0.0, stream=None):
        """One iteration for g(t) = diag*t + rest; t is updated in place
        unless ||g|| < tol.  Returns (||g||, energy or None, converged)."""
        for x in (t, rest, diag):
            if x.numel() != self.n:
                raise GeccoHipError('diis.step: vector length mismatch')
        gn = ctypes.c_double()
        en = ctypes.c_double()
        cv = I64()
        s = VP(0 if stream is None else int(stream))
        _check(_lib.gecco_hip_diis_step(self._h, _ptr(t), _ptr(rest), _ptr(diag),
                                        None if energy is None else _ptr(energy), float(tol),
                                        ctypes.byref(gn), ctypes.byref(en), ctypes.byref(cv), s))
        return gn.value, (en.value if energy is not None else None), bool(cv.value)

    def step_async(self, t, rest, diag, energy=None, tol=0.0, stream=None):
        """The same step with the subspace bookkeeping and the small solve on
        the device: nothing is synchronised, and every call queues the same
        launches (an iteration can be captured as a graph).  Its record is
        read with status(i) for the i-th asynchronous step since reset."""
        for x in (t, rest, diag):
            if x.numel() != self.n:
                raise GeccoHipError('diis.step_async: vector length mismatch')
        s = VP(0 if stream is None else int(stream))
        _check(_lib.gecco_hip_diis_step_async(self._h, _ptr(t), _ptr(rest), _ptr(diag),
                                              None if energy is None else _ptr(energy), float(tol), s))

    def set_scale(self, scale):
        """Device float64 vector of length n (or None): later steps use
        r = scale * g / D (learned per-amplitude factor)."""
        if scale is not None and (str(scale.dtype) != 'torch.float64' or scale.numel() != self.n):
            raise GeccoHipError('set_scale: float64 tensor of the amplitude length expected')
        _check(_lib.gecco_hip_diis_set_scale(self._h, None if scale is None else _ptr(scale)))
        self._scale = scale

    def set_flag(self, flag):
        """One-element device int32 tensor (or None) that an asynchronous
        step sets to 1 when it meets tol; while it is set, later asynchronous
        steps leave t and the subspace unchanged and repeat the converged
        record.  The caller clears it on restart."""
        if flag is not None and (str(flag.dtype) != 'torch.int32' or flag.numel() != 1):
            raise GeccoHipError('set_flag: one-element int32 tensor expected')
        _check(_lib.gecco_hip_diis_set_flag(self._h, None if flag is None else _ptr(flag)))
        self._flag = flag

    def status(self, step):
        """(||g||, energy, code) of asynchronous step `step` once the stream
        has passed it; code 1 converged, 0 extrapolated, -1 singular."""
        gn = ctypes.c_double()
        en = ctypes.c_double()
        cv = I64()
        _check(_lib.gecco_hip_diis_status(self._h, int(step), ctypes.byref(gn), ctypes.byref(en),
                                          ctypes.byref(cv)))
        return gn.value, en.value, int(cv.value)

    def weights(self):
        w = (ctypes.c_double * 16)()
        m = I64()
        _check(_lib.gecco_hip_diis_weights(self._h, w, 16, ctypes.byref(m)))
        return np.array(w[:m.value])

    def matrix(self):
        b = (ctypes.c_double * 256)()
        m = I64()
        _check(_lib.gecco_hip_diis_matrix(self._h, b, 256, ctypes.byref(m)))
        return np.array(b[:m.value * m.value]).reshape(m.value, m.value)

    def __del__(self):
        try:
            if self._h:
                _lib.gecco_hip_diis_destroy(self._h)
                self._h = VP()
        except Exception:
            pass


# ------------------------------------------------------------------ learned preconditioner
def _mlp_check(t, dtype, n, what):
    if str(t.dtype) != dtype or t.numel() != n or not t.is_contiguous() or not t.is_cuda:
        raise GeccoHipError(f'{what}: contiguous {dtype} device tensor of {n} elements expected')


def mlp_infer(params, x, stream=None):
    """y for feature rows x (n x 4, float32 device tensor)."""
    import torch
    n = x.shape[0]
    _mlp_check(params, 'torch.float32', MLP_P, 'mlp_infer params')
    _mlp_check(x, 'torch.float32', 4 * n, 'mlp_infer x')
    y = torch.empty(n, dtype=torch.float32, device=x.device)
    _check(_lib.gecco_hip_mlp_infer(_ptr(params), _ptr(x), n, _ptr(y), VP(0 if stream is None else int(stream))))
    return y


def mlp_scale(params, diag, rest, t, it, scale, work, stream=None):
    """scale = 1 + y(features of every amplitude); work: float64 device
    tensor of MLP_SCALE_WORK elements."""
    n = diag.numel()
    _mlp_check(params, 'torch.float32', MLP_P, 'mlp_scale params')
    for v, nm in ((diag, 'diag'), (rest, 'rest'), (t, 't'), (scale, 'scale')):
        _mlp_check(v, 'torch.float64', n, f'mlp_scale {nm}')
    _mlp_check(work, 'torch.float64', MLP_SCALE_WORK, 'mlp_scale work')
    _check(_lib.gecco_hip_mlp_scale(_ptr(params), _ptr(diag), _ptr(rest), _ptr(t), n, float(it),
                                    _ptr(scale), _ptr(work), VP(0 if stream is None else int(stream))))


def mlp_grad(params, x, ut, dt, grad, loss, work, stream=None):
    """Loss gradient over the pairs (x, ut, dt) into grad (float32,
    MLP_P) and the loss into loss[0] (float64 device tensor or None)."""
    n = ut.numel()
    _mlp_check(params, 'torch.float32', MLP_P, 'mlp_grad params')
    _mlp_check(x, 'torch.float32', 4 * n, 'mlp_grad x')
    _mlp_check(ut, 'torch.float32', n, 'mlp_grad ut')
    _mlp_check(dt, 'torch.float32', n, 'mlp_grad dt')
    _mlp_check(grad, 'torch.float32', MLP_P, 'mlp_grad grad')
    _mlp_check(work, 'torch.float32', MLP_GRAD_WORK, 'mlp_grad work')
    _check(_lib.gecco_hip_mlp_grad(_ptr(params), _ptr(x), _ptr(ut), _ptr(dt), n, _ptr(grad),
                                   None if loss is None else _ptr(loss), _ptr(work),
                                   VP(0 if stream is None else int(stream))))


def mlp_adam(params, m, v, grad, step, lr, stream=None):
    for t, nm in ((params, 'params'), (m, 'm'), (v, 'v'), (grad, 'grad')):
        _mlp_check(t, 'torch.float32', MLP_P, f'mlp_adam {nm}')
    _check(_lib.gecco_hip_mlp_adam(_ptr(params), _ptr(m), _ptr(v), _ptr(grad), int(step), float(lr),
                                   VP(0 if stream is None else int(stream))))


class DIISGroup:
    """Asynchronous steps of many DIIS subspaces in three launches
    (gecco_hip_diis_group_*): member k steps t[k] with rest[k], diag[k] and
    energy[k].  Set each member's flag and scale before creating the group."""

    def __init__(self, diis, t, rest, diag, energy=None):
        n = len(diis)
        for k in range(n):
            for x in (t[k], rest[k], diag[k]):
                if x.numel() != diis[k].n:
                    raise GeccoHipError('DIISGroup: vector length mismatch')
        arr = lambda xs: (VP * n)(*[None if x is None else _ptr(x).value for x in xs])
        self._keep = (list(diis), list(t), list(rest), list(diag), list(energy or []))
        h = VP()
        hd = (VP * n)(*[d._h.value for d in diis])
        _check(_lib.gecco_hip_diis_group_create(hd, n, arr(t), arr(rest),
                                                arr(diag), None if energy is None else arr(energy),
                                                ctypes.byref(h)))
        self._h = h

    def step_async(self, tol=0.0, stream=None):
        _check(_lib.gecco_hip_diis_group_step_async(self._h, float(tol),
                                                    VP(0 if stream is None else int(stream))))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            _lib.gecco_hip_diis_group_destroy(h)
            self._h = None


def task_order(patch, cost):
    """Launch order of the strided-GEMM launch's tasks (gecco_hip_task_order;
    host only): patch and cost per task, tasks in cost order.  Returns the
    task index at every launch position (int64 numpy array)."""
    import numpy as np
    patch = np.ascontiguousarray(patch, dtype=np.int32)
    cost = np.ascontiguousarray(cost, dtype=np.int64)
    if patch.shape != cost.shape or patch.ndim != 1:
        raise ValueError('patch and cost must be 1-d arrays of one length')
    perm = np.empty(len(patch), dtype=np.int64)
    _check(_lib.gecco_hip_task_order(len(patch), patch.ctypes.data, cost.ctypes.data, perm.ctypes.data))
    return perm
