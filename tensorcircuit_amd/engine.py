"""Python side of the engine: owners of the C-ABI handles (gate programs,
Pauli sums), device workspaces, and the torch.autograd.Function bridges.

The reference's gradients are reverse-mode AD of the framework through the
tensordot network (backends/pytorch_backend.py:497-544) and its batching is
torch.vmap (backends/pytorch_backend.py:605-617).  Here the state is produced
by one HIP launch sequence and differentiated by the adjoint reverse sweep;
the autograd.Functions below carry explicit vmap rules so that
``torch.func.grad_and_value`` / ``torch.vmap`` (what the backend's grad /
vmap / vvag call) compose with the native calls, the pattern of the
reference's torch bridge (interfaces/torch.py:68-111).
"""
from __future__ import annotations

import collections
import ctypes as C
import json
import os
import threading
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

# ---------------------------------------------------------------------------
# device / dtype helpers
# ---------------------------------------------------------------------------

_CDT = {"complex64": torch.complex64, "complex128": torch.complex128}
_RDT = {"complex64": torch.float32, "complex128": torch.float64}
_CODE = {"complex64": _lib.TCM_C64, "complex128": _lib.TCM_C128}


def device() -> torch.device:
    """The HIP device every state lives on; raises without one (no CPU path)."""
    if not torch.cuda.is_available():
        raise RuntimeError("tensorcircuit_amd computes on a HIP device and none is available")
    return torch.device("cuda", torch.cuda.current_device())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _workspace(nbytes: int, dev: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


# objects whose device tables a HIP graph under capture reads (CapturedStep
# holds them for the graph's lifetime); one stack per thread
_SCOPE = threading.local()


def _hold(*objs: Any) -> None:
    stack = getattr(_SCOPE, "stack", None)
    if stack:
        stack[-1].extend(o for o in objs if o is not None)


def _i64() -> C.c_int64:
    return C.c_int64(0)


# ---------------------------------------------------------------------------
# gate programs
# ---------------------------------------------------------------------------


class GateSpec(collections.namedtuple("GateSpec", "kind qubits param theta scale mat")):
    """One gate for the C ABI: tcm_gate_kind, qubit tuple, parameter slot
    (-1 = fixed), fixed angle, angle scale, constant / generator matrix."""

    def key(self) -> Tuple:
        m = None if self.mat is None else np.ascontiguousarray(self.mat, dtype=np.complex128).tobytes()
        return (self.kind, tuple(self.qubits), self.param >= 0, float(self.theta) if self.param < 0 else 0.0,
                float(self.scale), m)


def barrier() -> GateSpec:
    return GateSpec(_lib.G_BARRIER, (), -1, 0.0, 1.0, None)


class Program:
    """A built gate program (tcm_program_build).  ``fuse`` selects LDS-tiled
    fused passes (1) or one kernel per gate (0, the plain path)."""

    def __init__(self, n: int, specs: Sequence[GateSpec], nparams: int, fuse: int = 1, tile_bits: int = 0,
                 adj_tile_bits: int = 0, ckpt_every: int = 0, reg_slots: int = 0, mfma: int = 0, jit: int = 0,
                 greedy: int = 0):
        self.n = int(n)
        self.nparams = int(nparams)
        self.specs = list(specs)
        arr = (_lib.TcmGate * max(len(self.specs), 1))()
        keep = []
        for i, g in enumerate(self.specs):
            arr[i].kind = g.kind
            arr[i].nq = len(g.qubits)
            for j, q in enumerate(g.qubits):
                arr[i].q[j] = int(q)
            arr[i].param = int(g.param)
            arr[i].theta = float(g.theta)
            arr[i].scale = float(g.scale)
            if g.mat is not None:
                m = np.ascontiguousarray(np.asarray(g.mat, dtype=np.complex128)).view(np.float64).copy()
                keep.append(m)
                arr[i].mat = m.ctypes.data_as(C.POINTER(C.c_double))
        opts = _lib.TcmOpts()
        opts.fuse, opts.tile_bits, opts.adj_tile_bits = fuse, tile_bits, adj_tile_bits
        opts.ckpt_every, opts.reg_slots, opts.mfma, opts.jit = ckpt_every, reg_slots, mfma, jit
        opts.greedy = greedy
        h = C.c_void_p()
        check(lib.tcm_program_build(arr, len(self.specs), self.n, self.nparams, C.byref(opts), C.byref(h)),
              "tcm_program_build")
        self.handle = h
        info = _lib.TcmProgramInfo()
        check(lib.tcm_program_get_info(self.handle, C.byref(info)), "tcm_program_get_info")
        self.nckpt = info.nckpt
        self.npasses = info.npasses
        self.ngates = info.ngates
        self.tile_bits = info.tile_bits
        self.fuse = fuse

    def __del__(self, _free=lib.tcm_program_free):
        h = getattr(self, "handle", None)
        if h:
            _free(h)
            self.handle = None

    def compile(self, dtype: str = "complex64", mode: int = 0) -> float:
        """Compile the pass kernels of one schedule now (programs built with
        ``jit=1``); returns the compile seconds (0.0 when cached)."""
        sec = C.c_double(0.0)
        check(lib.tcm_program_compile(self.handle, _CODE[dtype], int(mode), C.byref(sec)), "tcm_program_compile")
        return sec.value

    def describe(self, dtype: str = "complex64", mode: int = 0) -> Dict:
        n = lib.tcm_program_describe(self.handle, _CODE[dtype], mode, None, 0)
        buf = C.create_string_buffer(int(n) + 1)
        lib.tcm_program_describe(self.handle, _CODE[dtype], mode, buf, int(n) + 1)
        return json.loads(buf.value.decode())

    # --- forward -------------------------------------------------------------
    def forward(self, params: Optional[torch.Tensor], batch: int, dtype: str = "complex64",
                inputs: Optional[torch.Tensor] = None, want_ckpts: bool = False,
                param_stride: Optional[int] = None,
                out: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
        _hold(self)
        dev = device()
        N = 1 << self.n
        code = _CODE[dtype]
        if out is None:
            if inputs is not None:
                state = inputs.to(device=dev, dtype=_CDT[dtype]).reshape(batch, N).clone()
            else:
                state = torch.empty(batch, N, dtype=_CDT[dtype], device=dev)
        else:
            state = out
            if inputs is not None:
                state.copy_(inputs.reshape(batch, N))
        p = None
        if self.nparams > 0:
            p = params.to(device=dev, dtype=_RDT[dtype]).contiguous()
        pstride = self.nparams if param_stride is None else param_stride
        ck = torch.empty((self.nckpt if want_ckpts else 0, batch, N), dtype=_CDT[dtype], device=dev)
        nb = _i64()
        check(lib.tcm_forward_workspace_bytes(self.handle, code, batch, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_forward(self.handle, code, _ptr(p), pstride, state.data_ptr(), batch,
                              1 if inputs is None else 0, _ptr(ck) if want_ckpts and self.nckpt else None,
                              ws.data_ptr(), ws.numel(), _stream()), "tcm_forward")
        return state, ck

    # --- vector-Jacobian product ---------------------------------------------
    def vjp(self, params: Optional[torch.Tensor], ckpts: Optional[torch.Tensor], state: torch.Tensor,
            grad_state: torch.Tensor, batch: int, dtype: str, want_input: bool,
            param_stride: Optional[int] = None) -> Tuple[torch.Tensor, Optional[torch.Tensor]]:
        _hold(self)
        dev = device()
        code = _CODE[dtype]
        p = None if self.nparams == 0 else params.to(device=dev, dtype=_RDT[dtype]).contiguous()
        pstride = self.nparams if param_stride is None else param_stride
        st = state.to(_CDT[dtype]).contiguous()
        gs = grad_state.to(_CDT[dtype]).contiguous()
        ck = None if (ckpts is None or self.nckpt == 0) else ckpts.to(_CDT[dtype]).contiguous()
        gp = torch.empty(batch, max(self.nparams, 1), dtype=torch.float64, device=dev)
        gi = torch.empty_like(st) if want_input else None
        nb = _i64()
        check(lib.tcm_adjoint_workspace_bytes(self.handle, None, code, batch, 1, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_adjoint_vjp(self.handle, code, _ptr(p), pstride, _ptr(ck), st.data_ptr(), gs.data_ptr(),
                                  batch, gp.data_ptr() if self.nparams else None, _ptr(gi), ws.data_ptr(),
                                  ws.numel(), _stream()), "tcm_adjoint_vjp")
        return gp[:, : self.nparams], gi

    # --- fused energy + gradient ---------------------------------------------
    def energy_grad(self, params: torch.Tensor, ham: "PauliSum", batch: int, dtype: str = "complex64",
                    inputs: Optional[torch.Tensor] = None,
                    param_stride: Optional[int] = None) -> Tuple[torch.Tensor, torch.Tensor]:
        _hold(self, ham)
        dev = device()
        code = _CODE[dtype]
        p = params.to(device=dev, dtype=_RDT[dtype]).contiguous() if self.nparams else None
        pstride = self.nparams if param_stride is None else param_stride
        inp = None if inputs is None else inputs.to(device=dev, dtype=_CDT[dtype]).contiguous()
        e = torch.empty(batch, dtype=torch.float64, device=dev)
        g = torch.empty(batch, max(self.nparams, 1), dtype=torch.float64, device=dev)
        nb = _i64()
        check(lib.tcm_adjoint_workspace_bytes(self.handle, ham.handle, code, batch, 0, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_adjoint_energy_grad(self.handle, code, _ptr(p), pstride, ham.handle, _ptr(inp), batch,
                                          e.data_ptr(), g.data_ptr(), ws.data_ptr(), ws.numel(), _stream()),
              "tcm_adjoint_energy_grad")
        return e, g[:, : self.nparams]

    def adjoint_workspace_bytes(self, ham: Optional["PauliSum"], batch: int, dtype: str, mode: int) -> int:
        nb = _i64()
        check(lib.tcm_adjoint_workspace_bytes(self.handle, None if ham is None else ham.handle, _CODE[dtype],
                                              batch, mode, C.byref(nb)), "workspace")
        return nb.value


def forward_multi(items: Sequence[Tuple[Program, Optional[torch.Tensor], torch.Tensor]],
                  dtype: str = "complex64", init_zero: bool = True) -> None:
    """Run many programs (e.g. circuits of different qubit counts) on their
    states [batch, 2^n] in pass-synchronous launches (tcm_forward_multi)."""
    _hold(*(it[0] for it in items))
    dev = device()
    code = _CODE[dtype]
    descs = (_lib.TcmMultiDesc * max(len(items), 1))()
    keep = []
    for i, (prog, params, state) in enumerate(items):
        b = state.shape[0]
        p = None
        if prog.nparams:
            p = params.to(device=dev, dtype=_RDT[dtype]).reshape(b, prog.nparams).contiguous()
            keep.append(p)
        descs[i].prog = prog.handle.value
        descs[i].params = None if p is None else p.data_ptr()
        descs[i].param_stride = prog.nparams
        descs[i].state = state.data_ptr()
        descs[i].batch = b
    nb = _i64()
    check(lib.tcm_multi_workspace_bytes(descs, len(items), code, C.byref(nb)), "workspace")
    ws = _workspace(nb.value, dev)
    check(lib.tcm_forward_multi(descs, len(items), code, 1 if init_zero else 0, ws.data_ptr(), ws.numel(),
                                _stream()), "tcm_forward_multi")


_TUNE: Dict[str, Any] = {}
_TUNE_FILE = None


class MultiPlan:
    """Compiled multi-program forward launches for a fixed list of programs
    (tcm_multi_plan_build): pass r of every program runs in one kernel, so
    many small circuits share the GPU with large ones instead of running one
    after another.  ``forward(items)`` takes (program, params, state) in the
    plan's program order."""

    def __init__(self, programs: Sequence[Program], dtype: str = "complex64",
                 batches: Optional[Sequence[int]] = None):
        """``batches``: the expected batch of each program (weights the
        kernels' block shape toward the programs with the most amplitudes)."""
        self.programs = list(programs)
        self.dtype = dtype
        arr = (C.c_void_p * len(self.programs))(*[p.handle.value for p in self.programs])
        bt = None if batches is None else (C.c_int64 * len(self.programs))(*[int(b) for b in batches])
        h = C.c_void_p()
        sec = C.c_double(0.0)
        check(lib.tcm_multi_plan_build(arr, bt, len(self.programs), _CODE[dtype], C.byref(h), C.byref(sec)),
              "tcm_multi_plan_build")
        self.handle = h
        self.compile_seconds = sec.value

    def __del__(self, _free=lib.tcm_multi_plan_free):
        h = getattr(self, "handle", None)
        if h:
            _free(h)
            self.handle = None

    def forward(self, items: Sequence[Tuple[Program, Optional[torch.Tensor], torch.Tensor]],
                init_zero: bool = True) -> None:
        _hold(self, *self.programs)
        dev = device()
        code = _CODE[self.dtype]
        descs = (_lib.TcmMultiDesc * max(len(items), 1))()
        keep = []
        for i, (prog, params, state) in enumerate(items):
            b = state.shape[0]
            p = None
            if prog.nparams:
                p = params.to(device=dev, dtype=_RDT[self.dtype]).reshape(b, prog.nparams).contiguous()
                keep.append(p)
            descs[i].prog = prog.handle.value
            descs[i].params = _ptr(p)
            descs[i].param_stride = prog.nparams
            descs[i].state = state.data_ptr()
            descs[i].batch = b
        nb = _i64()
        check(lib.tcm_multi_workspace_bytes(descs, len(items), code, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_forward_multi_plan(self.handle, descs, len(items), 1 if init_zero else 0, ws.data_ptr(),
                                         ws.numel(), _stream()), "tcm_forward_multi_plan")


def _tune_path() -> str:
    import os

    d = os.environ.get("TCM_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "tensorcircuit_amd")
    return os.path.join(d, "autotune.json")


def tune_key(opts: Dict[str, int]) -> str:
    """Name of a (tile bits, register slots) choice; {} is "default"."""
    if not opts:
        return "default"
    return "k%d_r%d" % (opts.get("tile_bits", 0), opts.get("reg_slots", 0))


def _time_best(fn: Any, reps: int) -> float:
    """Best of ``reps`` single runs in ms by HIP events on the current
    stream (one slow run, e.g. a clock ramp, does not decide)."""
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def autotune_program(n: int, specs: Sequence[GateSpec], nparams: int, batch: int, dtype: str = "complex64",
                     tiles: Sequence[int] = (12, 13), slots: Sequence[int] = (4, 5), reps: int = 3,
                     margin: float = 0.03, **opts: Any) -> Tuple[Dict[str, int], Dict[str, float]]:
    """Pick the fusion width (register slots: qubits whose gates apply from
    registers between LDS exchanges) and the tile (log2 amplitudes per LDS
    tile) of a program by timing its forward on the real batch: the default
    configuration and every (tile, slots) candidate, best of ``reps``.  A
    candidate replaces the default only when faster by more than ``margin``
    (so a tuned program is never slower than the default by this
    measurement).  Returns (options for Program, ms per candidate); the
    choice is cached on disk keyed by (n, dtype, batch, structure), and a
    cached call returns an empty timing dict.  Staging cost, reported apart
    from running time like the reference's jit staging
    (utils.py:205-232)."""
    import hashlib

    global _TUNE_FILE
    jit = int(opts.get("jit", 0))
    h = hashlib.sha1(repr([s.key() for s in specs]).encode()).hexdigest()[:16]
    key = "prog/%d/%s/%d/%d/%s" % (n, dtype, batch, jit, h)
    if _TUNE_FILE is None:
        _TUNE_FILE = _tune_path()
        try:
            with open(_TUNE_FILE) as f:
                _TUNE.update(json.load(f))
        except Exception:
            pass
    if key in _TUNE:
        return dict(_TUNE[key]), {}
    dev = device()
    p = torch.zeros(batch, max(nparams, 1), dtype=_RDT[dtype], device=dev)[:, :nparams] if nparams else None
    st = torch.empty(batch, 1 << n, dtype=_CDT[dtype], device=dev)
    cands: List[Dict[str, int]] = [{}]
    for k in tiles:
        for r in (slots if jit else (4,)):
            if k <= n and k - r >= 6:
                cands.append({"tile_bits": int(k), "reg_slots": int(r)})
    times: Dict[str, float] = {}
    for c in cands:
        try:
            prog = Program(n, specs, nparams, fuse=1, **dict(opts, **c))
            if jit:
                prog.compile(dtype, 0)
            times[tune_key(c)] = _time_best(lambda: prog.forward(p, batch, dtype, out=st), reps)
        except (ValueError, NotImplementedError):
            continue  # a configuration this program cannot use
    best = min(cands, key=lambda c: times.get(tune_key(c), float("inf")))
    if times.get(tune_key(best), float("inf")) > times["default"] * (1.0 - margin):
        best = {}
    _TUNE[key] = best
    try:
        os.makedirs(os.path.dirname(_TUNE_FILE), exist_ok=True)
        with open(_TUNE_FILE, "w") as f:
            json.dump(_TUNE, f)
    except Exception:
        pass
    return dict(best), times


def tune_step(candidates: Dict[str, Any], default: str, reps: int = 3,
              margin: float = 0.03) -> Tuple[str, Dict[str, float]]:
    """Step-level choice: time each candidate step (a callable enqueuing a
    whole step on the current stream), best of ``reps``, and return the
    fastest; a candidate replaces ``default`` only when faster than it by
    more than ``margin``.  Returns (name, ms per candidate)."""
    times = {name: _time_best(fn, reps) for name, fn in candidates.items()}
    best = min(times, key=times.get)
    if times[best] > times[default] * (1.0 - margin):
        best = default
    return best, times


_PROG_CACHE: "collections.OrderedDict[Tuple, Program]" = collections.OrderedDict()
_PROG_LOCK = threading.Lock()
_PROG_CACHE_MAX = 128


def get_program(n: int, specs: Sequence[GateSpec], nparams: int, **opts: Any) -> Program:
    """Build-once cache keyed by structure (gate kinds, qubits, constant
    matrices, parameter slots) and options -- the reference rebuilds its
    tensor network on every call (basecircuit.py:120-245)."""
    if "jit" not in opts and opts.get("fuse", 1):
        from . import cons

        opts["jit"] = cons.jit_for(n)
    key = (n, nparams, tuple(s.key() + (s.param,) for s in specs), tuple(sorted(opts.items())))
    with _PROG_LOCK:
        p = _PROG_CACHE.get(key)
        if p is not None:
            _PROG_CACHE.move_to_end(key)
            return p
    p = Program(n, specs, nparams, **opts)
    with _PROG_LOCK:
        _PROG_CACHE[key] = p
        while len(_PROG_CACHE) > _PROG_CACHE_MAX:
            _PROG_CACHE.popitem(last=False)
    return p


# ---------------------------------------------------------------------------
# Pauli sums
# ---------------------------------------------------------------------------


class PauliSum:
    """Sum_k w_k P_k with P_k given by codes 0=I,1=X,2=Y,3=Z per qubit
    (the ``ps`` structures of quantum.py:1025-1068 / any_measurements)."""

    def __init__(self, n: int, codes: Any, coeffs: Any, jit: Optional[int] = None):
        """``jit``: compile the tiled passes (1), keep the generic kernel (0),
        or follow the package policy for the qubit count (None)."""
        codes = np.ascontiguousarray(np.asarray(codes, dtype=np.int8).reshape(-1, n))
        coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64).reshape(-1))
        if codes.shape[0] != coeffs.shape[0]:
            raise ValueError("codes and coeffs disagree on the number of terms")
        self.n = n
        self.codes = codes
        self.coeffs = coeffs
        h = C.c_void_p()
        check(lib.tcm_pauli_build(codes.ctypes.data_as(C.POINTER(C.c_int8)),
                                  coeffs.ctypes.data_as(C.POINTER(C.c_double)), codes.shape[0], n, C.byref(h)),
              "tcm_pauli_build")
        self.handle = h
        if jit is None:
            from . import cons

            jit = cons.jit_for(n)
        self.jit = int(jit)
        self._compiled = set()

    def compile(self, dtype: str = "complex64") -> float:
        """Compile the tiled passes for dtype now; returns compile seconds."""
        sec = C.c_double(0.0)
        check(lib.tcm_pauli_compile(self.handle, _CODE[dtype], C.byref(sec)), "tcm_pauli_compile")
        self._compiled.add(_CODE[dtype])
        return sec.value

    def _prepare(self, code: int) -> None:
        if self.jit and code not in self._compiled:
            check(lib.tcm_pauli_compile(self.handle, code, None), "tcm_pauli_compile")
            self._compiled.add(code)

    def __del__(self, _free=lib.tcm_pauli_free):
        h = getattr(self, "handle", None)
        if h:
            _free(h)
            self.handle = None

    def expect(self, state: torch.Tensor) -> torch.Tensor:
        """complex128 [batch] of <psi|H|psi> (fp64 accumulation)."""
        _hold(self)
        dev = device()
        code = _CODE["complex64" if state.dtype == torch.complex64 else "complex128"]
        self._prepare(code)
        s = state.reshape(-1, 1 << self.n).contiguous()
        b = s.shape[0]
        out = torch.empty(b, 2, dtype=torch.float64, device=dev)
        nb = _i64()
        check(lib.tcm_pauli_workspace_bytes(self.handle, code, b, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_expect_pauli_sum(self.handle, code, s.data_ptr(), b, out.data_ptr(), ws.data_ptr(),
                                       ws.numel(), _stream()), "tcm_expect_pauli_sum")
        return torch.view_as_complex(out)

    def apply(self, state: torch.Tensor, alpha: Optional[torch.Tensor] = None) -> torch.Tensor:
        """alpha_b * H psi_b as a new state tensor."""
        _hold(self)
        code = _CODE["complex64" if state.dtype == torch.complex64 else "complex128"]
        self._prepare(code)
        s = state.reshape(-1, 1 << self.n).contiguous()
        out = torch.empty_like(s)
        a = None if alpha is None else alpha.to(torch.float64).contiguous()
        check(lib.tcm_apply_pauli_sum(self.handle, code, s.data_ptr(), out.data_ptr(), s.shape[0], _ptr(a),
                                      _stream()), "tcm_apply_pauli_sum")
        return out


    def expect_terms(self, state: torch.Tensor) -> torch.Tensor:
        """complex128 [batch, nterms]: w_t <psi_b|P_t|psi_b> for every term, in
        the order given, from one launch (tcm_expect_pauli_terms)."""
        _hold(self)
        dev = device()
        code = _CODE["complex64" if state.dtype == torch.complex64 else "complex128"]
        s = state.reshape(-1, 1 << self.n).contiguous()
        b = s.shape[0]
        nt = self.codes.shape[0]
        out = torch.empty(b, nt, 2, dtype=torch.float64, device=dev)
        nb = _i64()
        check(lib.tcm_pauli_terms_workspace_bytes(self.handle, b, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, dev)
        check(lib.tcm_expect_pauli_terms(self.handle, code, s.data_ptr(), b, out.data_ptr(), ws.data_ptr(),
                                         ws.numel(), _stream()), "tcm_expect_pauli_terms")
        return torch.view_as_complex(out)

    def apply_terms(self, state: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
        """sum_t weights[b, t] w_t P_t psi_b as a new state tensor
        (tcm_apply_pauli_terms); weights real [batch, nterms]."""
        _hold(self)
        code = _CODE["complex64" if state.dtype == torch.complex64 else "complex128"]
        s = state.reshape(-1, 1 << self.n).contiguous()
        w = weights.to(device=s.device, dtype=torch.float64).reshape(s.shape[0], self.codes.shape[0]).contiguous()
        out = torch.empty_like(s)
        check(lib.tcm_apply_pauli_terms(self.handle, code, s.data_ptr(), out.data_ptr(), s.shape[0], w.data_ptr(),
                                        _stream()), "tcm_apply_pauli_terms")
        return out


_PAULI_CACHE: "collections.OrderedDict[Tuple, PauliSum]" = collections.OrderedDict()


def get_pauli(n: int, codes: Any, coeffs: Any) -> PauliSum:
    codes = np.asarray(codes, dtype=np.int8).reshape(-1, n)
    coeffs = np.asarray(coeffs, dtype=np.float64).reshape(-1)
    key = (n, codes.tobytes(), coeffs.tobytes())
    with _PROG_LOCK:
        h = _PAULI_CACHE.get(key)
        if h is not None:
            _PAULI_CACHE.move_to_end(key)
            return h
    h = PauliSum(n, codes, coeffs)
    with _PROG_LOCK:
        _PAULI_CACHE[key] = h
        while len(_PAULI_CACHE) > 256:
            _PAULI_CACHE.popitem(last=False)
    return h


def inner(a: torch.Tensor, b: torch.Tensor, n: int) -> torch.Tensor:
    """complex128 [batch] of <a|b>."""
    dev = device()
    code = _CODE["complex64" if a.dtype == torch.complex64 else "complex128"]
    a2 = a.reshape(-1, 1 << n).contiguous()
    b2 = b.to(a.dtype).reshape(-1, 1 << n).contiguous()
    bt = a2.shape[0]
    out = torch.empty(bt, 2, dtype=torch.float64, device=dev)
    nb = _i64()
    check(lib.tcm_inner_workspace_bytes(n, bt, C.byref(nb)), "workspace")
    ws = _workspace(nb.value, dev)
    check(lib.tcm_inner(n, code, a2.data_ptr(), b2.data_ptr(), bt, out.data_ptr(), ws.data_ptr(), ws.numel(),
                        _stream()), "tcm_inner")
    return torch.view_as_complex(out)


# ---------------------------------------------------------------------------
# measurement
# ---------------------------------------------------------------------------


def _states(state: torch.Tensor, n: int) -> Tuple[torch.Tensor, int]:
    dev = device()
    if state.dtype not in (torch.complex64, torch.complex128):
        state = state.to(torch.complex64)
    s = state.to(dev).reshape(-1, 1 << n).contiguous()
    return s, _CODE["complex64" if s.dtype == torch.complex64 else "complex128"]


def probability(state: torch.Tensor, n: int) -> torch.Tensor:
    """|psi|^2 in the state's real dtype, shape [batch, 2^n] (tcm_probability)."""
    s, code = _states(state, n)
    out = torch.empty(s.shape, dtype=_RDT["complex64" if code == _lib.TCM_C64 else "complex128"], device=s.device)
    check(lib.tcm_probability(n, code, s.data_ptr(), s.shape[0], out.data_ptr(), _stream()), "tcm_probability")
    return out


def sample_indices(state: torch.Tensor, n: int, uniforms: torch.Tensor) -> torch.Tensor:
    """int64 [batch, shots]: inverse-CDF draws from |psi_b|^2 for the uniforms
    [batch, shots] in [0, 1] (tcm_sample; the reference's
    backend.probability_sample, abstract_backend.py:1124-1157)."""
    s, code = _states(state, n)
    bt = s.shape[0]
    u = uniforms.to(device=s.device, dtype=torch.float64).reshape(bt, -1).contiguous()
    shots = u.shape[1]
    out = torch.empty(bt, shots, dtype=torch.int64, device=s.device)
    nb = _i64()
    check(lib.tcm_sample_workspace_bytes(n, bt, C.byref(nb)), "workspace")
    ws = _workspace(nb.value, s.device)
    check(lib.tcm_sample(n, code, s.data_ptr(), bt, u.data_ptr(), shots, out.data_ptr(), ws.data_ptr(), ws.numel(),
                         _stream()), "tcm_sample")
    return out


def pauli_coo(ham: "PauliSum", dtype: str = "complex64") -> torch.Tensor:
    """The Pauli sum as a coalesced 2^n x 2^n torch sparse COO matrix on the
    device; entries assembled by tcm_pauli_coo (ps2coo_core,
    quantum.py:1461-1485), duplicates summed (PauliStringSum2COO)."""
    dev = device()
    nnz = _i64()
    check(lib.tcm_pauli_nnz(ham.handle, C.byref(nnz)), "tcm_pauli_nnz")
    N = 1 << ham.n
    rows = torch.empty(nnz.value, dtype=torch.int64, device=dev)
    cols = torch.empty(nnz.value, dtype=torch.int64, device=dev)
    vals = torch.empty(nnz.value, dtype=_CDT[dtype], device=dev)
    if nnz.value:
        check(lib.tcm_pauli_coo(ham.handle, _CODE[dtype], rows.data_ptr(), cols.data_ptr(), vals.data_ptr(),
                                _stream()), "tcm_pauli_coo")
    return torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, (N, N)).coalesce()


class CsrMatrix:
    """A 2^n x 2^n sparse matrix on the device in CSR form (int64 row
    pointers and columns, complex values of the state dtype), with the CSR
    form of its conjugate transpose for the backward pass."""

    def __init__(self, h: torch.Tensor, n: int, dtype: torch.dtype):
        # format conversions of a constant matrix: run with torch.func's
        # transforms off, so that a matrix captured from outside a grad /
        # vmap stays a plain tensor
        import warnings

        with torch._C._DisableFuncTorch(), warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # CSR support is "beta" in torch
            self._build(h, n, dtype)

    def _build(self, h: torch.Tensor, n: int, dtype: torch.dtype) -> None:
        dev = device()
        hc = h.detach().to(device=dev, dtype=dtype)
        if hc.layout == torch.sparse_coo:
            hc = hc.coalesce()
        elif hc.layout != torch.sparse_csr:
            hc = hc.to_sparse()
        N = 1 << n
        if tuple(hc.shape) != (N, N):
            raise ValueError("sparse matrix of shape %s for %d qubits" % (tuple(hc.shape), n))
        self.n = n
        self.dtype = dtype
        self.fwd = self._csr(hc)
        # A^dagger: swap the coordinates of the COO form, conjugate the values
        coo = hc.to_sparse_coo().coalesce() if hc.layout != torch.sparse_coo else hc
        idx = coo.indices()
        dag = torch.sparse_coo_tensor(torch.stack([idx[1], idx[0]]), torch.conj(coo.values()).resolve_conj(),
                                      (N, N)).coalesce()
        self.dag = self._csr(dag)

    @staticmethod
    def _csr(h: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        c = h.to_sparse_csr() if h.layout != torch.sparse_csr else h
        return (c.crow_indices().to(torch.int64).contiguous(), c.col_indices().to(torch.int64).contiguous(),
                c.values().contiguous())

    def expect(self, st: torch.Tensor) -> torch.Tensor:
        """complex128 <psi|A|psi> (tcm_csr_expect, fp64 accumulation)."""
        _hold(self)
        s, code = _states(st, self.n)
        rowptr, cols, vals = self.fwd
        out = torch.empty(2, dtype=torch.float64, device=s.device)
        nb = _i64()
        check(lib.tcm_csr_expect_workspace_bytes(self.n, C.byref(nb)), "workspace")
        ws = _workspace(nb.value, s.device)
        check(lib.tcm_csr_expect(self.n, code, rowptr.data_ptr(), cols.data_ptr(), vals.data_ptr(), s.data_ptr(),
                                 out.data_ptr(), ws.data_ptr(), ws.numel(), _stream()), "tcm_csr_expect")
        return torch.complex(out[0], out[1])

    def matvec(self, x: torch.Tensor, dagger: bool = False) -> torch.Tensor:
        """A x_b (or A^dagger x_b) for x [batch, 2^n] (tcm_csr_matvec)."""
        _hold(self)
        s, code = _states(x, self.n)
        rowptr, cols, vals = self.dag if dagger else self.fwd
        out = torch.empty_like(s)
        check(lib.tcm_csr_matvec(self.n, code, rowptr.data_ptr(), cols.data_ptr(), vals.data_ptr(), s.data_ptr(),
                                 out.data_ptr(), s.shape[0], _stream()), "tcm_csr_matvec")
        return out


def csr_expectation(h: Any, state: torch.Tensor, n: int) -> torch.Tensor:
    """complex <psi|H|psi> for a sparse (COO or CSR) H by tcm_csr_expect,
    differentiable in psi (CsrExpFn)."""
    st = state.reshape(-1)
    m = h if isinstance(h, CsrMatrix) else CsrMatrix(h, n, st.dtype)
    return CsrExpFn.apply(st, m)


def sample_prob_indices(p: torch.Tensor, n: int, uniforms: torch.Tensor) -> torch.Tensor:
    """int64 [batch, shots]: inverse-CDF draws from probability vectors p
    [batch, 2^n] (normalised by their sum) for the uniforms [batch, shots]
    (tcm_sample_prob)."""
    dev = device()
    pt = p.to(dev)
    if pt.dtype not in (torch.float32, torch.float64):
        pt = pt.to(torch.float32)
    pt = pt.reshape(-1, 1 << n).contiguous()
    code = _lib.TCM_C64 if pt.dtype == torch.float32 else _lib.TCM_C128
    bt = pt.shape[0]
    u = uniforms.to(device=dev, dtype=torch.float64).reshape(bt, -1).contiguous()
    shots = u.shape[1]
    out = torch.empty(bt, shots, dtype=torch.int64, device=dev)
    nb = _i64()
    check(lib.tcm_sample_workspace_bytes(n, bt, C.byref(nb)), "workspace")
    ws = _workspace(nb.value, dev)
    check(lib.tcm_sample_prob(n, code, pt.data_ptr(), bt, u.data_ptr(), shots, out.data_ptr(), ws.data_ptr(),
                              ws.numel(), _stream()), "tcm_sample_prob")
    return out


def measure(state: torch.Tensor, n: int, qubits: Sequence[int], status: torch.Tensor
            ) -> Tuple[torch.Tensor, torch.Tensor]:
    """Sequential measurement of `qubits` with external randomness status
    [batch, len(qubits)] (tcm_measure; measure_jit, basecircuit.py:371-445):
    outcomes fp64 [batch, k] and the probability of that record fp64 [batch]."""
    s, code = _states(state, n)
    bt = s.shape[0]
    q = (C.c_int32 * max(len(qubits), 1))(*[int(x) for x in qubits])
    st = status.to(device=s.device, dtype=torch.float64).reshape(bt, len(qubits)).contiguous()
    outc = torch.empty(bt, len(qubits), dtype=torch.float64, device=s.device)
    prob = torch.ones(bt, dtype=torch.float64, device=s.device)
    nb = _i64()
    check(lib.tcm_measure_workspace_bytes(n, bt, C.byref(nb)), "workspace")
    ws = _workspace(nb.value, s.device)
    check(lib.tcm_measure(n, code, s.data_ptr(), bt, q, len(qubits), st.data_ptr(), outc.data_ptr(), prob.data_ptr(),
                          ws.data_ptr(), ws.numel(), _stream()), "tcm_measure")
    return outc, prob


# ---------------------------------------------------------------------------
# autograd bridges with vmap rules
# ---------------------------------------------------------------------------


def _to_batch(x: Optional[torch.Tensor], dim: Optional[int], B: int, tail: Tuple[int, ...]) -> Optional[torch.Tensor]:
    """Physical batched view [B, *tail] of a vmap argument."""
    if x is None:
        return None
    if dim is None:
        return x.reshape((1,) + tail).expand((B,) + tail).contiguous()
    return x.movedim(dim, 0).reshape((B,) + tail).contiguous()


class StateFn(torch.autograd.Function):
    """params [P] (and optional input state [N]) -> (psi [N], checkpoints).
    Backward = adjoint VJP through StateVJPFn (which has its own vmap rule)."""

    @staticmethod
    def forward(params, inputs, prog, dtype):
        N = 1 << prog.n
        st, ck = prog.forward(params.reshape(1, -1) if prog.nparams else None, 1, dtype,
                              inputs=None if inputs is None else inputs.reshape(1, N), want_ckpts=prog.nckpt > 0)
        return st.reshape(N), ck.reshape(ck.shape[0], N)

    @staticmethod
    def setup_context(ctx, inputs, output):
        params, inp, prog, dtype = inputs
        st, ck = output
        ctx.save_for_backward(params, st, ck)
        ctx.prog, ctx.dtype, ctx.has_input = prog, dtype, inp is not None
        ctx.mark_non_differentiable(ck)

    @staticmethod
    def backward(ctx, g_state, g_ck):
        params, st, ck = ctx.saved_tensors
        want_input = bool(ctx.has_input and ctx.needs_input_grad[1])
        gp, gi = StateVJPFn.apply(params, ck, st, g_state, ctx.prog, ctx.dtype, want_input)
        return (gp if ctx.needs_input_grad[0] else None), (gi if want_input else None), None, None

    @staticmethod
    def vmap(info, in_dims, params, inputs, prog, dtype):
        B = info.batch_size
        N = 1 << prog.n
        P = params.shape[-1] if params.dim() else 0
        p = _to_batch(params, in_dims[0], B, (P,))
        inp = _to_batch(inputs, in_dims[1], B, (N,)) if inputs is not None else None
        # one batched evaluation that regular autograd can also see through
        st, ck = StateBatchFn.apply(p, inp, prog, dtype)
        return (st, ck), (0, 1)


class StateBatchFn(torch.autograd.Function):
    """params [B, P], inputs [B, N] or None -> (psi [B, N], checkpoints
    [C, B, N]) in one batched launch; backward = batched reverse sweep."""

    @staticmethod
    def forward(params, inputs, prog, dtype):
        B = params.shape[0]
        return prog.forward(params if prog.nparams else None, B, dtype, inputs=inputs,
                            want_ckpts=prog.nckpt > 0)

    @staticmethod
    def setup_context(ctx, inputs, output):
        params, inp, prog, dtype = inputs
        st, ck = output
        ctx.save_for_backward(params, st, ck)
        ctx.prog, ctx.dtype, ctx.has_input = prog, dtype, inp is not None
        ctx.mark_non_differentiable(ck)

    @staticmethod
    def backward(ctx, g_state, g_ck):
        params, st, ck = ctx.saved_tensors
        prog = ctx.prog
        want_input = bool(ctx.has_input and ctx.needs_input_grad[1])
        B = params.shape[0]
        gp, gi = prog.vjp(params if prog.nparams else None, ck, st, g_state.contiguous(), B, ctx.dtype, want_input)
        gp = gp.to(params.dtype).reshape(params.shape) if ctx.needs_input_grad[0] else None
        return gp, (gi if want_input else None), None, None


class StateVJPFn(torch.autograd.Function):
    """(params, ckpts, psi, dL/dpsi) -> (dL/dparams, dL/dinput) by the reverse sweep."""

    @staticmethod
    def forward(params, ck, st, g_state, prog, dtype, want_input):
        N = 1 << prog.n
        gp, gi = prog.vjp(params.reshape(1, -1) if prog.nparams else None, ck.reshape(ck.shape[0], 1, N),
                          st.reshape(1, N), g_state.reshape(1, N), 1, dtype, want_input)
        gp = gp.reshape(params.shape).to(params.dtype)
        gi = gi.reshape(-1) if want_input else torch.zeros(0, dtype=st.dtype, device=st.device)
        return gp, gi

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("higher-order derivatives through the adjoint sweep are not supported")

    @staticmethod
    def vmap(info, in_dims, params, ck, st, g_state, prog, dtype, want_input):
        B = info.batch_size
        N = 1 << prog.n
        P = params.shape[-1] if params.dim() else 0
        p = _to_batch(params, in_dims[0], B, (P,)) if prog.nparams else None
        C_ = ck.shape[0] if in_dims[1] != 0 else ck.shape[1]
        if in_dims[1] is None:
            ckb = ck.reshape(C_, 1, N).expand(C_, B, N).contiguous()
        else:
            ckb = ck.movedim(in_dims[1], 1).reshape(C_, B, N).contiguous()
        s = _to_batch(st, in_dims[2], B, (N,))
        g = _to_batch(g_state, in_dims[3], B, (N,))
        gp, gi = prog.vjp(p, ckb, s, g, B, dtype, want_input)
        gp = gp.to(params.dtype).reshape((B,) + tuple(params.shape[-1:]) if params.dim() else (B,))
        if not want_input:
            gi = torch.zeros(B, 0, dtype=s.dtype, device=s.device)
        return (gp, gi), (0, 0)


class PauliExpFn(torch.autograd.Function):
    """psi [N] -> <psi|H|psi> (complex, the state dtype), fp64-accumulated."""

    @staticmethod
    def forward(st, ham):
        return ham.expect(st.reshape(1, -1)).to(st.dtype).reshape(())

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, ham = inputs
        ctx.save_for_backward(st)
        ctx.ham = ham

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        return PauliExpVJPFn.apply(st, g, ctx.ham), None

    @staticmethod
    def vmap(info, in_dims, st, ham):
        B = info.batch_size
        s = _to_batch(st, in_dims[0], B, (1 << ham.n,))
        return PauliExpBatchFn.apply(s, ham), 0


class PauliExpBatchFn(torch.autograd.Function):
    """psi [B, N] -> <psi_b|H|psi_b> [B]; backward 2 Re(g_b) H psi_b."""

    @staticmethod
    def forward(st, ham):
        return ham.expect(st).to(st.dtype)

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, ham = inputs
        ctx.save_for_backward(st)
        ctx.ham = ham

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        alpha = (2.0 * torch.real(g)).to(torch.float64).reshape(-1).to(st.device)
        return ctx.ham.apply(st, alpha), None


class PauliExpVJPFn(torch.autograd.Function):
    """dL/dpsi = 2 Re(g) H psi for L depending on E = <psi|H|psi> (H Hermitian)."""

    @staticmethod
    def forward(st, g, ham):
        alpha = (2.0 * torch.real(g)).to(torch.float64).reshape(1).to(st.device)
        return ham.apply(st.reshape(1, -1), alpha).reshape(st.shape)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("higher-order derivatives of expectations are not supported")

    @staticmethod
    def vmap(info, in_dims, st, g, ham):
        B = info.batch_size
        N = 1 << ham.n
        s = _to_batch(st, in_dims[0], B, (N,))
        gg = _to_batch(g, in_dims[1], B, ())
        alpha = (2.0 * torch.real(gg)).to(torch.float64).reshape(B).to(s.device)
        return ham.apply(s, alpha), 0


class OpExpFn(torch.autograd.Function):
    """psi -> <psi|O|psi> for a general operator O given as two programs
    (O and O^dagger); dL/dpsi = conj(g) O psi + g O^dagger psi."""

    @staticmethod
    def forward(st, prog, prog_dag, dtype):
        phi, _ = prog.forward(None, 1, dtype, inputs=st.reshape(1, -1))
        return inner(st.reshape(1, -1), phi, prog.n).to(st.dtype).reshape(())

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, prog, prog_dag, dtype = inputs
        ctx.save_for_backward(st)
        ctx.progs = (prog, prog_dag, dtype)

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        prog, prog_dag, dtype = ctx.progs
        with torch.no_grad():
            a, _ = prog.forward(None, 1, dtype, inputs=st.reshape(1, -1))
            b, _ = prog_dag.forward(None, 1, dtype, inputs=st.reshape(1, -1))
            gg = g.to(st.dtype)
            out = (torch.conj(gg) * a + gg * b).reshape(st.shape)
        return out, None, None, None


class CsrExpFn(torch.autograd.Function):
    """psi [N] -> <psi|A|psi> for a sparse A (complex, the state dtype);
    dL/dpsi = conj(g) A psi + g A^dagger psi (the same convention as
    OpExpFn), both products by the CSR mat-vec kernel."""

    @staticmethod
    def forward(st, m):
        return m.expect(st).to(st.dtype).reshape(())

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, m = inputs
        ctx.save_for_backward(st)
        ctx.m = m

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        return CsrVJPFn.apply(st, g, ctx.m), None

    @staticmethod
    def vmap(info, in_dims, st, m):
        B = info.batch_size
        s = _to_batch(st, in_dims[0], B, (1 << m.n,))
        return torch.stack([CsrExpFn.apply(s[b], m) for b in range(B)]), 0


class CsrVJPFn(torch.autograd.Function):
    @staticmethod
    def forward(st, g, m):
        s = st.reshape(1, -1)
        gg = g.to(st.dtype).reshape(())
        return (torch.conj(gg) * m.matvec(s) + gg * m.matvec(s, dagger=True)).reshape(st.shape)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("higher-order derivatives of sparse expectations are not supported")

    @staticmethod
    def vmap(info, in_dims, st, g, m):
        B = info.batch_size
        N = 1 << m.n
        s = _to_batch(st, in_dims[0], B, (N,))
        gg = _to_batch(g, in_dims[1], B, ()).to(s.dtype).reshape(B, 1)
        return torch.conj(gg) * m.matvec(s) + gg * m.matvec(s, dagger=True), 0


class PauliTermsFn(torch.autograd.Function):
    """psi [N] -> [nterms] of w_t <psi|P_t|psi> (complex, the state dtype)
    from one launch; backward sum_t 2 Re(g_t) w_t P_t psi (one launch)."""

    @staticmethod
    def forward(st, ham):
        return ham.expect_terms(st.reshape(1, -1)).to(st.dtype).reshape(-1)

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, ham = inputs
        ctx.save_for_backward(st)
        ctx.ham = ham

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        return PauliTermsVJPFn.apply(st, g, ctx.ham), None

    @staticmethod
    def vmap(info, in_dims, st, ham):
        B = info.batch_size
        s = _to_batch(st, in_dims[0], B, (1 << ham.n,))
        return PauliTermsBatchFn.apply(s, ham), 0


class PauliTermsBatchFn(torch.autograd.Function):
    """psi [B, N] -> [B, nterms]; backward sum_t 2 Re(g_bt) w_t P_t psi_b."""

    @staticmethod
    def forward(st, ham):
        return ham.expect_terms(st).to(st.dtype)

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, ham = inputs
        ctx.save_for_backward(st)
        ctx.ham = ham

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        return ctx.ham.apply_terms(st, 2.0 * torch.real(g)), None


class PauliTermsVJPFn(torch.autograd.Function):
    @staticmethod
    def forward(st, g, ham):
        w = (2.0 * torch.real(g)).reshape(1, -1)
        return ham.apply_terms(st.reshape(1, -1), w).reshape(st.shape)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("higher-order derivatives of expectations are not supported")

    @staticmethod
    def vmap(info, in_dims, st, g, ham):
        B = info.batch_size
        N = 1 << ham.n
        s = _to_batch(st, in_dims[0], B, (N,))
        gg = _to_batch(g, in_dims[1], B, (ham.codes.shape[0],))
        return ham.apply_terms(s, 2.0 * torch.real(gg)), 0


class ProbFn(torch.autograd.Function):
    """psi [..., N] -> |psi|^2 (tcm_probability); backward 2 g psi (the torch
    convention for real losses of complex inputs)."""

    @staticmethod
    def forward(st, n):
        return probability(st, n).reshape(st.shape)

    @staticmethod
    def setup_context(ctx, inputs, output):
        st, _ = inputs
        ctx.save_for_backward(st)

    @staticmethod
    def backward(ctx, g):
        (st,) = ctx.saved_tensors
        return 2.0 * g.to(st.real.dtype) * st, None

    @staticmethod
    def vmap(info, in_dims, st, n):
        B = info.batch_size
        s = _to_batch(st, in_dims[0], B, (1 << n,))
        return ProbFn.apply(s, n), 0


# ---------------------------------------------------------------------------
# HIP graph capture of a whole step
# ---------------------------------------------------------------------------


class CapturedStep:
    """A step -- a callable that only enqueues device work on the current
    stream through this engine (forward / forward_multi / expectations /
    energy_grad) -- captured once into a HIP graph and replayed with one
    launch.  Inputs are the tensors the callable closes over: update them in
    place between replays; ``replay()`` returns the tensors the captured call
    returned, overwritten by each replay.

    The callable runs ``warmup`` times eagerly first (on a side stream of
    default priority) so that device tables, compiled kernels and workspace
    sizes exist before capture; the entry points record no memset nodes.

    Lifetime: the graph reads the device tables and compiled kernels of
    every Program, PauliSum and CsrMatrix the callable used during capture.
    The step holds references to all of them (and to the callable), so the
    engine's caches may evict them without freeing memory the graph still
    uses; they are released when the step is."""

    def __init__(self, fn, warmup: int = 2):
        device()
        self.fn = fn
        stack = getattr(_SCOPE, "stack", None)
        if stack is None:
            stack = _SCOPE.stack = []
        stack.append([])
        try:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(max(1, warmup)):
                    fn()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.out = fn()
        finally:
            used = stack.pop()
        # one reference per distinct object
        self.held = list({id(o): o for o in used}.values())

    def replay(self):
        self.graph.replay()
        return self.out


class UnevenBatch:
    """Many circuit structures (different gates or qubit counts), each with a
    batch of parameter sets, run forward -- with a Pauli-sum energy per state
    where a Hamiltonian is given -- as one step replayed from a HIP graph: the
    reference's vmap over circuit structures (backends/pytorch_backend.py:
    605-617) without one launch per structure and gate.

    Staging: every structure is compiled into pass kernels; those holding at
    least ``large_frac`` of all amplitudes get their fusion width (register
    slots) and tile bits autotuned on their real batch (autotune_program).
    Then whole candidate steps are captured as HIP graphs and timed
    (tune_step): the programs' passes in sequence; one compiled
    multi-program launch per pass for all (MultiPlan); the hybrid (large
    programs on their own passes, the small, launch-bound ones in one
    multi-program launch per pass); and the hybrid with the small part on a
    second stream (default priority) joined at the end of the step -- each
    with the default and with the tuned programs.  The default (passes in
    sequence, default programs) is kept unless a candidate is faster by
    ``margin``, so the chosen step is never slower by its own measurement.

    ``structures[i]`` = (n, specs, nparams).  ``params[i]`` ([batch_i,
    nparams_i], the dtype's real type) are inputs: update them in place
    between runs.  ``states[i]`` ([batch_i, 2^n_i]) are the outputs.
    ``run()`` replays the chosen step and returns the energies (a complex128
    [batch_i] tensor per structure, None where no Hamiltonian was given),
    overwritten by each run."""

    def __init__(self, structures: Sequence[Tuple[int, Sequence[GateSpec], int]], batches: Sequence[int],
                 hams: Optional[Sequence[Optional["PauliSum"]]] = None, dtype: str = "complex64",
                 params: Optional[Sequence[torch.Tensor]] = None,
                 options: Optional[Sequence[Dict[str, int]]] = None, tune: bool = True,
                 large_frac: float = 0.05, margin: float = 0.03, reps: int = 3):
        dev = device()
        k = len(structures)
        if len(batches) != k or k == 0:
            raise ValueError("one batch size per structure (at least one structure)")
        if dtype not in _CODE:
            raise ValueError("dtype must be complex64 or complex128")
        rdt = torch.float32 if dtype == "complex64" else torch.float64
        cdt = torch.complex64 if dtype == "complex64" else torch.complex128
        self.dtype = dtype
        self.sizes = [(int(n), int(b)) for (n, _, _), b in zip(structures, batches)]
        if params is None:
            params = [torch.zeros(b, int(P), dtype=rdt, device=dev) for (_, _, P), b in zip(structures, batches)]
        for (n, _, P), b, pa in zip(structures, batches, params):
            if tuple(pa.shape) != (int(b), int(P)) or pa.dtype != rdt or not pa.is_contiguous():
                raise ValueError("params[i] must be a contiguous [%d, %d] %s tensor" % (b, P, rdt))
        self.params = list(params)
        self.states = [torch.empty(b, 1 << n, dtype=cdt, device=dev) for n, b in self.sizes]
        self.hams = list(hams) if hams is not None else [None] * k
        if len(self.hams) != k:
            raise ValueError("one Hamiltonian (or None) per structure")
        opts = [dict(o) for o in options] if options is not None else [{} for _ in range(k)]
        if len(opts) != k:
            raise ValueError("one options dict per structure")
        total = sum(b << n for n, b in self.sizes)
        self.large = [(b << n) >= large_frac * total for n, b in self.sizes]

        def build(olist: List[Dict[str, int]]) -> List[Program]:
            progs = []
            for (n, specs, P), o in zip(structures, olist):
                prog = Program(n, specs, P, **{"fuse": 1, "jit": 1, **o})
                prog.compile(dtype, 0)
                progs.append(prog)
            return progs

        variants = {"default": build(opts)}
        self.tuning_ms: Dict[str, Dict[str, float]] = {}
        tuned = [dict(o) for o in opts]
        if tune:
            for i, ((n, specs, P), b) in enumerate(zip(structures, batches)):
                if self.large[i] and not opts[i]:
                    o, t = autotune_program(n, specs, P, b, dtype, jit=1)
                    self.tuning_ms["%dq_x%d" % (n, b)] = t
                    tuned[i] = o
        self.tuned_options = tuned
        if tuned != opts:
            variants["tuned"] = build(tuned)
        self._side = torch.cuda.Stream()
        steps: Dict[str, Any] = {}
        mixed = any(self.large) and not all(self.large)
        for name, progs in variants.items():
            steps["sequence_" + name] = self._sequence(progs)
            steps["multi_" + name] = self._hybrid(progs, [False] * k, False)
            if mixed:
                steps["hybrid_" + name] = self._hybrid(progs, self.large, False)
                steps["hybrid_streams_" + name] = self._hybrid(progs, self.large, True)
        graphs = {name: CapturedStep(fn) for name, fn in steps.items()}
        self.choice, self.candidate_ms = tune_step({name: g.replay for name, g in graphs.items()},
                                                   "sequence_default", reps, margin)
        # keep only the chosen step (the others' graphs hold their own memory)
        self.step = steps[self.choice]
        self.graph = graphs[self.choice]
        self.programs = variants["tuned" if self.choice.endswith("tuned") else "default"]
        self.multi_plan = getattr(self.step, "plan", None)

    def _energy(self, i: int) -> Optional[torch.Tensor]:
        h = self.hams[i]
        return None if h is None else h.expect(self.states[i])

    def _sequence(self, progs: List[Program]) -> Any:
        def step():
            for prog, pa, st in zip(progs, self.params, self.states):
                prog.forward(pa, pa.shape[0], self.dtype, out=st)
            return [self._energy(i) for i in range(len(progs))]
        return step

    def _hybrid(self, progs: List[Program], alone: List[bool], streams: bool) -> Any:
        """programs with alone[i] on their own passes, the rest in one
        multi-program launch per pass (on a second stream if ``streams``)"""
        rest = [i for i, a in enumerate(alone) if not a]
        plan = (MultiPlan([progs[i] for i in rest], self.dtype, batches=[self.sizes[i][1] for i in rest])
                if rest else None)
        items = [(progs[i], self.params[i], self.states[i]) for i in rest]

        def run_rest(res):
            if plan is not None:
                plan.forward(items)
            for i in rest:
                res[i] = self._energy(i)

        def step():
            res: List[Optional[torch.Tensor]] = [None] * len(progs)
            cur = torch.cuda.current_stream()
            if streams:
                self._side.wait_stream(cur)
                with torch.cuda.stream(self._side):
                    run_rest(res)
            else:
                run_rest(res)
            for i, a in enumerate(alone):
                if a:
                    progs[i].forward(self.params[i], self.params[i].shape[0], self.dtype, out=self.states[i])
                    res[i] = self._energy(i)
            if streams:
                cur.wait_stream(self._side)
            return res
        step.plan = plan
        return step

    def run(self) -> List[Optional[torch.Tensor]]:
        """Replay the chosen step (one graph launch)."""
        return self.graph.replay()

    def run_eager(self) -> List[Optional[torch.Tensor]]:
        """The chosen step without the graph (same kernels, launched one by one)."""
        return self.step()
