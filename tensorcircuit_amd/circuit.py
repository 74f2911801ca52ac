"""``Circuit``: the reference's statevector Circuit API (circuit.py,
basecircuit.py, abstractcircuit.py) on the HIP engine.

Gate calls record the QIR exactly as the reference does
(abstractcircuit.py:111-167: {"gatef", "index", "name", "split", "mpo",
"parameters"}) and lower to engine gates; ``wavefunction`` / ``expectation``
run the cached program on the device.  Gate methods are registered by
reflection over the same name lists (abstractcircuit.py:217-326), including
the upper-case and alias names and list/range fancy indexing.
"""
from __future__ import annotations

from typing import Any, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib, cons, engine
from . import gates as G
from .engine import GateSpec

Tensor = Any


def _is_sequence(x: Any) -> bool:
    return isinstance(x, (list, tuple, range)) or (isinstance(x, np.ndarray) and x.ndim > 0) or (
        isinstance(x, torch.Tensor) and x.dim() > 0)


class _Op:
    __slots__ = ("kind", "qubits", "theta", "scale", "mat")

    def __init__(self, kind, qubits, theta=None, scale=1.0, mat=None):
        self.kind, self.qubits, self.theta, self.scale, self.mat = kind, qubits, theta, scale, mat


def _float_of(x: Any, what: str) -> float:
    """A host float of a gate parameter the engine treats as a constant; a
    tensor that requires a gradient (or is traced by torch.func) raises
    instead of being silently detached."""
    if isinstance(x, torch.Tensor):
        if x.requires_grad:
            raise NotImplementedError(
                "%s: this gate is evaluated on the host as a constant, so its parameters cannot carry a "
                "gradient" % what)
        try:
            return float(x.detach().cpu().real.item())
        except Exception as e:  # functorch-wrapped tensors cannot be read
            raise NotImplementedError(
                "%s: this gate is evaluated on the host and its parameters cannot be traced tensors" % what) from e
    return float(np.real(x))


# fused matrix-core programs of constant circuits, by structure
_MFMA_CACHE: Dict[Any, Any] = {}


def _any_tensor(vars: Dict[str, Any]) -> bool:
    return any(isinstance(v, torch.Tensor) for v in vars.values())


class Circuit:
    """Statevector circuit on ``nqubits`` qubits (circuit.py:43-122).

    ``inputs`` may hold 2^n or 2^(2n) amplitudes (circuit.py:86-96: the
    latter lets a matrix be fed as a state, gates acting on the first n legs).
    """

    is_dm = False
    sgates = G.sgates
    vgates = G.vgates
    gate_aliases = G.gate_aliases

    def __init__(self, nqubits: int, inputs: Optional[Tensor] = None, mps_inputs: Any = None,
                 split: Optional[Dict[str, Any]] = None) -> None:
        if mps_inputs is not None:
            raise NotImplementedError("mps_inputs: MPS simulation is outside this engine")
        self._nqubits = int(nqubits)
        self._qir: List[Dict[str, Any]] = []
        self._extra_qir: List[Dict[str, Any]] = []
        self._ops: List[_Op] = []
        self.split = split
        self.circuit_param = {"nqubits": nqubits, "inputs": inputs, "mps_inputs": None, "split": split}
        self._nstate = self._nqubits
        self.inputs = None
        if inputs is not None:
            t = inputs if isinstance(inputs, torch.Tensor) else torch.as_tensor(np.asarray(inputs))
            t = t.reshape(-1)
            N = t.shape[0]
            m = int(round(np.log2(N)))
            assert m == nqubits or m == 2 * nqubits
            self._nstate = m
            self.inputs = t
        self.state_tensor = None

    # ------------------------------------------------------------ recording --
    def _norm_index(self, index: Sequence[int]) -> Tuple[int, ...]:
        idx = tuple(int(i) if int(i) >= 0 else self._nqubits + int(i) for i in index)
        assert len(idx) == len(set(idx))
        for i in idx:
            if i < 0 or i >= self._nqubits:
                raise ValueError("qubit index %d out of range for %d qubits" % (i, self._nqubits))
        return idx

    def _record(self, op: Union[_Op, List[_Op]], ir: Dict[str, Any]) -> None:
        """Append one QIR entry and the engine ops it lowers to (one op, or
        the factor list of a decomposed gate)."""
        self._ops.append(op if isinstance(op, list) else [op])
        self._qir.append(ir)
        self.state_tensor = None  # refresh the state cache (basecircuit.py:245)

    def apply_general_gate(self, gate: Any, *index: int, name: Optional[str] = None,
                           split: Any = None, mpo: bool = False, ir_dict: Optional[Dict] = None) -> None:
        """Apply an explicit gate matrix / Gate on ``index`` (basecircuit.py:120-245)."""
        if mpo:
            raise NotImplementedError("MPO gates are outside the dense statevector engine")
        idx = self._norm_index(index)
        m = G.matrix_of(gate)
        if m.shape[0] != 2 ** len(idx):
            raise ValueError("gate of dimension %d on %d qubits" % (m.shape[0], len(idx)))
        ir = ir_dict or {"gate": gate, "index": idx, "name": name or "any", "split": split, "mpo": mpo}
        self._record(_Op(_lib.G_MAT, idx, mat=m), ir)

    apply = apply_general_gate

    def _apply_named(self, gname: str, index: Tuple[int, ...], vars: Dict[str, Any], localname: str,
                     split: Any) -> None:
        idx = self._norm_index(index)
        ir = {"gatef": getattr(G, gname), "index": idx, "target": list(idx), "name": localname,
              "split": split, "mpo": False, "parameters": dict(vars)}
        if gname in G.CONST:
            self._record(_Op(_lib.G_MAT, idx, mat=np.asarray(G.CONST[gname], dtype=np.complex128)), ir)
            return
        if gname in G.TRAINABLE:
            kind, scale = G.TRAINABLE[gname]
            self._record(_Op(kind, idx, theta=vars.get("theta", 0.0), scale=scale), ir)
            return
        if gname in ("exp1", "exp"):
            u = G.matrix_of(vars["unitary"])
            theta = vars.get("theta", 0.0)
            half = bool(vars.get("half", False)) if gname == "exp1" else False
            if G.is_involution(u) and len(idx) <= 2:
                # exp1(U, t) = exp(-i t U) = cos t I - i sin t U for U^2 = I
                self._record(_Op(_lib.G_EXP1, idx, theta=theta, scale=0.5 if half else 1.0, mat=u), ir)
            elif gname == "exp" and G.is_generator(u) and len(idx) <= 2:
                self._record(_Op(_lib.G_EXPG, idx, theta=theta, scale=1.0, mat=u), ir)
            else:
                th = _float_of(theta, gname)
                m = G.exp1_matrix(u, th, half) if gname == "exp1" else G.exp_matrix(u, th)
                self._record(_Op(_lib.G_MAT, idx, mat=m), ir)
            return
        if gname in G.GENERATOR and _any_tensor(vars):
            g, scale = G.GENERATOR[gname]
            theta = vars.get("theta", G.PARAM_DEFAULT.get(gname, 0.0))
            self._record(_Op(_lib.G_EXPG, idx, theta=theta, scale=scale, mat=g), ir)
            return
        if gname in G.DECOMPOSE and _any_tensor(vars):
            ops = []
            for kind, pos, pname, scale, g in G.DECOMPOSE[gname]:
                ops.append(_Op(kind, tuple(idx[i] for i in pos), theta=vars.get(pname, 0.0), scale=scale, mat=g))
            self._record(ops, ir)
            return
        if gname in ("any", "unitary"):
            m = G.matrix_of(vars["unitary"])
            self._record(_Op(_lib.G_MAT, idx, mat=m), ir)
            return
        if gname in G.HOST_PARAM:
            f, keys = G.HOST_PARAM[gname]
            kw = {k: _float_of(vars[k], gname) for k in keys if k in vars}
            self._record(_Op(_lib.G_MAT, idx, mat=np.asarray(f(**kw), dtype=np.complex128)), ir)
            return
        raise ValueError("unknown gate %s" % gname)

    @classmethod
    def _make_sgate(cls, gname: str):
        def apply(self, *index, split=None, name=None):
            if isinstance(index[0], (int, np.integer)):
                self._apply_named(gname, index, {}, name or gname, split)
            elif _is_sequence(index[0]):
                for ind in zip(*index):
                    self._apply_named(gname, tuple(int(i) for i in ind), {}, name or gname, split)
            else:
                raise ValueError("Illegal index specification")

        apply.__doc__ = "Apply **%s** gate (reference gates.py)." % gname.upper()
        return apply

    @classmethod
    def _make_vgate(cls, gname: str):
        def apply(self, *index, **vars):
            split = vars.pop("split", None)
            localname = vars.pop("name", gname)
            if isinstance(index[0], (int, np.integer)):
                self._apply_named(gname, index, vars, localname, split)
            elif _is_sequence(index[0]):
                for i, ind in enumerate(zip(*index)):
                    nvars = {}
                    for k, v in vars.items():
                        try:
                            nvars[k] = v[i]
                        except Exception:  # scalar parameter shared by all
                            nvars[k] = v
                    self._apply_named(gname, tuple(int(j) for j in ind), nvars, localname, split)
            else:
                raise ValueError("Illegal index specification")

        apply.__doc__ = "Apply **%s** gate with parameters (reference gates.py)." % gname.upper()
        return apply

    @classmethod
    def _meta_apply(cls) -> None:
        """abstractcircuit.py:217-326."""
        for g in G.sgates:
            setattr(cls, g, cls._make_sgate(g))
            setattr(cls, g.upper(), cls._make_sgate(g))
        for g in G.vgates:
            setattr(cls, g, cls._make_vgate(g))
            setattr(cls, g.upper(), cls._make_vgate(g))
        for alias in G.gate_aliases:
            for a in alias[1:]:
                setattr(cls, a, getattr(cls, alias[0]))
                setattr(cls, a.upper(), getattr(cls, alias[0]))

    @staticmethod
    def apply_general_variable_gate_delayed(gatef: Any, name: Optional[str] = None, mpo: bool = False):
        """Register a user gate factory as a method (tests/test_gates.py:128-137);
        the gate matrix is built on the host from the call's parameters."""
        if name is None:
            name = getattr(gatef, "n", "any")

        def apply(self, *index, **vars):
            localname = vars.pop("name", name)
            split = vars.pop("split", None)
            if isinstance(gatef, G._GateFactory):
                # one of the library's own gates: lowered (and differentiated) as such
                self._apply_named(gatef.n, tuple(index), vars, localname, split)
                return
            mat = gatef(**{k: (_float_of(v, name) if k != "unitary" else v) for k, v in vars.items()})
            idx = self._norm_index(index)
            ir = {"gatef": gatef, "index": idx, "target": list(idx), "name": localname, "split": None,
                  "mpo": mpo, "parameters": vars}
            self._record(_Op(_lib.G_MAT, idx, mat=G.matrix_of(mat)), ir)

        return apply

    def barrier_instruction(self, *index: Any) -> None:
        """Layer marker; the adjoint engine may checkpoint the state here."""
        self._record(_Op(_lib.G_BARRIER, ()), {"name": "barrier", "index": (), "gatef": None})

    # ------------------------------------------------------------- lowering --
    def _lower(self) -> Tuple[List[GateSpec], List[Any]]:
        specs: List[GateSpec] = []
        thetas: List[Any] = []
        for op in (o for entry in self._ops for o in entry):
            if op.kind == _lib.G_BARRIER:
                specs.append(engine.barrier())
            elif op.kind == _lib.G_MAT:
                specs.append(GateSpec(_lib.G_MAT, op.qubits, -1, 0.0, 1.0, op.mat))
            else:
                specs.append(GateSpec(op.kind, op.qubits, len(thetas), 0.0, op.scale, op.mat))
                thetas.append(op.theta)
        return specs, thetas

    def _params(self, thetas: List[Any], dev: torch.device) -> torch.Tensor:
        rdt = getattr(torch, cons.rdtypestr)
        if not thetas:
            return torch.zeros(0, dtype=rdt, device=dev)
        if all(not isinstance(t, torch.Tensor) for t in thetas):
            return torch.tensor([float(np.real(t)) for t in thetas], dtype=rdt, device=dev)
        ts = []
        for t in thetas:
            if not isinstance(t, torch.Tensor):
                t = torch.tensor(float(np.real(t)), dtype=rdt)
            if t.is_complex():
                t = torch.real(t)
            ts.append(t.reshape(()).to(device=dev, dtype=rdt))
        return torch.stack(ts)

    def _program(self, specs: List[GateSpec], nparams: int) -> engine.Program:
        return engine.get_program(self._nstate, specs, nparams, fuse=1)

    def _constant_specs(self) -> Optional[List[GateSpec]]:
        """Every gate as a constant matrix when no parameter is a tensor
        (nothing to differentiate), else None."""
        specs: List[GateSpec] = []
        for op in (o for entry in self._ops for o in entry):
            if op.kind == _lib.G_BARRIER:
                continue
            if op.kind == _lib.G_MAT:
                specs.append(GateSpec(_lib.G_MAT, op.qubits, -1, 0.0, 1.0, op.mat))
                continue
            if isinstance(op.theta, torch.Tensor):
                return None
            a = op.scale * float(np.real(op.theta))
            if op.kind == _lib.G_EXP1:
                m = G.exp1_matrix(op.mat, a)
            elif op.kind in (_lib.G_RXX, _lib.G_RYY, _lib.G_RZZ):
                gen = {_lib.G_RXX: G._xx_matrix, _lib.G_RYY: G._yy_matrix, _lib.G_RZZ: G._zz_matrix}[op.kind]
                m = G.exp1_matrix(gen, a, half=True)
            else:
                m = G.factor_matrix(op.kind, op.mat, a)
            specs.append(GateSpec(_lib.G_MAT, op.qubits, -1, 0.0, 1.0, np.asarray(m, dtype=np.complex128)))
        return specs

    def _mfma_program(self) -> Optional[engine.Program]:
        """A constant circuit on >= 20 qubits in complex64: its gates fused
        into dense 5-qubit blocks applied on the FP32 matrix cores (LDS tile
        passes of consecutive blocks), or None when the circuit does not
        qualify (tensor parameters, fewer qubits, complex128, or fusion
        would not reduce the number of state passes)."""
        if not cons.mfma_for(self._nstate) or self._nstate < 5:
            return None
        specs = self._constant_specs()
        if not specs:
            return None
        from . import fusion

        key = ("mfma", self._nstate, tuple(s.key() for s in specs))
        prog = _MFMA_CACHE.get(key)
        if prog is None:
            blocks = fusion.fuse(specs, self._nstate)
            if 4 * len(blocks) > len(specs):
                return None  # too few gates per block to pay for dense 32 x 32 blocks
            prog = engine.Program(self._nstate, blocks, 0, fuse=1, mfma=2)
            _MFMA_CACHE[key] = prog
            while len(_MFMA_CACHE) > 16:
                _MFMA_CACHE.pop(next(iter(_MFMA_CACHE)))
        return prog

    def _run(self) -> torch.Tensor:
        dev = engine.device()
        mp = self._mfma_program()
        if mp is not None:
            inp = None
            if self.inputs is not None:
                inp = self.inputs.to(device=dev).to(engine._CDT[cons.dtypestr]).reshape(1, -1)
            st, _ = mp.forward(None, 1, cons.dtypestr, inputs=inp)
            return st.reshape(-1)
        specs, thetas = self._lower()
        params = self._params(thetas, dev)
        prog = self._program(specs, len(thetas))
        inp = None
        if self.inputs is not None:
            inp = self.inputs.to(device=dev)
            inp = inp.to(engine._CDT[cons.dtypestr])
        st, _ = engine.StateFn.apply(params, inp, prog, cons.dtypestr)
        return st

    def _state(self, reuse: bool = True) -> torch.Tensor:
        if reuse and self.state_tensor is not None:
            return self.state_tensor
        st = self._run()
        if reuse:
            self.state_tensor = st
        return st

    # -------------------------------------------------------------- outputs --
    def wavefunction(self, form: str = "default") -> Tensor:
        """circuit.py:792-810: "default" [-1], "ket" [-1, 1], "bra" [1, -1] (no conj)."""
        s = self._state()
        if form == "default":
            return s.reshape(-1)
        if form == "ket":
            return s.reshape(-1, 1)
        if form == "bra":
            return s.reshape(1, -1)
        raise ValueError("unknown form %r" % form)

    state = wavefunction

    def get_state(self) -> Tensor:
        return self.wavefunction()

    def amplitude(self, l: Union[str, Tensor]) -> Tensor:
        """<l|psi> for a bitstring (basecircuit.py:447-508)."""
        s = self._state()
        if isinstance(l, str):
            idx = int(l, 2)
            return s[idx]
        bits = torch.as_tensor(l).to(s.device)
        weights = 2 ** torch.arange(self._nstate - 1, -1, -1, device=s.device)
        return s[(bits.to(torch.int64) * weights).sum()]

    def probability(self) -> Tensor:
        """|psi|^2 over the computational basis (basecircuit.py:510-524),
        differentiable."""
        return engine.ProbFn.apply(self._state(), self._nstate)

    # ------------------------------------------------------- measurement --
    def measure_jit(self, *index: int, with_prob: bool = False, status: Optional[Tensor] = None
                    ) -> Tuple[Tensor, Tensor]:
        """Measure the given qubits one after another in the Z basis without
        collapsing the circuit (basecircuit.py:371-445): outcome k is 1 iff
        status[k] > P(0 | earlier outcomes) - 0.31415926e-12.  Returns the
        outcomes (float 0/1) and, with ``with_prob``, the probability of the
        record (else -1.0)."""
        idx = [int(i) if int(i) >= 0 else self._nqubits + int(i) for i in index]
        if status is None:
            u = torch.rand(len(idx), dtype=torch.float64)
        else:
            u = torch.as_tensor(status, dtype=torch.float64).reshape(-1)[: len(idx)]
        s = self._state()
        outc, prob = engine.measure(s.detach(), self._nstate, idx, u.reshape(1, -1))
        rdt = s.real.dtype
        sample = outc[0].to(rdt)
        if with_prob:
            return sample, prob[0].to(rdt)
        return sample, -1.0

    measure = measure_jit

    def perfect_sampling(self, status: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
        """All qubits measured in order, with the record's probability
        (basecircuit.py:359-369)."""
        return self.measure_jit(*range(self._nqubits), with_prob=True, status=status)

    def readouterror_bs(self, readout_error: Optional[Sequence[Any]] = None, p: Optional[Tensor] = None) -> Tensor:
        """Noisy bitstring probabilities: qubit i's readout matrix
        [[e_i0, 1 - e_i1], [1 - e_i0, e_i1]] applied to p (basecircuit.py:760-803)."""
        n = len(readout_error)  # type: ignore
        q = torch.as_tensor(p)
        q = (torch.real(q) if q.is_complex() else q).to(torch.float64).reshape(-1)
        for i in range(n):
            e0, e1 = float(readout_error[i][0]), float(readout_error[i][1])  # type: ignore
            m = torch.tensor([[e0, 1 - e1], [1 - e0, e1]], dtype=torch.float64, device=q.device)
            v = q.reshape(2**i, 2, -1)
            q = torch.einsum("ab,ibj->iaj", m, v).reshape(-1)
        return q

    def sample(self, batch: Optional[int] = None, allow_state: bool = False,
               readout_error: Optional[Sequence[Any]] = None, format: Optional[str] = None,
               random_generator: Any = None, status: Optional[Tensor] = None, **kws: Any) -> Any:
        """Bitstring samples of the output state (basecircuit.py:526-616).
        Every draw comes from |psi|^2 on the device: with ``allow_state``
        (and for ``batch`` draws without it) by inverse CDF over the prefix
        sums (the reference's probability_sample), else one sequential
        qubit-by-qubit draw (perfect_sampling); the reference's perfect
        sampling samples the same distribution through a different map from
        random numbers to outcomes.  ``format`` as in
        quantum.measurement_counts; ``status`` [batch] uniforms replace the
        generator."""
        from . import quantum as qu

        if format is None and "format_" in kws:
            format = kws.pop("format_")
        n = self._nqubits
        if not allow_state and batch is None and readout_error is None:
            u = qu._uniforms((n,), None if status is None else status, random_generator)
            r = self.perfect_sampling(u)
            if format is None:
                return r
            ch = qu.sample_bin2int(r[0].to(torch.int64)[None, :], n)
            return qu.sample2all(sample=ch, n=n, format=format, jittable=True)
        nb = 1 if batch is None else int(batch)
        p = self.probability().detach()
        if readout_error is not None:
            p = self.readouterror_bs(readout_error, p).to(p.dtype)
        u = qu._uniforms((1, nb), status, random_generator)
        ch = engine.sample_prob_indices(p.reshape(1, -1), n, u)[0]
        if format is None:
            confg = qu.sample_int2bin(ch, n)
            prob = (p / torch.sum(p))[ch]
            r = list(zip(confg, prob))
            return r[0] if batch is None else r
        return qu.sample2all(sample=ch, n=n, format=format, jittable=True)

    def sample_expectation_ps(self, x: Optional[Sequence[int]] = None, y: Optional[Sequence[int]] = None,
                              z: Optional[Sequence[int]] = None, shots: Optional[int] = None,
                              random_generator: Any = None, status: Optional[Tensor] = None,
                              readout_error: Optional[Sequence[Any]] = None, noise_conf: Any = None,
                              nmc: int = 1000, statusc: Any = None, **kws: Any) -> Tensor:
        """Pauli-string expectation from Z-basis measurement after basis
        rotation (H for x, rx(pi/2) for y; basecircuit.py:618-758): exact
        from the probabilities when ``shots`` is None, else the mean parity
        of ``shots`` device draws."""
        from . import quantum as qu

        if noise_conf is not None:
            raise NotImplementedError("noisy sampling is outside this engine")
        x, y, z = list(x or []), list(y or []), list(z or [])
        c = type(self)(self._nqubits, inputs=self._state())
        for i in x:
            c.h(i)
        for i in y:
            c.rx(i, theta=np.pi / 2)
        p = c.probability()
        if readout_error is not None:
            p = self.readouterror_bs(readout_error, p).to(p.dtype)
        idx = [int(i) if int(i) >= 0 else self._nqubits + int(i) for i in x + y + z]
        if shots is None:
            return qu.correlation_from_counts(idx, p)
        u = qu._uniforms((1, int(shots)), status, random_generator)
        ch = engine.sample_prob_indices(p.detach().reshape(1, -1), self._nqubits, u)[0]
        return qu.correlation_from_samples(idx, ch, self._nqubits)

    sexpps = sample_expectation_ps

    def expectation_ps(self, x: Optional[Sequence[int]] = None, y: Optional[Sequence[int]] = None,
                       z: Optional[Sequence[int]] = None, ps: Optional[Sequence[int]] = None,
                       reuse: bool = True, noise_conf: Any = None, nmc: int = 1000,
                       status: Any = None, **kws: Any) -> Tensor:
        """Pauli-string expectation (abstractcircuit.py:1208-1288); ``ps``
        (codes 0=I,1=X,2=Y,3=Z) overrides x/y/z."""
        if noise_conf is not None:
            raise NotImplementedError("noisy expectations are outside this engine")
        if ps is not None:
            from .quantum import ps2xyz

            d = ps2xyz(ps)
            x, y, z = d.get("x"), d.get("y"), d.get("z")
        codes = np.zeros(self._nstate, dtype=np.int8)
        seen = set()
        for lst, code in ((x, 1), (y, 2), (z, 3)):
            for i in lst or []:
                i = int(i)
                i = i if i >= 0 else self._nqubits + i
                if i in seen:
                    raise ValueError("Cannot measure two operators in one index")
                seen.add(i)
                codes[i] = code
        return self._pauli_expectation(codes[None, :], np.ones(1), reuse)

    def _pauli_expectation(self, codes: np.ndarray, coeffs: np.ndarray, reuse: bool = True) -> Tensor:
        ham = engine.get_pauli(self._nstate, codes, coeffs)
        return engine.PauliExpFn.apply(self._state(reuse), ham)

    def expectation(self, *ops: Tuple[Any, List[int]], reuse: bool = True, enable_lightcone: bool = False,
                    noise_conf: Any = None, nmc: int = 1000, status: Any = None, **kws: Any) -> Tensor:
        """<psi|O_1...O_m|psi>, unnormalised, complex (circuit.py:914-990)."""
        if noise_conf is not None:
            raise NotImplementedError("noisy expectations are outside this engine")
        occupied = set()
        norm_ops = []
        for op, index in ops:
            if isinstance(index, (int, np.integer)):
                index = [index]
            idx = tuple(int(i) if int(i) >= 0 else self._nqubits + int(i) for i in index)
            for i in idx:
                if i in occupied:
                    raise ValueError("Cannot measure two operators in one index")
                occupied.add(i)
            norm_ops.append((G.matrix_of(op), idx))
        # Pauli strings take the fused fp64 kernel
        codes = np.zeros(self._nstate, dtype=np.int8)
        all_pauli = True
        for m, idx in norm_ops:
            c = G.pauli_code(m) if len(idx) == 1 else None
            if c is None:
                all_pauli = False
                break
            codes[idx[0]] = c
        if all_pauli:
            return self._pauli_expectation(codes[None, :], np.ones(1), reuse)
        specs = [GateSpec(_lib.G_MAT, idx, -1, 0.0, 1.0, m) for m, idx in norm_ops]
        dspecs = [GateSpec(_lib.G_MAT, idx, -1, 0.0, 1.0, m.conj().T) for m, idx in norm_ops]
        prog = engine.get_program(self._nstate, specs, 0, fuse=1)
        prog_dag = engine.get_program(self._nstate, dspecs, 0, fuse=1)
        return engine.OpExpFn.apply(self._state(reuse), prog, prog_dag, cons.dtypestr)

    # ------------------------------------------------------------ QIR / misc --
    def to_qir(self) -> List[Dict[str, Any]]:
        return self._qir

    @classmethod
    def from_qir(cls, qir: List[Dict[str, Any]], circuit_params: Optional[Dict[str, Any]] = None) -> "Circuit":
        """Rebuild a circuit from ``to_qir()`` output (abstractcircuit.py:367-426);
        without ``nqubits`` in ``circuit_params`` the width is 1 + the largest
        index used."""
        params = dict(circuit_params or {})
        if "nqubits" not in params:
            params["nqubits"] = 1 + max([max(d["index"]) for d in qir if len(d.get("index", ()))] or [0])
        c = cls(**params)
        for d in qir:
            gatef = d.get("gatef")
            index = tuple(d.get("index", ()))
            if gatef is None and d.get("name") == "barrier":
                c.barrier_instruction()
            elif isinstance(gatef, G._GateFactory):
                c._apply_named(gatef.n, index, dict(d.get("parameters", {})), d.get("name", gatef.n),
                               d.get("split"))
            elif gatef is None:
                c.apply_general_gate(d["gate"], *index, name=d.get("name"), split=d.get("split"))
            else:
                c.apply_general_variable_gate_delayed(gatef, d.get("name"), mpo=d.get("mpo", False))(
                    c, *index, **dict(d.get("parameters", {})))
        return c

    def to_json(self, file: Optional[str] = None, simplified: bool = False) -> Any:
        """JSON gate list of the circuit (abstractcircuit.py:932-950)."""
        import json

        from .translation import qir2json

        tcqasm = qir2json(self.to_qir(), simplified=simplified)
        if file is not None:
            with open(file, "w") as f:
                json.dump(tcqasm, f)
        return json.dumps(tcqasm)

    @classmethod
    def from_json(cls, jsonstr: Any, circuit_params: Optional[Dict[str, Any]] = None) -> "Circuit":
        """abstractcircuit.py:1036-1056."""
        import json

        from .translation import json2qir

        if isinstance(jsonstr, str):
            jsonstr = json.loads(jsonstr)
        return cls.from_qir(json2qir(jsonstr), circuit_params)

    @classmethod
    def from_json_file(cls, file: str, circuit_params: Optional[Dict[str, Any]] = None) -> "Circuit":
        """abstractcircuit.py:1058-1074."""
        import json

        with open(file, "r") as f:
            return cls.from_json(json.load(f), circuit_params)

    @classmethod
    def from_qsim_file(cls, file: str, circuit_params: Optional[Dict[str, Any]] = None) -> "Circuit":
        """A circuit in qsim's text format: qubit count, then one gate per
        line (abstractcircuit.py:952-966)."""
        with open(file, "r") as f:
            lines = f.readlines()
        params = dict(circuit_params or {})
        if "nqubits" not in params:
            params["nqubits"] = int(lines[0])
        return cls._apply_qsim(cls(**params), lines)

    @staticmethod
    def _apply_qsim(c: "Circuit", qsim_str: List[str]) -> "Circuit":
        from .translation import apply_qsim

        return apply_qsim(c, qsim_str[1:])

    def gate_count(self, gate_list: Optional[Union[str, Sequence[str]]] = None) -> int:
        """abstractcircuit.py gate_count."""
        if gate_list is None:
            return len([d for d in self._qir if d.get("name") != "barrier"])
        if isinstance(gate_list, str):
            gate_list = [gate_list]
        names = set()
        for g in gate_list:
            names.add(g.lower())
            for alias in G.gate_aliases:
                if g.lower() in alias:
                    names.update(alias)
        return len([d for d in self._qir if str(d.get("name", "")).lower() in names])

    def inverse(self, circuit_params: Optional[Dict[str, Any]] = None) -> "Circuit":
        """U^dagger as a new circuit of the same class, with the reference's
        QIR rewriting (abstractcircuit.py:428-507): gates run in reverse
        order; sd/td become s/t; r, cr and the other rotations take
        theta -> -theta (iswap's default theta is 1); u and cu take
        (phi, lbd) -> (pi - lbd, pi - phi); any / exp / exp1, wroot and
        user gates become their adjoint (QIR name "<gate>d").  Unlike the
        reference, which re-applies s and t unchanged, s and t become sd and
        td, so that inverse() always inverts."""
        params = dict(circuit_params or {})
        params.setdefault("nqubits", self._nqubits)
        c = type(self)(**params)
        for ops, d in zip(reversed(self._ops), reversed(self._qir)):
            gatef = d.get("gatef")
            index = tuple(d.get("index", ()))
            name, split = d.get("name"), d.get("split")
            if gatef is None and name == "barrier":
                c.barrier_instruction()
                continue
            gname = gatef.n if isinstance(gatef, G._GateFactory) else None
            if gname in G.sgates and gname not in ("wroot",):
                inv = {"sd": "s", "td": "t", "s": "sd", "t": "td"}.get(gname, gname)
                c._apply_named(inv, index, {}, name, split)
            elif gname in ("r", "cr") or (gname in G.vgates and gname not in ("u", "cu", "any", "exp", "exp1")):
                vs = dict(d.get("parameters", {}))
                vs["theta"] = -vs.get("theta", G.PARAM_DEFAULT.get(gname, 0.0))
                c._apply_named(gname, index, vs, name, split)
            elif gname in ("u", "cu"):
                vs = dict(d.get("parameters", {}))
                phi, lbd = vs.get("phi", 0.0), vs.get("lbd", 0.0)
                vs["lbd"] = -phi + np.pi
                vs["phi"] = -lbd + np.pi
                c._apply_named(gname, index, vs, name, split)
            else:
                # adjoint of the gate as recorded: the inverse of each
                # lowered op in reverse order
                inv_ops = []
                for op in reversed(ops):
                    if op.kind == _lib.G_MAT:
                        inv_ops.append(_Op(_lib.G_MAT, op.qubits, mat=np.asarray(op.mat).conj().T))
                    else:
                        inv_ops.append(_Op(op.kind, op.qubits, theta=-op.theta, scale=op.scale, mat=op.mat))
                adj = G.adjoint_factory(gatef) if gatef is not None else None
                ir = {"gatef": adj, "index": index, "target": list(index), "name": name, "split": split,
                      "mpo": d.get("mpo", False)}
                if "parameters" in d:
                    ir["parameters"] = dict(d["parameters"])
                if gatef is None and d.get("gate") is not None:
                    ir["gate"] = G.Gate(G.matrix_of(d["gate"]).conj().T, name or "any")
                c._record(inv_ops, ir)
        return c

    def _entries(self, indices: Optional[Sequence[int]] = None):
        """(ops, QIR entry) pairs of this circuit, qubits mapped through
        ``indices`` (qubit i -> indices[i])."""
        for ops, d in zip(self._ops, self._qir):
            d = dict(d)
            if indices is None:
                yield list(ops), d
                continue
            m = [int(i) for i in indices]
            if "index" in d:
                d["index"] = tuple(m[i] for i in d["index"])
                d["target"] = list(d["index"])
            yield [_Op(o.kind, tuple(m[q] for q in o.qubits), theta=o.theta, scale=o.scale, mat=o.mat)
                   for o in ops], d

    def append(self, c: "Circuit", indices: Optional[List[int]] = None) -> "Circuit":
        """Append the gates of ``c`` after this circuit's, qubit i of ``c``
        on ``indices[i]`` (abstractcircuit.py:1148-1190); trainable gates
        stay trainable."""
        for ops, d in list(c._entries(indices)):
            self._record(ops, d)
        return self

    def prepend(self, c: "Circuit") -> "Circuit":
        """Put the gates of ``c`` before this circuit's (abstractcircuit.py
        prepend)."""
        mine = list(self._entries())
        self._ops, self._qir = [], []
        for ops, d in list(c._entries()) + mine:
            self._record(ops, d)
        return self

    def copy(self) -> "Circuit":
        c = type(self)(**{k: v for k, v in self.circuit_param.items() if k != "mps_inputs"})
        return c.append(self)

    def matrix(self) -> Tensor:
        """The circuit unitary (circuit.py matrix): columns U|j> for every j."""
        n = self._nqubits
        dev = engine.device()
        specs, thetas = self._lower()
        params = self._params(thetas, dev)
        prog = engine.get_program(n, specs, len(thetas), fuse=1)
        eye = torch.eye(2**n, dtype=engine._CDT[cons.dtypestr], device=dev)
        p = params.reshape(1, -1).expand(2**n, -1).contiguous() if len(thetas) else None
        st, _ = prog.forward(p, 2**n, cons.dtypestr, inputs=eye)
        return st.T

    def __repr__(self) -> str:
        return "Circuit(nqubits=%d, gates=%d)" % (self._nqubits, len(self._qir))


Circuit._meta_apply()


def expectation(*ops: Tuple[Any, List[int]], ket: Tensor, bra: Optional[Tensor] = None,
                conj: bool = True, normalization: bool = False) -> Tensor:
    """Module-level <bra|O|ket> (circuit.py:997-1129)."""
    dev = engine.device()
    ket = torch.as_tensor(ket).to(device=dev).to(engine._CDT[cons.dtypestr]).reshape(-1)
    n = int(round(np.log2(ket.shape[0])))
    if bra is None:
        bra = ket
    else:
        bra = torch.as_tensor(bra).to(device=dev).to(ket.dtype).reshape(-1)
    specs = []
    for op, index in ops:
        if isinstance(index, (int, np.integer)):
            index = [index]
        idx = tuple(int(i) if int(i) >= 0 else n + int(i) for i in index)
        specs.append(GateSpec(_lib.G_MAT, idx, -1, 0.0, 1.0, G.matrix_of(op)))
    prog = engine.get_program(n, specs, 0, fuse=1)
    phi, _ = prog.forward(None, 1, cons.dtypestr, inputs=ket.reshape(1, -1))
    b = bra if conj else torch.conj(bra)
    r = engine.inner(b.reshape(1, -1), phi, n).to(ket.dtype).reshape(())
    if normalization:
        nb = engine.inner(bra.reshape(1, -1), bra.reshape(1, -1), n).real
        nk = engine.inner(ket.reshape(1, -1), ket.reshape(1, -1), n).real
        r = r / torch.sqrt(nb * nk).to(r.dtype).reshape(())
    return r
