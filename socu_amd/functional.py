"""Functional front-end of the ND block-tridiagonal Cholesky on torch tensors,
the counterpart of socu's JAX interface (socu/jax.py:270-337):

* ``cholesky_factor(L, E)``                 -> (L_factor, E_factor)
* ``cholesky_solve(L_factor, E_factor, b)`` -> x
* ``cholesky_factor_and_solve(L, E, b)``    -> (L_factor, E_factor, x)
* ``forward_substitution(L_factor, E_factor, b)``  -> y
* ``backward_substitution(L_factor, E_factor, y)`` -> x

Unlike the launch factories these never modify their arguments: each call
works on copies and returns new tensors.  As in the reference:

* shapes are validated with the reference's messages (jax.py:25-62);
* ``E`` may be given short of the ND storage length, e.g. as the N-1 coupling
  blocks, and is zero-padded to ``calculate_off_diag_storage_len(N)``
  (jax.py:72-82); the factors come back at full storage length;
* with ``pad_problem`` the block size is padded to the multiple
  ``optimal_problem_settings`` names, with zeros off the diagonal and an
  identity on the padded diagonal, so the real blocks are unchanged
  (jax.py:85-129); results are sliced back to n;
* a 3-D call is a batch of one (jax.py:65-69), and under ``torch.func.vmap``
  every mapped call becomes ONE batched launch, unmapped operands broadcast
  (the custom_vmap rule of jax.py:223-263).  Leading batch dimensions beyond
  one are also accepted and flattened into one launch.

Gradients (no counterpart in the reference, whose JAX calls are not
differentiable), through ``torch.autograd``:

* ``cholesky_factor_and_solve``: x = Psi^{-1} b with respect to L (lower
  triangle), E (the coupling blocks; the fill-in region gets zeros) and b,
  by one transpose solve with the factor and the Psi-pattern adjoint kernel;
  the returned factors carry no gradient.
* ``cholesky_solve`` / ``forward_substitution`` / ``backward_substitution``:
  with respect to the right-hand side (the transposed substitution) and to the
  stored factor entries (tril of each L block and every E block, the
  factor-pattern adjoint kernel).
* ``cholesky_factor``: its outputs carry no gradient.

All work runs in the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import torch
import torch.nn.functional as F

from . import _lib
from .block_tridiag_solver import calculate_off_diag_storage_len, optimal_problem_settings

__all__ = [
    "cholesky_factor",
    "cholesky_solve",
    "cholesky_factor_and_solve",
    "forward_substitution",
    "backward_substitution",
]


# ---------------------------------------------------------------------------
# validation and padding (jax.py:25-129)
# ---------------------------------------------------------------------------

def _validate_problem_shapes(L, E, b=None):
    """jax.py:25-62, with any number (>= 0) of leading batch dimensions."""
    if L.ndim < 3:
        raise ValueError(f"L must be 3- or 4-dimensional, got shape {tuple(L.shape)}")
    if E.ndim != L.ndim:
        raise ValueError(f"E must have the same rank as L, got L: {tuple(L.shape)}, E: {tuple(E.shape)}")
    horizon_axis = L.ndim - 3
    block_axis = L.ndim - 2
    N = L.shape[horizon_axis]
    n = L.shape[block_axis]
    if L.shape[-1] != n:
        raise ValueError(f"L must be square in last two dimensions, got shape {tuple(L.shape)}")
    if tuple(E.shape[horizon_axis + 1:]) != tuple(L.shape[block_axis:]):
        raise ValueError(f"E block dimensions must match L's n={n}, got E shape {tuple(E.shape)}")
    if L.ndim > 3 and E.shape[:horizon_axis] != L.shape[:horizon_axis]:
        raise ValueError(f"L and E must have the same batch size, got L: {tuple(L.shape)}, E: {tuple(E.shape)}")
    if b is not None:
        if b.ndim != L.ndim:
            raise ValueError(f"b must have the same rank as L, got L: {tuple(L.shape)}, b: {tuple(b.shape)}")
        if b.shape[horizon_axis] != N:
            raise ValueError(f"b horizon dimension must be {N}, got b shape {tuple(b.shape)}")
        if b.shape[block_axis] != n:
            raise ValueError(f"b block dimension must be {n}, got b shape {tuple(b.shape)}")
        if L.ndim > 3 and b.shape[:horizon_axis] != L.shape[:horizon_axis]:
            raise ValueError(f"L and b must have the same batch size, got L: {tuple(L.shape)}, b: {tuple(b.shape)}")
        if L.dtype != b.dtype:
            raise ValueError(f"L and b must have same dtype, got L: {L.dtype}, b: {b.dtype}")
    if L.dtype != E.dtype:
        raise ValueError(f"L and E must have same dtype, got L: {L.dtype}, E: {E.dtype}")
    for name, t in (("L", L), ("E", E), ("b", b)):
        if t is not None and t.device.type != "cuda":
            raise ValueError(f"{name} is on {t.device}: socu_amd runs on an AMD Instinct GPU and has no CPU path")
    if L.dtype not in (torch.float32, torch.float64):
        raise RuntimeError("dtype must be float32 or float64")


def _ensure_E_size(N, E):
    """Zero-pads E (B, k, n, n) to the ND storage length (jax.py:72-82)."""
    nE = calculate_off_diag_storage_len(N)
    if E.shape[1] < nE:
        return F.pad(E, (0, 0, 0, 0, 0, nE - E.shape[1]))
    return E


def _pad_factor_data(L, E, b, factor):
    """Pads the block size to the multiple of the settings table, identity on
    the padded diagonal (jax.py:85-129).  Differentiable torch ops."""
    n = L.shape[-1]
    pm = optimal_problem_settings(n, L.dtype)["pad_multiple"]["factor" if factor else "solve"]
    n_pad = (n + pm - 1) // pm * pm
    if n_pad == n:
        return L, E, b
    w = n_pad - n
    eye = torch.zeros(n_pad, n_pad, dtype=L.dtype, device=L.device)
    eye[n:, n:] = torch.eye(w, dtype=L.dtype, device=L.device)
    L = F.pad(L, (0, w, 0, w)) + eye
    E = F.pad(E, (0, w, 0, w))
    if b is not None:
        b = F.pad(b, (0, 0, 0, w))
    return L, E, b


# ---------------------------------------------------------------------------
# calls into the HIP library (4-D, padded, full storage length, contiguous)
# ---------------------------------------------------------------------------

def _code(t):
    return _lib.DTYPE_F64 if t.dtype == torch.float64 else _lib.DTYPE_F32


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _dims(L, b=None):
    B, N, n = L.shape[0], L.shape[1], L.shape[2]
    m = b.shape[3] if b is not None else 1
    return B, N, n, m


def _substitute(op, Lf, Ef, rhs):
    """New tensor: forward / backward substitution or solve of rhs with the
    factor (Lf, Ef); the inputs are only read."""
    lib = _lib.load()
    B, N, n, m = _dims(Lf, rhs)
    x = rhs.contiguous().clone()
    if B == 0:
        return x
    code = _code(Lf)
    Lc, Ec = Lf.contiguous(), Ef.contiguous()
    dev = x.device
    if op == "solve_own":
        # a factor made by this library's factor: its diagonal-block
        # inverses are in place, no workspace or inverse launch needed
        with torch.cuda.device(dev):
            _lib.check(lib.socu_cholesky_solve_with_inverses(Lc.data_ptr(), Ec.data_ptr(), x.data_ptr(), None, B, N,
                                                             n, m, code, _stream(dev)), "cholesky_solve")
        return x
    work = torch.empty(lib.socu_solve_workspace_size(B, N, n, code), dtype=torch.uint8, device=dev)
    fn = {"forward": lib.socu_cholesky_forward_substitution, "backward": lib.socu_cholesky_backward_substitution,
          "solve": lib.socu_cholesky_solve}[op]
    with torch.cuda.device(dev):
        _lib.check(fn(Lc.data_ptr(), Ec.data_ptr(), x.data_ptr(), None, B, N, n, m, code, work.data_ptr(),
                      _stream(dev)), f"cholesky_{op}")
    return x


def _factor_grads(Lf, Ef, u, v, p=None, q=None):
    """(dL, dE) = -(u v^T + p q^T) on the stored factor's pattern."""
    lib = _lib.load()
    B, N, n, m = _dims(Lf, u)
    dL = torch.empty_like(Lf, memory_format=torch.contiguous_format)
    dE = torch.empty_like(Ef, memory_format=torch.contiguous_format)
    if B == 0:
        return dL.zero_(), dE.zero_()
    u, v = u.contiguous(), v.contiguous()
    p = p.contiguous() if p is not None else None
    q = q.contiguous() if q is not None else None
    dev = Lf.device
    with torch.cuda.device(dev):
        _lib.check(lib.socu_factor_adjoint_scatter(u.data_ptr(), v.data_ptr(), None if p is None else p.data_ptr(),
                                                   None if q is None else q.data_ptr(), dL.data_ptr(), dE.data_ptr(),
                                                   None, B, N, n, m, _code(Lf), _stream(dev)),
                   "factor adjoint scatter")
    return dL, dE


def _flatten_vmapped(info, in_dims, *args):
    """Moves each mapped dimension to the front (expanding unmapped operands
    to the vmap size) and merges it with the batch axis: (V, B, ...) ->
    (V*B, ...).  Returns the merged tensors and V."""
    V = info.batch_size
    out = []
    for t, d in zip(args, in_dims):
        if d is None:
            t = t.unsqueeze(0).expand(V, *t.shape)
        else:
            t = t.movedim(d, 0)
        out.append(t.reshape(V * t.shape[1], *t.shape[2:]))
    return out, V


def _unflatten(V, *outs):
    return tuple(o.reshape(V, o.shape[0] // V, *o.shape[1:]) for o in outs)


class _Factor(torch.autograd.Function):
    @staticmethod
    def forward(L, E):
        lib = _lib.load()
        B, N, n, _ = _dims(L)
        # out of place: the factor is written to fresh buffers straight from
        # the inputs (no copies first)
        Ls, Es = L.contiguous(), E.contiguous()
        Lf, Ef = torch.empty_like(Ls), torch.empty_like(Es)
        if B > 0:
            dev = Lf.device
            with torch.cuda.device(dev):
                _lib.check(lib.socu_cholesky_factor_from(Ls.data_ptr(), Es.data_ptr(), Lf.data_ptr(), Ef.data_ptr(),
                                                         None, B, N, n, _code(Lf), _stream(dev)), "cholesky_factor")
        return Lf, Ef

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.mark_non_differentiable(*output)

    @staticmethod
    def backward(ctx, gL, gE):
        raise RuntimeError("cholesky_factor: the factor blocks carry no gradient")

    @staticmethod
    def vmap(info, in_dims, L, E):
        (L, E), V = _flatten_vmapped(info, in_dims, L, E)
        return _unflatten(V, *_Factor.apply(L, E)), (0, 0)


class _FactorAndSolve(torch.autograd.Function):
    @staticmethod
    def forward(L, E, b):
        lib = _lib.load()
        B, N, n, m = _dims(L, b)
        Ls, Es = L.contiguous(), E.contiguous()
        Lf, Ef = torch.empty_like(Ls), torch.empty_like(Es)
        x = b.contiguous().clone()
        if B > 0:
            dev = Lf.device
            with torch.cuda.device(dev):
                _lib.check(lib.socu_cholesky_factor_and_solve_from(Ls.data_ptr(), Es.data_ptr(), Lf.data_ptr(),
                                                                   Ef.data_ptr(), x.data_ptr(), None, B, N, n, m,
                                                                   _code(Lf), _stream(dev)),
                           "cholesky_factor_and_solve")
        return Lf, Ef, x

    @staticmethod
    def setup_context(ctx, inputs, output):
        Lf, Ef, x = output
        ctx.mark_non_differentiable(Lf, Ef)
        ctx.save_for_backward(Lf, Ef, x)

    @staticmethod
    def backward(ctx, gL, gE, gx):
        Lf, Ef, x = ctx.saved_tensors
        B, N, n, m = _dims(Lf, x)
        bbar = _substitute("solve_own", Lf, Ef, gx)
        dL = dE = None
        if (ctx.needs_input_grad[0] or ctx.needs_input_grad[1]) and B > 0:
            lib = _lib.load()
            dL = torch.empty_like(Lf)
            dE = torch.empty_like(Ef)
            dev = x.device
            with torch.cuda.device(dev):
                _lib.check(lib.socu_adjoint_scatter(bbar.data_ptr(), x.data_ptr(), dL.data_ptr(), dE.data_ptr(), None,
                                                    B, N, n, m, _code(Lf), _stream(dev)), "adjoint scatter")
        elif ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            dL, dE = torch.zeros_like(Lf), torch.zeros_like(Ef)
        return (dL if ctx.needs_input_grad[0] else None, dE if ctx.needs_input_grad[1] else None,
                bbar if ctx.needs_input_grad[2] else None)

    @staticmethod
    def vmap(info, in_dims, L, E, b):
        (L, E, b), V = _flatten_vmapped(info, in_dims, L, E, b)
        return _unflatten(V, *_FactorAndSolve.apply(L, E, b)), (0, 0, 0)


class _Solve(torch.autograd.Function):
    """mode "solve": x = P^T Lp^{-T} Lp^{-1} P b; "forward": y = Lp^{-1} P b;
    "backward": x = P^T Lp^{-T} y."""

    @staticmethod
    def forward(Lf, Ef, b, mode):
        return _substitute(mode, Lf, Ef, b)

    @staticmethod
    def setup_context(ctx, inputs, output):
        Lf, Ef, b, mode = inputs
        ctx.mode = mode
        ctx.save_for_backward(Lf, Ef, b, output)

    @staticmethod
    def backward(ctx, gout):
        Lf, Ef, b, out = ctx.saved_tensors
        mode = ctx.mode
        need_factor = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        dL = dE = None
        if mode == "forward":
            # y = Lp^{-1} b:  bbar = Lp^{-T} ybar,  G = -bbar y^T
            bbar = _substitute("backward", Lf, Ef, gout)
            if need_factor:
                dL, dE = _factor_grads(Lf, Ef, bbar, out)
        elif mode == "backward":
            # x = Lp^{-T} y:  ybar = w = Lp^{-1} xbar,  G = -x w^T
            bbar = _substitute("forward", Lf, Ef, gout)
            if need_factor:
                dL, dE = _factor_grads(Lf, Ef, out, bbar)
        elif need_factor:
            # both: w = Lp^{-1} xbar, bbar = Lp^{-T} w, y = Lp^{-1} b,
            # G = -(x w^T + bbar y^T)
            w = _substitute("forward", Lf, Ef, gout)
            bbar = _substitute("backward", Lf, Ef, w)
            y = _substitute("forward", Lf, Ef, b)
            dL, dE = _factor_grads(Lf, Ef, out, w, bbar, y)
        else:
            bbar = _substitute("solve", Lf, Ef, gout)
        return (dL if ctx.needs_input_grad[0] else None, dE if ctx.needs_input_grad[1] else None,
                bbar if ctx.needs_input_grad[2] else None, None)

    @staticmethod
    def vmap(info, in_dims, Lf, Ef, b, mode):
        (Lf, Ef, b), V = _flatten_vmapped(info, in_dims[:3], Lf, Ef, b)
        return _unflatten(V, _Solve.apply(Lf, Ef, b, mode))[0], 0


# ---------------------------------------------------------------------------
# public functions (jax.py:270-337)
# ---------------------------------------------------------------------------

def _to_batched(*ts):
    """Merges all leading batch dimensions into one (a 3-D call gets a batch
    of one, jax.py:65-69); returns the 4-D tensors and the batch shape."""
    lead = ts[0].shape[:-3]
    B = 1
    for d in lead:
        B *= d
    return [None if t is None else t.reshape(B, *t.shape[-3:]) for t in ts], lead


def _restore(lead, t):
    return t.reshape(*lead, *t.shape[1:])


def cholesky_factor(L: torch.Tensor, E: torch.Tensor, pad_problem: bool = True):
    """ND Cholesky factor of (L, E); returns (L_factor, E_factor), E_factor at
    full storage length (jax.py:270-277).  Inputs are not modified."""
    _validate_problem_shapes(L, E)
    (L4, E4), lead = _to_batched(L, E)
    n = L4.shape[2]
    E4 = _ensure_E_size(L4.shape[1], E4)
    if pad_problem:
        L4, E4, _ = _pad_factor_data(L4, E4, None, factor=True)
    Lf, Ef = _Factor.apply(L4, E4)
    return _restore(lead, Lf[..., :n, :n]), _restore(lead, Ef[..., :n, :n])


def cholesky_factor_and_solve(L: torch.Tensor, E: torch.Tensor, b: torch.Tensor, pad_problem: bool = True):
    """Factor and solve; returns (L_factor, E_factor, x) (jax.py:296-305).
    x is differentiable with respect to L, E and b."""
    _validate_problem_shapes(L, E, b)
    (L4, E4, b4), lead = _to_batched(L, E, b)
    n = L4.shape[2]
    E4 = _ensure_E_size(L4.shape[1], E4)
    if pad_problem:
        L4, E4, b4 = _pad_factor_data(L4, E4, b4, factor=True)
    Lf, Ef, x = _FactorAndSolve.apply(L4, E4, b4)
    return (_restore(lead, Lf[..., :n, :n]), _restore(lead, Ef[..., :n, :n]), _restore(lead, x[..., :n, :]))


def _substitution(mode, L, E, b, pad_problem):
    _validate_problem_shapes(L, E, b)
    (L4, E4, b4), lead = _to_batched(L, E, b)
    n = L4.shape[2]
    E4 = _ensure_E_size(L4.shape[1], E4)
    if pad_problem:
        L4, E4, b4 = _pad_factor_data(L4, E4, b4, factor=False)
    out = _Solve.apply(L4, E4, b4, mode)
    return _restore(lead, out[..., :n, :])


def cholesky_solve(L: torch.Tensor, E: torch.Tensor, b: torch.Tensor, pad_problem: bool = True):
    """x = Psi^{-1} b with the factor (L, E) from cholesky_factor (jax.py:283-290)."""
    return _substitution("solve", L, E, b, pad_problem)


def forward_substitution(L: torch.Tensor, E: torch.Tensor, b: torch.Tensor, pad_problem: bool = True):
    """y = Lp^{-1} P b with the factor (L, E) (jax.py:311-320)."""
    return _substitution("forward", L, E, b, pad_problem)


def backward_substitution(L: torch.Tensor, E: torch.Tensor, y: torch.Tensor, pad_problem: bool = True):
    """x = P^T Lp^{-T} y with the factor (L, E) (jax.py:326-335)."""
    return _substitution("backward", L, E, y, pad_problem)
