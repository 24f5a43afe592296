"""Training of a SIREN level on the HIP library: the hand-written forward,
input gradient and second-order backward behind torch.autograd.Function, and
the fused residual-training step.

The reference trains with autograd: model(x) -> f, gradient(f, x) with
create_graph=True (i3d/diff_operators.py:259-265), losses on f and grad f
(loss_add_detail_fine, i3d/loss_functions.py:327-394), then a double
backward (experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252).
Here the same quantities come from two library calls per level:

  mp_train_fwd: f and grad f (the forward with the three coordinate
  tangents: every activation block carries the value and the tangents, one
  matrix-core GEMM per layer applies W_j to all of them, the activation runs
  in the GEMM's epilogue);
  mp_train_bwd: the parameter gradients of sum_p [s_p f_p + g_p . grad f_p]
  for the loss seeds s = dL/df, g = dL/dgrad f (the chain rule through the
  tangent products included), recomputing the activations unless the
  forward kept them.

Every product runs on the library's own GEMMs (FP32, BF16 or FP16 operands
with FP32 accumulation); the loss itself stays in torch (it is elementwise on
n values and gradients) and only supplies the seeds.
"""
import torch

from . import _lib
from .fused import model_stack
from .stack import _linear_layers


def _layer_shapes(model):
    """[(out, in, omega)] of every Linear of the level."""
    return [(w.shape[0], w.shape[1], om) for w, _, om in _linear_layers(model)]


def level_forward(model, x, precision="fp32", keep=False, stack=None, level=0):
    """(f [n], grad f [n, 3]) of one level at x [n, 3] (float32, on the GPU).

    keep: leave the activations in the stack's workspace for the next
    level_backward of the same points (the training step's backward)."""
    lib = _lib.load()
    st = stack if stack is not None else model_stack(model, precision)
    st.refresh()
    x = x.float().contiguous()
    n = x.shape[0]
    f = torch.empty(n, device=x.device, dtype=torch.float32)
    g = torch.empty(n, 3, device=x.device, dtype=torch.float32)
    with torch.cuda.device(x.device):
        _lib.check(lib.mp_train_fwd(st._handle, level, _lib.ptr(x), n, _lib.ptr(f), _lib.ptr(g), int(bool(keep)),
                                    _lib.stream_handle()))
    return f, g


def level_backward(model, x, s=None, g=None, precision="fp32"):
    """Gradients of sum_p [s_p f_p + g_p . grad f_p] with respect to the
    module's parameters, in module order (frequencies unbaked: the kernels
    differentiate the baked weights omega W, omega b)."""
    lib = _lib.load()
    st = model_stack(model, precision)
    st.refresh()
    x = x.float().contiguous()
    n = x.shape[0]
    count = int(lib.mp_train_param_count(st._handle, 0))
    if count < 0:
        _lib.check(_lib.MP_EINVAL)
    dparams = torch.empty(count, device=x.device, dtype=torch.float32)
    s = s.float().contiguous() if s is not None else None
    g = g.float().contiguous() if g is not None else None
    with torch.cuda.device(x.device):
        _lib.check(lib.mp_train_bwd(st._handle, 0, _lib.ptr(x), n, _lib.ptr(s), _lib.ptr(g), _lib.ptr(dparams),
                                    _lib.stream_handle()))
    shapes = _layer_shapes(model)
    n_w = sum(o * i for o, i, _ in shapes)
    grads = []
    wo, bo = 0, n_w
    for out, inp, om in shapes:
        dw = dparams[wo:wo + out * inp].view(out, inp)
        db = dparams[bo:bo + out]
        grads += [dw * om, db * om] if om != 1.0 else [dw, db]
        wo += out * inp
        bo += out
    return grads


class _Holder:
    """Carries the module and precision through autograd (non-tensor argument)."""

    def __init__(self, model, precision):
        self.model = model
        self.precision = precision


class _SirenValue(torch.autograd.Function):
    """f = SIREN(x); backward gives s * grad f for x and s * df/dtheta for the
    parameters through _SirenGrad, so gradient(f, x, create_graph=True) can be
    differentiated again."""

    @staticmethod
    def forward(ctx, x, holder, *params):
        ctx.holder = holder
        ctx.save_for_backward(x, *params)
        f = model_stack(holder.model, holder.precision).eval(x.detach().float().contiguous())
        return f.unsqueeze(-1)

    @staticmethod
    def backward(ctx, fbar):
        x, *params = ctx.saved_tensors
        outs = _SirenGrad.apply(x, fbar.reshape(-1).float().contiguous(), ctx.holder, *params)
        gx, gparams = outs[0], outs[1:]
        return (gx.reshape(x.shape), None, *gparams)


class _SirenGrad(torch.autograd.Function):
    """(s * grad_x f, s * df/dtheta); its backward is the second-order pass."""

    @staticmethod
    def forward(ctx, x, s, holder, *params):
        xf = x.detach().float().reshape(-1, 3).contiguous()
        need_params = any(ctx.needs_input_grad[3:])
        _, fd = level_forward(holder.model, xf, holder.precision, keep=need_params)
        gx = s[:, None] * fd
        if need_params:
            grads = level_backward(holder.model, xf, s=s, precision=holder.precision)
        else:
            grads = [None] * len(params)
        ctx.holder = holder
        ctx.save_for_backward(x, s, fd, *params)
        return (gx, *grads)

    @staticmethod
    def backward(ctx, gbar, *pbar):
        if any(p is not None and bool((p != 0).any()) for p in pbar):
            raise NotImplementedError("second derivatives with respect to the parameters are not implemented")
        x, s, fd, *params = ctx.saved_tensors
        holder = ctx.holder
        xf = x.detach().float().reshape(-1, 3).contiguous()
        n = xf.shape[0]
        gbar = gbar.reshape(n, 3).float()
        ds = (gbar * fd).sum(-1)
        grads = [None] * len(params)
        if any(ctx.needs_input_grad[3:]):
            grads = level_backward(holder.model, xf, g=(s[:, None] * gbar).contiguous(), precision=holder.precision)
        return (None, ds, None, *grads)


def siren_apply(model, coords):
    """model_out of ``model`` at ``coords`` with a differentiable graph through
    the hand-written kernels (first and second order)."""
    holder = _Holder(model, getattr(model, "precision", "fp32"))
    params = [p for seq in model.net for p in (seq[0].weight, seq[0].bias)]
    shape = coords.shape
    x = coords.reshape(-1, 3)
    y = _SirenValue.apply(x, holder, *params)
    return y.reshape(*shape[:-1], 1)


def config_loss(f, grad, f_t, n_t):
    """mean|f - f_t| + mean||grad f| - 1| + mean(1 - cos(grad f, n_t))."""
    gn = grad.norm(dim=-1)
    cos = (grad * n_t).sum(-1) / (gn * n_t.norm(dim=-1)).clamp_min(1e-8)
    return (f - f_t).abs().mean() + (gn - 1).abs().mean() + (1 - cos).mean()


def frozen_value_grad(frozen, x):
    """Sum of the frozen levels' values and gradients at x.  FP32 stacks of
    levels with >= 2 sine layers run the training forward (FP32 matrix-core
    GEMMs over all points and tangents at once); otherwise the level kernels
    (mp_eval with gradients)."""
    models = frozen.models
    if frozen.precision == "fp32" and all(len(m.net) >= 3 for m in models):
        fc = torch.zeros(x.shape[0], device=x.device)
        gc = torch.zeros(x.shape[0], 3, device=x.device)
        for k, m in enumerate(models):
            fk, gk = level_forward(m, x, stack=frozen, level=k)
            fc += fk
            gc += gk
        return fc, gc
    return frozen.eval(x, grad=True)


def train_step(model, x, loss_fn, frozen=None, optimizer=None, precision="fp32"):
    """One fused training step of ``model`` (the trainable residual level).

    frozen: optional LevelStack of the lower levels (their value and gradient
    are added, not trained).  loss_fn(F, grad F) -> scalar tensor, F and
    grad F the totals.  The forward keeps its activations, the loss gives the
    seeds of both terms (value and gradient), and one backward replaces the
    forward / create_graph gradient / double backward of autograd.
    Returns (loss, F, grad F).
    """
    x = x.float().contiguous()
    f, fd = level_forward(model, x, precision, keep=True)
    if frozen is not None:
        fc, gc = frozen_value_grad(frozen, x)
        f_tot, g_tot = f + fc, fd + gc
    else:
        f_tot, g_tot = f, fd
    F = f_tot.detach().requires_grad_(True)
    G = g_tot.detach().requires_grad_(True)
    with torch.enable_grad():
        loss = loss_fn(F, G)
        s, g = torch.autograd.grad(loss, (F, G))
    grads = level_backward(model, x, s=s, g=g, precision=precision)
    params = [p for seq in model.net for p in (seq[0].weight, seq[0].bias)]
    for p, gp in zip(params, grads):
        p.grad = gp.reshape(p.shape).to(p.dtype) if p.grad is None else p.grad + gp.reshape(p.shape)
    if optimizer is not None:
        optimizer.step()
    return loss.detach(), f_tot, g_tot
