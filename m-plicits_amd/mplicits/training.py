"""Hand-written forward / input-gradient / backward of a SIREN level behind
torch.autograd.Function, and the fused residual-training step.

The reference trains with autograd: model(x) -> f, gradient(f, x) with
create_graph=True (i3d/diff_operators.py:259-265), losses on f and grad f
(loss_add_detail_fine, i3d/loss_functions.py:327-394), then a double
backward (experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252).
Here the same quantities come from explicit passes:

  forward with tangents: every activation block carries the value and T input
  tangents ([n][1 + T][W]); one GEMM per layer applies W_j to all of them, a
  HIP kernel adds the bias and applies sin / cos (csrc/train.hip).  With the
  three unit tangents the output tangents are grad f.
  backward: seeds on f (s) and on the tangent outputs (g_t), propagated with
  the chain rule including the second-order term of the tangent products
  (-sin a_j * ad_j * pdbar_j); weight gradients are GEMMs over all point and
  tangent rows.

The matrix products are plain GEMMs issued through torch (hipBLASLt): FP32, or
BF16 / FP16 operands with FP32 accumulation and output.  Activations are not
kept between the forward and the second-order backward: the backward
recomputes them (only x, the seeds and grad f are saved).
"""
import torch

from . import _lib
from .stack import _linear_layers

_DT = {"fp32": torch.float32, "bf16": torch.bfloat16, "fp16": torch.float16}
_CODE = {"fp32": 0, "bf16": 1, "fp16": 2}  # element type code of the C ABI


def baked_layers(model):
    """[(W_eff, b_eff, omega)] with the sine frequency folded in (views of the
    parameters scaled by omega; the last layer unscaled)."""
    out = []
    layers = _linear_layers(model)
    for k, (w, b, om) in enumerate(layers):
        if om != 1.0:
            out.append((w * om, b * om, om))
        else:
            out.append((w, b, 1.0))
    return out


def _mm(a, b, precision):
    """a @ b with FP32 output (BF16 / FP16 operands accumulate in FP32)."""
    if precision == "fp32":
        return torch.mm(a, b)
    return torch.mm(a, b, out_dtype=torch.float32)


def _gemm_weight(w, precision):
    return w if precision == "fp32" else w.to(_DT[precision])


def unit_tangents(n, device):
    """[n, 3, 3]: the three coordinate directions for every point."""
    return torch.eye(3, device=device, dtype=torch.float32).expand(n, 3, 3).contiguous()


@torch.no_grad()
def _blocked_dw(precision, n, nt, wout, win):
    """Whether the weight gradient of a wout x win hidden layer runs as the
    split-K kernel on K-blocked operand copies (16-bit, 3 tangents)."""
    return (precision != "fp32" and nt == 3 and n % 16 == 0 and wout % 256 == 0 and win % 256 == 0)


def forward_tangents(layers, x, V, precision="fp32"):
    """f [n] and its directional derivatives fd [n, T] along V [n, T, 3].

    Returns (f, fd, cache); the cache holds the pre-activations Z_j (FP32,
    [n, 1+T, W]) and the GEMM inputs X_j of every layer for the backward (and,
    where the blocked weight-gradient kernel applies, their K-blocked copies)."""
    lib = _lib.load()
    n = x.shape[0]
    nt = 0 if V is None else V.shape[1]
    dt = _DT[precision]
    stream = _lib.stream_handle()
    W0, b0, _ = layers[0]
    w = W0.shape[0]
    hidden = [lay[0].shape for lay in layers[1:-1]]  # (wout, win) of the hidden GEMMs

    def blocked(j, width):
        # X_j feeds hidden GEMM j + 1
        if j < len(hidden) and _blocked_dw(precision, n, nt, hidden[j][0], width):
            return torch.empty(n * (1 + nt) * width, device=x.device, dtype=dt)
        return None

    Z = torch.empty(n, 1 + nt, w, device=x.device, dtype=torch.float32)
    X = torch.empty(n, 1 + nt, w, device=x.device, dtype=dt)
    Xb = blocked(0, w)
    _lib.check(lib.mp_train_first(_lib.ptr(x), _lib.ptr(V), _lib.ptr(W0.detach().contiguous()),
                                  _lib.ptr(b0.detach().contiguous()), n, w, nt, _lib.ptr(Z), _lib.ptr(X),
                                  _CODE[precision], _lib.ptr(Xb), stream))
    Zs, Xs, Xbs = [Z], [X], [Xb]
    b_sine = None
    for j, (Wj, bj, _) in enumerate(layers[1:-1], start=1):
        wout = Wj.shape[0]
        Zj = _mm(Xs[-1].view(-1, Xs[-1].shape[-1]), _gemm_weight(Wj.detach(), precision).t(), precision)
        Zj = Zj.view(n, 1 + nt, wout)
        if j == len(layers) - 2:
            # the last sine layer's X is no GEMM operand: the output layer
            # reads Z and adds this layer's bias to it in place
            b_sine = bj.detach().contiguous()
            Zs.append(Zj)
            Xs.append(None)
            Xbs.append(None)
            break
        Xj = torch.empty(n, 1 + nt, wout, device=x.device, dtype=dt)
        Xbj = blocked(j, wout)
        _lib.check(lib.mp_train_act(_lib.ptr(Zj), _lib.ptr(bj.detach().contiguous()), n, wout, nt, _lib.ptr(Xj),
                                    _CODE[precision], _lib.ptr(Xbj), stream))
        Zs.append(Zj)
        Xs.append(Xj)
        Xbs.append(Xbj)
    WL, bL, _ = layers[-1]
    # output layer in FP32 from the FP32 pre-activations of the last sine layer
    f = torch.empty(n, device=x.device, dtype=torch.float32)
    fd = torch.empty(n, nt, device=x.device, dtype=torch.float32)
    _lib.check(lib.mp_train_out_fwd(_lib.ptr(Zs[-1]), _lib.ptr(b_sine), _lib.ptr(WL.detach().contiguous()),
                                    _lib.ptr(bL.detach().contiguous()), n, Zs[-1].shape[-1], nt, _lib.ptr(f),
                                    _lib.ptr(fd if nt else None), stream))
    return f, fd, (Zs, Xs, Xbs)


@torch.no_grad()
def backward_tangents(layers, x, V, cache, s=None, g=None, precision="fp32"):
    """Parameter gradients of sum_i [s_i f_i + sum_t g_it fd_it] (effective,
    frequency-baked parameters).  Returns [(dW_k, db_k)] per layer."""
    lib = _lib.load()
    Zs, Xs, Xbs = cache
    n = x.shape[0]
    nt = Zs[0].shape[1] - 1
    dt = _DT[precision]
    stream = _lib.stream_handle()
    WL, bL, _ = layers[-1]
    wl = WL.detach()[0].contiguous()
    zl = Zs[-1]
    wlast = zl.shape[-1]
    L = len(layers)
    nblk = int(lib.mp_train_out_bwd_blocks(n))
    gc = g.contiguous() if g is not None else None
    dbl = s.sum().reshape(1) if s is not None else torch.zeros(1, device=x.device)
    grads = [None] * L

    def adjoint(j):
        # Abar_j and, where the split-K kernel applies, its K-blocked copy
        A = torch.empty(n, 1 + nt, layers[j][0].shape[0], device=x.device, dtype=dt)
        return A, (torch.empty_like(A) if Xbs[j - 1] is not None else None)

    def hidden_bwd(j, A, Ab, db):
        # dW_j from Abar_j and X_{j-1}; returns the seeds [pbar; pdbar]_{j-1}
        Wj = layers[j][0].detach()
        wout, win = Wj.shape
        A2 = A.view(-1, wout)
        if Ab is not None:
            rows = n * (1 + nt)
            part = torch.empty(int(lib.mp_train_dw_chunks(rows, wout, win)), wout, win, device=x.device,
                               dtype=torch.float32)
            _lib.check(lib.mp_train_dw(_lib.ptr(Ab), _lib.ptr(Xbs[j - 1]), rows, wout, win, _CODE[precision],
                                       _lib.ptr(part), stream))
            dW = part.sum(0)
        else:
            dW = _mm(A2.t(), Xs[j - 1].view(-1, win), precision)
        grads[j] = (dW, db)
        if precision != "fp32" and win % 4 == 0 and win <= 512:
            # 16-bit seeds: their consumers round to 16 bit right after, and
            # the GEMM writes half the bytes
            return torch.mm(A2, _gemm_weight(Wj, precision)).view(n, 1 + nt, win)
        return _mm(A2, _gemm_weight(Wj, precision), precision).view(n, 1 + nt, win)

    def pcode(P):
        return {torch.float32: 0, torch.bfloat16: 1, torch.float16: 2}[P.dtype]

    if L >= 3 and wlast % 4 == 0 and wlast <= 512:
        # output layer and last sine layer in one pass: the rank-one seeds
        # [s w_L; g_t w_L] are never stored
        A, Ab = adjoint(L - 2)
        part = torch.empty(nblk, 2, wlast, device=x.device, dtype=torch.float32)
        _lib.check(lib.mp_train_out_act_bwd(_lib.ptr(zl), _lib.ptr(s), _lib.ptr(gc), _lib.ptr(wl), n, wlast, nt,
                                            _lib.ptr(A), _CODE[precision], _lib.ptr(Ab), _lib.ptr(part), stream))
        tot = part.sum(0)
        grads[-1] = (tot[0].reshape(1, -1), dbl)
        Pbar = hidden_bwd(L - 2, A, Ab, tot[1])
        first = L - 3
    else:
        # output layer: seeds of the last sine layer and dw_L in one pass
        Pbar = torch.empty(n, 1 + nt, wlast, device=x.device, dtype=torch.float32)
        part = torch.empty(nblk, wlast, device=x.device, dtype=torch.float32)
        _lib.check(lib.mp_train_out_bwd(_lib.ptr(zl), _lib.ptr(s), _lib.ptr(gc), _lib.ptr(wl), n, wlast, nt,
                                        _lib.ptr(Pbar), _lib.ptr(part), stream))
        grads[-1] = (part.sum(0).reshape(1, -1), dbl)
        first = L - 2
    for j in range(first, 0, -1):
        wout = layers[j][0].shape[0]
        A, Ab = adjoint(j)
        if wout % 4 == 0 and wout <= 512:
            dbp = torch.empty(nblk, wout, device=x.device, dtype=torch.float32)
            _lib.check(lib.mp_train_act_bwd_part(_lib.ptr(Zs[j]), _lib.ptr(Pbar), pcode(Pbar), n, wout, nt,
                                                 _lib.ptr(A), _CODE[precision], _lib.ptr(Ab), _lib.ptr(dbp), stream))
            db = dbp.sum(0)
        else:
            aval = torch.empty(n, wout, device=x.device, dtype=torch.float32)
            Pbar = Pbar.float()
            _lib.check(lib.mp_train_act_bwd(_lib.ptr(Zs[j]), _lib.ptr(Pbar), n, wout, nt, _lib.ptr(A),
                                            _CODE[precision], _lib.ptr(aval), _lib.ptr(Ab), stream))
            db = aval.sum(0)
        Pbar = hidden_bwd(j, A, Ab, db)
    w0 = Zs[0].shape[-1]
    if w0 % 4 == 0 and w0 <= 512:
        # layer-0 adjoints reduced against [x; v_t] in the same pass
        part = torch.empty(int(lib.mp_train_first_bwd_blocks(n)), w0, 4, device=x.device, dtype=torch.float32)
        _lib.check(lib.mp_train_first_bwd(_lib.ptr(Zs[0]), _lib.ptr(Pbar), pcode(Pbar), _lib.ptr(x),
                                          _lib.ptr(V.contiguous() if nt else None), n, w0, nt, _lib.ptr(part),
                                          stream))
        tot = part.sum(0)
        grads[0] = (tot[:, :3].contiguous(), tot[:, 3].contiguous())
        return grads
    A0 = torch.empty(n, 1 + nt, w0, device=x.device, dtype=torch.float32)
    aval0 = torch.empty(n, w0, device=x.device, dtype=torch.float32)
    _lib.check(lib.mp_train_act_bwd(_lib.ptr(Zs[0]), _lib.ptr(Pbar), n, w0, nt, _lib.ptr(A0), 0, _lib.ptr(aval0),
                                    None, stream))
    # dW0 = sum over value and tangent rows of A0 (x) [x | v_t]: one GEMM over
    # the contiguous [n (1 + T), w0] block
    X0 = x[:, None, :] if not nt else torch.cat([x[:, None, :], V], 1)
    dW0 = torch.mm(A0.view(-1, w0).t(), X0.reshape(-1, 3))
    grads[0] = (dW0, aval0.sum(0))
    return grads


def _unbake(layers, grads):
    """Gradients w.r.t. the module parameters (W, b) from those w.r.t. the
    baked (omega W, omega b)."""
    out = []
    for (_, _, om), (dW, db) in zip(layers, grads):
        out.append(dW * om if om != 1.0 else dW)
        out.append(db * om if om != 1.0 else db)
    return out


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
        from .fused import model_stack
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
        layers = baked_layers(holder.model)
        xf = x.detach().float().reshape(-1, 3).contiguous()
        n = xf.shape[0]
        V = unit_tangents(n, xf.device)
        f, fd, cache = forward_tangents(layers, xf, V, holder.precision)
        gx = s[:, None] * fd
        if any(ctx.needs_input_grad[3:]):
            grads = _unbake(layers, backward_tangents(layers, xf, V, cache, s=s, precision=holder.precision))
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
        layers = baked_layers(holder.model)
        xf = x.detach().float().reshape(-1, 3).contiguous()
        n = xf.shape[0]
        gbar = gbar.reshape(n, 3).float()
        ds = (gbar * fd).sum(-1)
        grads = [None] * len(params)
        if any(ctx.needs_input_grad[3:]):
            V = unit_tangents(n, xf.device)
            _, _, cache = forward_tangents(layers, xf, V, holder.precision)  # recomputed
            g = (s[:, None] * gbar).contiguous()
            grads = _unbake(layers, backward_tangents(layers, xf, V, cache, g=g, precision=holder.precision))
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


def train_step(model, x, loss_fn, frozen=None, optimizer=None, precision="fp32"):
    """One fused training step of ``model`` (the trainable residual level).

    frozen: optional LevelStack of the lower levels (their value and gradient
    are added, not trained).  loss_fn(F, grad F) -> scalar tensor, F and
    grad F the totals.  One forward with the three input tangents and one
    backward with the seeds of both loss terms (value and gradient) replace
    the forward / create_graph gradient / double backward of autograd.
    Returns (loss, F, grad F).
    """
    layers = baked_layers(model)
    x = x.float().contiguous()
    n = x.shape[0]
    V = unit_tangents(n, x.device)
    f, fd, cache = forward_tangents(layers, x, V, precision)
    if frozen is not None:
        if precision == "fp32" and max(max(m.hidden_widths()) for m in frozen.models) >= 256:
            # at these widths the FP32 level kernel computes gradients in forward
            # mode (four passes); the layered pass with FP32 GEMMs is faster
            # (6.8 ms against 9.1 ms for two 256x3 levels at 2^18 points)
            fc = torch.zeros_like(f)
            gc = torch.zeros_like(fd)
            for m in frozen.models:
                fm, gm, _ = forward_tangents(baked_layers(m), x, V, "fp32")
                fc += fm
                gc += gm
        else:
            fc, gc = frozen.eval(x, grad=True)
        f_tot, g_tot = f + fc, fd + gc
    else:
        f_tot, g_tot = f, fd
    F = f_tot.detach().requires_grad_(True)
    G = g_tot.detach().requires_grad_(True)
    with torch.enable_grad():
        loss = loss_fn(F, G)
        s, g = torch.autograd.grad(loss, (F, G))
    grads = _unbake(layers, backward_tangents(layers, x, V, cache, s=s.contiguous(), g=g.contiguous(),
                                              precision=precision))
    params = [p for seq in model.net for p in (seq[0].weight, seq[0].bias)]
    for p, gp in zip(params, grads):
        p.grad = gp.reshape(p.shape).to(p.dtype) if p.grad is None else p.grad + gp.reshape(p.shape)
    if optimizer is not None:
        optimizer.step()
    return loss.detach(), f_tot, g_tot
