"""SIREN evaluation restated on the CPU (oracle; test infrastructure only).

restates i3d/model.py:107-129: y = W_L sin(... sin(W_0 x + b_0) ...) + b_L with
the frequency baked into every layer but the last (i3d/model.py:143-180), and
the input gradient that i3d/diff_operators.py:259-265 obtains with autograd,
here written as the forward-mode chain rule
    J_0 = diag(cos a_0) W_0,  J_i = diag(cos a_i) W_i J_{i-1},  grad = W_L J_{L-1}.
"""
import numpy as np
import torch


def layers_from_model(model, dtype=np.float64):
    """[(W [out,in], b [out])] of an i3d-style SIREN with the frequencies baked."""
    out = []
    n = len(model.net)
    for k, seq in enumerate(model.net):
        lin = seq[0]
        w = lin.weight.detach().cpu().double().numpy()
        b = lin.bias.detach().cpu().double().numpy()
        if k < n - 1:
            om = float(seq[1].w0)
            w, b = w * om, b * om
        out.append((w.astype(dtype), b.astype(dtype)))
    return out


def layers_from_state_dict(sd, omegas=None, dtype=np.float64):
    """Same from a state_dict (net.{k}.0.weight / bias), omegas per sine layer."""
    keys = list(sd.keys())
    pairs = [(sd[keys[i]], sd[keys[i + 1]]) for i in range(0, len(keys), 2)]
    out = []
    for k, (w, b) in enumerate(pairs):
        w = w.detach().cpu().double().numpy()
        b = b.detach().cpu().double().numpy()
        if omegas is not None and k < len(pairs) - 1:
            w, b = w * omegas[k], b * omegas[k]
        out.append((w.astype(dtype), b.astype(dtype)))
    return out


def forward(layers, x):
    """f(x) [n] in the dtype of the layers (FP64 by default)."""
    h = np.asarray(x, dtype=layers[0][0].dtype)
    for w, b in layers[:-1]:
        h = np.sin(h @ w.T + b)
    w, b = layers[-1]
    return (h @ w.T + b)[:, 0]


def forward_grad(layers, x):
    """(f [n], grad f [n, 3]) by the forward-mode chain rule."""
    h = np.asarray(x, dtype=layers[0][0].dtype)
    jac = None
    for k, (w, b) in enumerate(layers[:-1]):
        a = h @ w.T + b
        c = np.cos(a)
        jac = (c[:, :, None] * w[None, :, :]) if k == 0 else c[:, :, None] * np.einsum("oi,nij->noj", w, jac)
        h = np.sin(a)
    w, b = layers[-1]
    return (h @ w.T + b)[:, 0], np.einsum("oi,nij->noj", w, jac)[:, 0, :]


def sum_levels(level_layers, x, grad=False):
    """Sum of several levels (create_mesh_multistage with deltas=[])."""
    if not grad:
        return sum(forward(L, x) for L in level_layers)
    fs, gs = zip(*(forward_grad(L, x) for L in level_layers))
    return sum(fs), sum(gs)


def torch_forward(layers, x):
    """Same computation in torch (any dtype / autograd); layers are torch tensors."""
    h = x
    for w, b in layers[:-1]:
        h = torch.sin(h @ w.T + b)
    w, b = layers[-1]
    return h @ w.T + b


def torch_forward_grad(layers, x):
    """forward_grad in torch (for FP64 evaluation on a device): (f [n], grad f
    [n, 3]) by the same forward-mode chain rule; layers are torch tensors of
    x's dtype and device."""
    h = x
    jac = None
    for k, (w, b) in enumerate(layers[:-1]):
        a = h @ w.T + b
        c = torch.cos(a)[:, :, None]
        jac = c * w[None, :, :] if k == 0 else c * torch.matmul(w, jac)
        h = torch.sin(a)
    w, b = layers[-1]
    return (h @ w.T + b)[:, 0], torch.matmul(w, jac)[:, 0, :]


def reference_init_draws(hidden, w0, ww, seed):
    """Restatement of the RNG call order of SIREN(3, 1, hidden, w0, ww)
    (i3d/model.py:87-105 then reset_weights 131-141): every nn.Linear is
    constructed in layer order (default kaiming-uniform weight, uniform bias),
    then the weights are overwritten in layer order, layer 0 with
    U(-1/fan_in, 1/fan_in) and the others with U(+-sqrt(6/fan_in)/ww).
    Returns [(W, b)] as float32 torch tensors."""
    g = torch.random.get_rng_state()
    torch.manual_seed(seed)
    try:
        widths = [3] + list(hidden) + [1]
        lins = [torch.nn.Linear(widths[k], widths[k + 1]) for k in range(len(widths) - 1)]
        with torch.no_grad():
            for k, lin in enumerate(lins):
                fan_in = lin.weight.shape[1]
                if k == 0:
                    lin.weight.uniform_(-1 / fan_in, 1 / fan_in)
                else:
                    bound = np.sqrt(6 / fan_in) / ww
                    lin.weight.uniform_(-bound, bound)
        return [(l.weight.detach().clone(), l.bias.detach().clone()) for l in lins]
    finally:
        torch.random.set_rng_state(g)
