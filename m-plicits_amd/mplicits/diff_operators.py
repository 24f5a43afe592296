"""Input gradient of a model output (i3d/diff_operators.py:259-265).

With this package's SIREN the call differentiates through the hand-written
forward/backward Functions (mplicits.training), so the result can itself be
differentiated (create_graph=True), as the residual losses require."""
import torch


def gradient(y, x, grad_outputs=None):
    """d y / d x with create_graph=True."""
    if grad_outputs is None:
        grad_outputs = torch.ones_like(y)
    return torch.autograd.grad(y, [x], grad_outputs=grad_outputs, create_graph=True)[0]
