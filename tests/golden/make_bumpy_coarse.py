"""Writes tests/golden/bumpy_coarse_256x3.pth: the coarse level of the
benchmark workload (config 1 of the workload list).

SIREN(3, 1, [256, 256, 256], w0 = ww = 30) with the reference initialisation
under torch.manual_seed(668123), fitted to the seeded bumpy sphere with 500 Adam
steps (lr 1e-4), 20,000 points per step (half uniform in [-1,1]^3, half on the
surface with N(0, 0.01) jitter), loss
    3e3 * mean|f - f_t| + 1e2 * mean(1 - cos(grad f, n_t)) + 5e1 * mean||grad f| - 1|,
then saved with the frequency baked (update_omegas(1)), as the reference's
trainers save checkpoints (train_sdf.py:275-285).  The fit runs in torch float32
on the CPU with autograd; it only produces fixture weights.

Usage: python tests/golden/make_bumpy_coarse.py [--steps 500]
"""
import argparse
import os.path as osp
import sys
import time

import torch

ROOT = osp.dirname(osp.dirname(osp.dirname(osp.abspath(__file__))))
sys.path.insert(0, osp.join(ROOT, "m-plicits_amd"))
sys.path.insert(0, ROOT)

from mplicits.model import SIREN  # noqa: E402
from mplicits.synthetic import MASTER_SEED, BumpySphere, training_batch  # noqa: E402

OUT = osp.join(osp.dirname(osp.abspath(__file__)), "bumpy_coarse_256x3.pth")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--batch", type=int, default=20000)
    ap.add_argument("--out", default=OUT)
    args = ap.parse_args()
    torch.manual_seed(MASTER_SEED)
    model = SIREN(3, 1, [256, 256, 256], w0=30, ww=30)
    shape = BumpySphere()
    params = [p for p in model.parameters()]
    opt = torch.optim.Adam(params, lr=1e-4)
    lins = [seq[0] for seq in model.net]
    omegas = [30.0] * (len(lins) - 1)
    t0 = time.time()
    for step in range(args.steps):
        x, f_t, n_t = training_batch(shape, args.batch, seed=MASTER_SEED + 1000 + 2 * step)
        x.requires_grad_(True)
        h = x
        for lin, om in zip(lins[:-1], omegas):
            h = torch.sin(om * lin(h))
        f = lins[-1](h).squeeze(-1)
        g = torch.autograd.grad(f.sum(), x, create_graph=True)[0]
        cos = torch.nn.functional.cosine_similarity(g, n_t, dim=-1)
        loss = (3e3 * (f - f_t).abs().mean() + 1e2 * (1 - cos).mean()
                + 5e1 * (g.norm(dim=-1) - 1).abs().mean())
        opt.zero_grad()
        loss.backward()
        opt.step()
        if step % 50 == 0 or step == args.steps - 1:
            print(f"step {step} loss {loss.item():.4f} sdf_l1 {(f - f_t).abs().mean().item():.5f} "
                  f"t {time.time() - t0:.1f}s", flush=True)
    model.update_omegas(1)
    torch.save(model.state_dict(), args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
