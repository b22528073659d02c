"""Writes tests/golden/corrector_fixture.npz: inputs, parameters, output and
parameter gradients of the corrector net computed by torch (CPU, float64)
with torch.nn.functional.conv3d and autograd, for a 7 x 9 x 10 grid.

Run from the repository root: python tests/golden/make_corrector_fixture.py
"""
import os

import numpy as np
import torch
import torch.nn.functional as F

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(16, 6, 3, 3, 3), (16,), (16, 16, 3, 3, 3), (16,), (2, 16, 3, 3, 3), (2,)]


def main():
    rng = np.random.default_rng(20261202)
    parts = []
    for shp in SHAPES:
        if len(shp) == 5:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / (shp[1] * 27)), shp).ravel())
        else:
            parts.append(rng.normal(0.0, 0.05, shp).ravel())
    params = np.concatenate(parts)
    nz, ny, nx = 7, 9, 10
    feats = rng.normal(0.0, 1.0, (nz, ny, nx, 6))
    grad_out = rng.normal(0.0, 1.0, (nz, ny, nx, 2))
    t = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    ws, o = [], 0
    for shp in SHAPES:
        n = int(np.prod(shp))
        ws.append(t[o : o + n].reshape(shp))
        o += n
    x = torch.tensor(np.moveaxis(feats, -1, 0)[None], dtype=torch.float64)
    a1 = F.relu(F.conv3d(x, ws[0], ws[1], padding=1))
    a2 = F.relu(F.conv3d(a1, ws[2], ws[3], padding=1))
    out = F.conv3d(a2, ws[4], ws[5], padding=1)
    g = torch.tensor(np.moveaxis(grad_out, -1, 0)[None], dtype=torch.float64)
    (out * g).sum().backward()
    np.savez_compressed(os.path.join(HERE, "corrector_fixture.npz"), params=params, features=feats,
                        grad_out=grad_out, out=np.moveaxis(out.detach().numpy()[0], 0, -1),
                        grad_params=t.grad.numpy())


if __name__ == "__main__":
    main()
