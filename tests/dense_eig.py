"""Lowest eigenvalues of the pair supermatrix V[P][Q] = (ik|jl) by a dense
FP64 eigensolve (torch.linalg.eigvalsh on the GPU), run in a child process
that imports torch before any other GPU library: the reference for the
iterative eigensolvers at sizes numpy cannot factor in seconds.  Test
infrastructure only."""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

_CHILD = r"""
import json, sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[4])
from slater_stoch_amd import basis as B
n, k = int(sys.argv[2]), int(sys.argv[3])
vals = np.load(sys.argv[1])
is_, js, ks, ls = B.quartet_mapping(n)
P = torch.from_numpy((is_ * (is_ - 1)) // 2 + ks - 1).cuda()
Q = torch.from_numpy((js * (js - 1)) // 2 + ls - 1).cuda()
N = n * (n + 1) // 2
V = torch.zeros((N, N), dtype=torch.float64, device="cuda")
x = torch.from_numpy(vals).cuda()
V[P, Q] = x
V[Q, P] = x
w = torch.linalg.eigvalsh(V)
print(json.dumps([float(v) for v in w[:k].cpu()]))
"""


def lowest_eigenvalues(n, vals, k, timeout=600):
    """The k lowest eigenvalues of V built from the kcp-ordered list vals of
    an n-AO basis."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "vals.npy")
        np.save(f, np.ascontiguousarray(vals, np.float64))
        out = subprocess.run([sys.executable, "-c", _CHILD, f, str(n), str(k), root],
                             capture_output=True, text=True, timeout=timeout)
    if out.returncode != 0:
        raise RuntimeError(f"dense eigensolve failed: {out.stderr[-2000:]}")
    return np.array(json.loads(out.stdout.strip().splitlines()[-1]))
