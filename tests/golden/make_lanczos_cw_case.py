"""Writes lanczos_cw_case.json (run on a GPU host): the FP64 Lanczos
coefficients of a device run on instance 178 of disorder_ensemble(512,
seed=0) (chain N=18, Δ = 1.48), 640 steps from the seed-0 device start
vector.  Its lowest Ritz value, the true E0 = -14.35599412 (eigsh), is
also an eigenvalue of T with its first row and column removed to within
the Cullum-Willoughby tolerance, so that test alone labels it spurious.
Usage: python tests/golden/make_lanczos_cw_case.py OUT.json"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main(path):
    import numpy as np
    import __graft_entry__ as ge
    ge.load_package()
    from energy_level_amd.batch import disorder_ensemble
    from energy_level_amd.lanczos import lanczos_run, _start_vector
    from energy_level_amd.plan import Plan
    m = disorder_ensemble(count=512, seed=0)[178]
    P = Plan(m, device='cuda')
    v0 = _start_vector(P.dim, 0, False, P.device)
    R = lanczos_run(P, v0, 640, 1, 0.0, want=2, store=False)
    beta_last = float(R.nrm[R.j].item())
    with open(path, 'w') as f:
        json.dump({'instance': 178, 'n': int(m.n), 'steps': int(R.j), 'alpha': [float(a) for a in R.a],
                   'beta': [float(b) for b in R.b], 'beta_last': beta_last,
                   'eigsh_E0': -14.355994122323, 'eigsh_E1': -14.203646413925}, f)
    P.close()


if __name__ == '__main__':
    main(sys.argv[1])
