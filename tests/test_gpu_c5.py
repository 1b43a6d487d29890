"""The C5 ensemble at its own size: 32 dry members of uneven extents
(nx, ny ~ U{96..256}, nz ~ U{40..60} from the workload's seed, dx = 3 km,
dt = 18 s, 4 substeps, the w damper over the top 5 km).

- two batched steps in one captured HIP graph (four member groups as
  parallel branches, rejoining after every step) leave every member bit
  for bit where the same two steps of arw_step, member after member,
  leave it;
- the smallest member is then bit-identical to the float32 oracle after
  the same two steps, so the batched launch is checked against the
  restatement of the reference step, not only against the single-domain
  kernels it shares.
"""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from oracle import dycore_np as O
from tests.helpers import DIAG, PROG, compare, oracle_and_gpu

pytestmark = pytest.mark.gpu

SEED = 20260731


def test_c5_batched_graph_matches_sequential_and_oracle():
    from arwen_amd import ensemble as E
    from arwen_amd.state import DomainState
    shapes = cases.ensemble_shapes(32, SEED)
    small = int(np.argmin([nx * ny * nz for nx, ny, nz in shapes]))
    print(f"C5 32 members, {sum(a * b * c for a, b, c in shapes)} points; "
          f"smallest member {small} {shapes[small]}", flush=True)
    states, o, cfg_small = [], None, None
    for i, (nx, ny, nz) in enumerate(shapes):
        h = cases.ensemble_member(nx, ny, nz, SEED + 1 + i)
        if i == small:
            o, g = oracle_and_gpu(h)
            cfg_small = h.cfg
            states.append(g)
        else:
            states.append(DomainState.from_host(h))
        del h
    cfg = states[0].cfg
    init = E.save_state(states)
    with E.Ensemble(states, cfg, groups=4, join_every=1) as ens:
        ens.capture(2)
        ens.launch()
    torch.cuda.synchronize()
    batched = E.save_state(states)
    E.restore_state(states, init)
    del init
    E.sequential_steps(states, cfg, 2)
    torch.cuda.synchronize()
    for m, (st, (sv, _)) in enumerate(zip(states, batched)):
        for a in E.STATE_ARRAYS:
            assert torch.equal(getattr(st, a), sv[a]), (m, a)
    for _ in range(2):
        O.step(o, cfg_small)
    res = compare(o, states[small], PROG + DIAG)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
