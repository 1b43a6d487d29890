"""GPU checks of the batched ensemble (arw_ensemble_*): members of uneven
extents advanced by one launch per pass agree bit for bit with stepping
each member alone through arw_step -- eagerly, through a replayed HIP
graph (with eager work between replays), with several member groups and
reordered work tables, on flat and terrain bases, in FP32 and FP64."""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from arwen_amd.ensemble import STATE_ARRAYS

pytestmark = pytest.mark.gpu

SHAPES = [(24, 20, 12), (37, 16, 10), (16, 33, 14), (50, 9, 11), (8, 8, 16)]


def _members(dtype=torch.float32, terrain=False):
    from arwen_amd.state import DomainState
    out = []
    for i, (nx, ny, nz) in enumerate(SHAPES):
        if terrain:
            case = cases.stratified_hill(nx=nx, ny=ny, nz=nz, seed=100 + i,
                                         radius=20000.0)
        else:
            case = cases.ensemble_member(nx, ny, nz, seed=100 + i)
        out.append((case, DomainState.from_host(case, dtype=dtype)))
    return out


def _pair(dtype=torch.float32, terrain=False):
    a = _members(dtype, terrain)
    b = _members(dtype, terrain)
    return a[0][0].cfg, [s for _, s in a], [s for _, s in b]


def _assert_same(sa, sb):
    for i, (x, y) in enumerate(zip(sa, sb)):
        for n in STATE_ARRAYS:
            assert torch.equal(getattr(x, n), getattr(y, n)), (i, n)


@pytest.mark.parametrize("terrain", [False, True])
def test_eager_ensemble_matches_member_steps_bitwise(terrain):
    from arwen_amd import dycore
    from arwen_amd.ensemble import Ensemble
    cfg, seq, ens_states = _pair(terrain=terrain)
    print(f"ensemble eager terrain={terrain}", flush=True)
    for st in seq:
        dycore.run_steps(st, st.cfg, 3)
    with Ensemble(ens_states, cfg) as ens:
        ens.step(3)
    torch.cuda.synchronize()
    _assert_same(seq, ens_states)


@pytest.mark.parametrize("groups,order,join,tile", [
    (1, None, 0, 256), (2, "largest_first", 0, 128),
    (3, "smallest_first", 1, 64)])
def test_graph_replays_match_member_steps_bitwise(groups, order, join, tile):
    from arwen_amd import dycore
    from arwen_amd.ensemble import Ensemble, member_orders
    cfg, seq, ens_states = _pair()
    print(f"ensemble graph groups={groups} order={order} join={join} "
          f"col_tile={tile}", flush=True)
    ordl = member_orders(ens_states)[order] if order else None
    with Ensemble(ens_states, cfg, order=ordl, groups=groups,
                  join_every=join, col_tile=tile) as ens:
        ens.capture(2)
        for rep in range(3):
            ens.launch()
            for st in seq:
                dycore.run_steps(st, st.cfg, 2)
            # eager work of the program's own between replays
            _ = [float(st.u.abs().max()) for st in seq]
            torch.cuda.synchronize()
            _assert_same(seq, ens_states)


def test_fp64_ensemble_matches_member_steps_bitwise():
    from arwen_amd import dycore
    from arwen_amd.ensemble import Ensemble
    cfg, seq, ens_states = _pair(dtype=torch.float64)
    print("ensemble fp64", flush=True)
    for st in seq:
        dycore.run_steps(st, st.cfg, 2)
    with Ensemble(ens_states, cfg, groups=2) as ens:
        ens.capture(2)
        ens.launch()
    torch.cuda.synchronize()
    _assert_same(seq, ens_states)


def test_autotune_leaves_state_and_picks_a_layout():
    from arwen_amd import ensemble as E
    cfg, _, states = _pair()
    print("ensemble autotune", flush=True)
    before = E.save_state(states)
    res = E.autotune(states, cfg, steps=1, reps=1, groups=(1, 2))
    torch.cuda.synchronize()
    for st, (sv, el) in zip(states, before):
        assert st.elapsed_seconds == el
        for n, t in sv.items():
            assert torch.equal(getattr(st, n), t), n
    assert res["groups"] in (1, 2) and sorted(res["order"]) == list(range(5))
    assert res["join_every"] in (1, 5)      # 0 measured as every `steps`
    assert res["col_tile"] in (64, 128, 256)
    # 3 orders, then 3 joins x 2 groups, then 2 more column tiles
    assert len(res["trials"]) == 8


def test_ensemble_rejects_bad_members():
    from arwen_amd.ensemble import Ensemble
    from arwen_amd.state import DomainState
    cfg, _, states = _pair()
    print("ensemble argument checks", flush=True)
    other = cases.ensemble_member(12, 12, 8, seed=1, dt=9.0)
    with pytest.raises(ValueError):
        Ensemble(states + [DomainState.from_host(other)], cfg)
    with pytest.raises(ValueError):
        Ensemble(states, cfg, order=[0, 0, 1, 2, 3])
    from arwen_amd import _lib
    with pytest.raises(_lib.ArwenError):
        Ensemble(states, cfg, groups=9)
    with pytest.raises(_lib.ArwenError, match="col_tile"):
        Ensemble(states, cfg, col_tile=96)


@pytest.mark.parametrize("terrain", [False, True])
def test_precision_form_ensemble_matches_member_steps_bitwise(terrain):
    """The precision form (perturbation EOS, p' array, FP64 tridiagonal)
    through the batched launches equals each member's own precision-form
    step."""
    from arwen_amd import dycore
    from arwen_amd.ensemble import Ensemble
    cfg, seq, ens_states = _pair(terrain=terrain)
    cfg = cfg.with_(precision_form=True)
    for st in seq + ens_states:
        st.cfg = st.cfg.with_(precision_form=True)
    print(f"ensemble precision form terrain={terrain}", flush=True)
    for st in seq:
        dycore.run_steps(st, st.cfg, 2)
    with Ensemble(ens_states, cfg, groups=2) as ens:
        ens.step(2)
    torch.cuda.synchronize()
    _assert_same(seq, ens_states)
