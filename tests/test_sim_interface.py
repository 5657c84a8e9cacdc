"""The reference's Sim constructor contract (sim.py:553-565) and its callers' keyword sets (viewer.py:137,
bake.py:59), checked without a GPU; the GPU tests construct Sim with the same keyword sets."""
import inspect

import pytest

from nebulaviewer_amd.sim import Sim


def test_reference_positional_order():
    names = list(inspect.signature(Sim.__init__).parameters)[1:11]
    assert names == ["scenario", "res", "params", "ic_mode", "rays", "fuv_grid", "ion_every", "device", "verbose",
                     "fast"]


def test_reference_defaults():
    p = inspect.signature(Sim.__init__).parameters
    assert p["fast"].default is True            # the viewer's interactive cadence (sim.py:565)
    assert p["device"].default == "cuda:0" and p["verbose"].default is True


@pytest.mark.parametrize("kw", [
    # viewer.py:137
    dict(args=("torus",), res=64, params={}, ic_mode="nebula3d", verbose=True),
    # bake.py:59
    dict(args=("torus",), res=64, params={}, ic_mode="nebula3d", rays=(64, 128), fuv_grid=(16, 16, 32), fast=False),
])
def test_caller_keyword_sets_bind(kw):
    kw = dict(kw)
    args = kw.pop("args")
    inspect.signature(Sim.__init__).bind(None, *args, **kw)


def test_matched_noise_is_refused_before_any_device_work():
    with pytest.raises(ValueError, match="matched"):
        Sim("torus", res=16, ic_mode="matched")
    with pytest.raises(ValueError, match="ic_mode"):
        Sim("torus", res=16, ic_mode="other")


def test_time_label_forms():
    s = Sim.__new__(Sim)
    s.phase, s.tp, s.t, s.P = "pre", 0.0123, 0.0, dict(pre_t=0.05)
    assert s.time_label() == "pre-turb 0.012/0.05"          # sim.py:1214-1215
    s.phase, s.t = "main", 0.25
    assert s.time_label() == "t=0.2500"


@pytest.mark.gpu
def test_reference_callers_construct_sim():
    """Sim built with exactly the keyword sets of viewer.py:137 and bake.py:59 (ic_mode='nebula3d': the seeded
    workloads), and the fast default's cadence (1-voxel rays, ionization every >= 3 steps, sim.py:743, 758)."""
    print("[interface] viewer and bake constructor calls", flush=True)
    v = Sim("torus", res=24, params={}, ic_mode="nebula3d", verbose=True)
    assert v.fast and v.dr == v.dx and v.ion_every == 3 and v.ic_mode == "nebula3d"
    v.step(2)
    bk = Sim("torus", res=24, params={}, ic_mode="nebula3d", rays=(32, 64), fuv_grid=(8, 8, 16), fast=False)
    assert not bk.fast and bk.dr == 0.5 * bk.dx and bk.ion_every == 2
    bk.step(2)
    assert bk.time_label().startswith("t=")


@pytest.mark.gpu
def test_star_formation_is_refused():
    """P['trigger'] > 0 asks for star formation (sim.py:1061-1097), which is not built: refused, not ignored."""
    with pytest.raises(NotImplementedError, match="trigger"):
        Sim("torus", res=16, params=dict(trigger=4.0), fast=False)
    s = Sim("torus", res=16, rays=(16, 32), fast=False, verbose=False)
    with pytest.raises(NotImplementedError, match="trigger"):
        s.set_param("trigger", 4.0)


def test_batch_residency_is_checked_before_any_device_work():
    from nebulaviewer_amd.batch import SimBatch
    with pytest.raises(ValueError, match="queue_wgs"):
        SimBatch([object()], queue_wgs=-1)
    with pytest.raises(ValueError, match="dispatch"):
        SimBatch([object()], dispatch="other")
